/* Drop-in header name of the reference build (<dft/dft.h>): everything is declared in libdft_mi355x.h;
   <grid/grid.h> adds the libgrid names drivers use next to it. */
#include "../grid/grid.h"
