/* Drop-in header name of the reference build (<grid/au.h>): the unit conversions are declared in
   libdft_mi355x.h; <grid/grid.h> adds libgrid's math names and HBAR. */
#include "grid.h"
