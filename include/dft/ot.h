/* Drop-in header name of the reference build (<dft/ot.h>): everything is declared in libdft_mi355x.h. */
#include "../libdft_mi355x.h"
