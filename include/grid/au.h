/* Drop-in header name of the reference build (<grid/au.h>): everything is declared in libdft_mi355x.h. */
#include "../libdft_mi355x.h"
