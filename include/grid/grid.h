/*
 * Drop-in header name of the reference build (<grid/grid.h>, libgrid).
 *
 * Everything the library exports is declared in libdft_mi355x.h.  This header
 * adds what reference drivers take from libgrid's own headers beyond the
 * functions: <complex.h> and <math.h> (drivers write I, cexp, sqrt),
 * libgrid's REAL-precision math names (SQRT, EXP, CREAL, ...: REAL = double
 * here, so each is the double routine), HBAR (atomic units) and the unit
 * conversions of <grid/au.h>.  Call sites, e.g.
 * examples/imag_time/imag_time.c:162,171,234, examples/dynamics/dynamics.c:49,124,
 * examples/gas-bubble/bubble.c:20, examples/dispersion/dispersion2.c:29.
 */
#ifndef LIBDFT_MI355X_GRID_COMPAT_H
#define LIBDFT_MI355X_GRID_COMPAT_H

#include <math.h>
#ifndef __cplusplus
#include <complex.h>
#endif

#include "../libdft_mi355x.h"

/* hbar in atomic units */
#ifndef HBAR
#define HBAR 1.0
#endif

/* REAL (double) math */
#define SQRT sqrt
#define EXP exp
#define LOG log
#define POW pow
#define FABS fabs
#define SIN sin
#define COS cos
#define TAN tan
#define ASIN asin
#define ACOS acos
#define ATAN atan
#define ATAN2 atan2
#define SINH sinh
#define COSH cosh
#define TANH tanh
#define FLOOR floor
#define CEIL ceil
#define FMOD fmod
#define ERF erf
#define FMIN fmin
#define ABS fabs
#define FMAX fmax
#define HYPOT hypot
#define CBRT cbrt

/* REAL complex math (C only) */
#ifndef __cplusplus
#define CEXP cexp
#define CSQRT csqrt
#define CLOG clog
#define CPOW cpow
#define CABS cabs
#define CARG carg
#define CONJ conj
#define CREAL creal
#define CIMAG cimag
#define CSIN csin
#define CCOS ccos
#endif

#endif /* LIBDFT_MI355X_GRID_COMPAT_H */
