/*
 * libdft_mi355x -- MI355X-native Orsay-Trento (OT) hot path of libdft.
 *
 * C ABI: plain structs, pointers and sizes.  REAL = double, INT = int64_t
 * at the API; the device storage precision (FP32 or FP64) is a property of
 * each grid handle, fixed when the grid is allocated.
 *
 * The functional entry points keep the reference names and signatures:
 *   dft_ot_alloc          src/ot.c:66       dft_ot_free              src/ot.c:268
 *   dft_ot_potential      src/ot.c:314      dft_ot_potential2        src/ot.c:431
 *   dft_ot_energy_density src/ot-energy.c:38
 *   dft_ot_energy_density_kc src/ot-energy.c:123, _bf src/ot-energy.c:202
 *   dft_ot_backflow_potential src/ot.c:840  dft_ot_init_params       src/ot.c:1002
 *   dft_ot_add_lennard_jones_potential src/ot.c:524
 *   dft_ot_add_local_correlation_potential src/ot.c:547
 *   dft_ot_bulk_*         src/helium-ot-bulk.c:21-463
 * and the libgrid subset they and their callers use (libgrid is not part of
 * the reference tree; call sites are cited per function).
 *
 * Semantics (same as the reference): dft_ot_potential ADDS V_OT to
 * `potential` and sets otf->density; dft_ot_energy_density overwrites the
 * output and sets otf->density.  Calls on one functional are ordered on the
 * library's HIP stream; functions returning host scalars synchronise.
 * Functions are not reentrant for one otf (shared workspaces).
 */
#ifndef LIBDFT_MI355X_H
#define LIBDFT_MI355X_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REAL double
#define INT int64_t
typedef double _Complex REAL_COMPLEX;
#define FMT_R "%le"
#define FMT_I "%ld"

/* Device storage precision of a grid. */
#define GRID_FP32 1
#define GRID_FP64 2

typedef struct rgrid_struct rgrid;
typedef struct cgrid_struct cgrid;

/* Real grid handle (libgrid rgrid; fields callers read are kept). */
struct rgrid_struct {
  INT nx, ny, nz;      /* row-major [i][j][k], k fastest */
  REAL step;           /* bohr */
  REAL x0, y0, z0;     /* origin; x_i = (i - nx/2) step - x0 */
  void *dev;           /* device storage (float or double), owned */
  INT dev_elems;       /* elements allocated in dev */
  int prec;            /* GRID_FP32 or GRID_FP64 */
  int flag;            /* claim flag (rgrid_claim / rgrid_release) */
  int spectral;        /* 1: holds a real half spectrum [nx][ny][nz/2+1] */
  char id[32];
  /* boundary function given to rgrid_alloc (value at an index outside the
     grid; only periodic boundaries exist here) and its parameter pointer */
  REAL (*value_outside)(struct rgrid_struct *, INT, INT, INT);
  void *outside_params_ptr;
};

/* Complex grid handle (libgrid cgrid). */
struct cgrid_struct {
  INT nx, ny, nz;
  REAL step;
  REAL x0, y0, z0;
  void *dev;           /* device storage (float2 or double2), owned */
  INT dev_elems;
  int prec;
  int flag;
  REAL kx0, ky0, kz0;  /* momentum frame (cgrid_set_momentum) */
  char id[32];
  REAL_COMPLEX (*value_outside)(struct cgrid_struct *, INT, INT, INT);
  void *outside_params_ptr;
};

#define WF_PERIODIC_BOUNDARY 1
#define WF_2ND_ORDER_FFT 2

/* Wave function (libgrid wf). */
typedef struct wf_struct {
  cgrid *grid;
  REAL mass;
  REAL norm;           /* target norm for grid_wf_normalize */
  char boundary;
  char propagator;
  cgrid *cworkspace;   /* propagator scratch (unused by the FFT propagator) */
} wf;

/* Boundary "functions" accepted by rgrid_alloc / cgrid_alloc (only periodic). */
REAL rgrid_value_outside_periodic(rgrid *grid, INT i, INT j, INT k);
REAL_COMPLEX cgrid_value_outside_periodic(cgrid *grid, INT i, INT j, INT k);
#define RGRID_PERIODIC_BOUNDARY rgrid_value_outside_periodic
#define CGRID_PERIODIC_BOUNDARY cgrid_value_outside_periodic

/* ---------------------------------------------------------------- setup -- */
/* Precision used by the allocations that follow (default GRID_FP64, the
   reference's REAL = double build). */
void grid_set_precision(int prec);
int grid_get_precision(void);
/* No-ops kept for drop-in source compatibility with reference drivers. */
void grid_threads_init(INT threads);
void grid_set_fftw_flags(char flags);
void grid_fft_read_wisdom(char *file);
void grid_fft_write_wisdom(char *file);
void cuda_enable(char flag, int ngpus, int *gpus);
/* Synchronise the library stream. */
void grid_synchronize(void);
/* The HIP stream (hipStream_t) every kernel of the library is launched on. */
void *grid_get_stream(void);

/* ------------------------------------------------------------ allocation -- */
rgrid *rgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL (*value_outside)(rgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id);
cgrid *cgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL_COMPLEX (*value_outside)(cgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id);
rgrid *rgrid_clone(rgrid *grid, char *id);
cgrid *cgrid_clone(cgrid *grid, char *id);
void rgrid_free(rgrid *grid);
void cgrid_free(cgrid *grid);
void rgrid_set_origin(rgrid *grid, REAL x0, REAL y0, REAL z0);
void cgrid_set_origin(cgrid *grid, REAL x0, REAL y0, REAL z0);
void rgrid_claim(rgrid *grid);
void rgrid_release(rgrid *grid);
wf *grid_wf_alloc(INT nx, INT ny, INT nz, REAL step, REAL mass, char boundary, char propagator, char *id);
wf *grid_wf_clone(wf *gwf, char *id);
void grid_wf_free(wf *gwf);

/* ------------------------------------------- host <-> device data transfer -- */
/* Host buffers are FP64 (REAL / REAL_COMPLEX), nx*ny*nz entries, row-major. */
void rgrid_copy_from_host(rgrid *grid, const REAL *src);
void rgrid_copy_to_host(const rgrid *grid, REAL *dst);
void cgrid_copy_from_host(cgrid *grid, const REAL_COMPLEX *src);
void cgrid_copy_to_host(const cgrid *grid, REAL_COMPLEX *dst);

/* ------------------------------------------------------------ grid files -- */
/* Checkpoint files (replace libgrid's rgrid_/cgrid_write_grid, _read_grid,
   _write, _read; call sites e.g. examples/gas-bubble/bubble.c:141-143,
   examples/vortex-trace/trace.c:155,177).  The file layout is this library's
   own: "LDFTGRD1", kind, nx, ny, nz (int64), step, x0, y0, z0 (double), then
   FP64 values row-major (complex as re, im).  *_write_grid writes <base>.grd;
   *_read_grid reads the named file into a grid of the same shape; *_read
   allocates the grid when given NULL.  Errors print a libdft: message (the
   readers then return NULL / leave the grid unchanged). */
void rgrid_write(rgrid *grid, FILE *fp);
rgrid *rgrid_read(rgrid *grid, FILE *fp);
void cgrid_write(cgrid *grid, FILE *fp);
cgrid *cgrid_read(cgrid *grid, FILE *fp);
void rgrid_write_grid(char *base, rgrid *grid);
void cgrid_write_grid(char *base, cgrid *grid);
void rgrid_read_grid(rgrid *grid, char *file);
void cgrid_read_grid(cgrid *grid, char *file);
REAL rgrid_value_at_index(rgrid *grid, INT i, INT j, INT k);
REAL_COMPLEX cgrid_value_at_index(cgrid *grid, INT i, INT j, INT k);
/* Host callback maps (evaluated on the host, uploaded once). */
void rgrid_map(rgrid *grid, REAL (*func)(void *arg, REAL x, REAL y, REAL z), void *farg);
void cgrid_map(cgrid *grid, REAL_COMPLEX (*func)(void *arg, REAL x, REAL y, REAL z), void *farg);
void grid_wf_map(wf *gwf, REAL_COMPLEX (*func)(void *arg, REAL x, REAL y, REAL z), void *farg);

/* -------------------------------------------------- grid operations (GPU) -- */
void rgrid_zero(rgrid *grid);
void cgrid_zero(cgrid *grid);
void rgrid_constant(rgrid *grid, REAL c);
void cgrid_constant(cgrid *grid, REAL_COMPLEX c);
void grid_wf_constant(wf *gwf, REAL_COMPLEX c);
void rgrid_copy(rgrid *dst, rgrid *src);
void cgrid_copy(cgrid *dst, cgrid *src);
void cgrid_add(cgrid *grid, REAL_COMPLEX c);
void cgrid_multiply(cgrid *grid, REAL_COMPLEX c);
void rgrid_add(rgrid *grid, REAL c);
void rgrid_multiply(rgrid *grid, REAL c);
/* dst = dst + s * a * b */
void rgrid_add_scaled_product(rgrid *dst, REAL s, rgrid *a, rgrid *b);
void grid_real_to_complex_re(cgrid *dst, rgrid *src);
void grid_add_real_to_complex_re(cgrid *dst, rgrid *src);
void grid_add_real_to_complex_im(cgrid *dst, rgrid *src);
void grid_wf_density(wf *gwf, rgrid *density);
/* v_d = Im(psi* d_d psi) / (M (|psi|^2 + eps)), periodic central differences. */
void grid_wf_velocity(wf *gwf, rgrid *vx, rgrid *vy, rgrid *vz, REAL eps);
void rgrid_gradient_x(rgrid *src, rgrid *dst);
void rgrid_gradient_y(rgrid *src, rgrid *dst);
void rgrid_gradient_z(rgrid *src, rgrid *dst);

/* Further libgrid calls of the reference drivers (examples/vortex-ring/vring.c:63-71,149,
   examples/gas-bubble/bubble.c:104).  cgrid_set_momentum sets the frame k0 of the
   propagator and the kinetic energy, whose symbol becomes (k - k0)^2 / 2M. */
void cgrid_set_momentum(cgrid *grid, REAL kx0, REAL ky0, REAL kz0);
/* dst = Laplacian of src (2nd-order central difference, periodic) */
void cgrid_fd_laplace(cgrid *src, cgrid *dst);
/* dst = conj(a) * b */
void cgrid_conjugate_product(cgrid *dst, cgrid *a, cgrid *b);
/* dst = |src|^p */
void cgrid_abs_power(cgrid *dst, cgrid *src, REAL p);
void rgrid_abs_power(rgrid *dst, rgrid *src, REAL p);
/* rot = |curl (fx, fy, fz)| (central differences, periodic) */
void rgrid_abs_rot(rgrid *rot, rgrid *fx, rgrid *fy, rgrid *fz);
/* j_d = (hbar / M) Im(psi* d_d psi), central differences; any output may be NULL */
void grid_wf_probability_flux(wf *gwf, rgrid *jx, rgrid *jy, rgrid *jz);
/* one component of it (examples/added_mass/added_mass.c:192-196) */
void grid_wf_probability_flux_x(wf *gwf, rgrid *flux_x);
void grid_wf_probability_flux_y(wf *gwf, rgrid *flux_y);
void grid_wf_probability_flux_z(wf *gwf, rgrid *flux_z);
/* dst = src^p, dst = a b, dst = a / (b + eps), the largest and smallest value
   (examples/soliton/soliton.c:96, examples/vortex-trace/trace.c:183-185) */
void rgrid_power(rgrid *dst, rgrid *src, REAL p);
void rgrid_product(rgrid *dst, rgrid *a, rgrid *b);
void rgrid_division_eps(rgrid *dst, rgrid *a, rgrid *b, REAL eps);
REAL rgrid_max(rgrid *grid);
REAL rgrid_min(rgrid *grid);
/* h^3 sum f(x, y, z) grid (f evaluated on the host) */
REAL rgrid_weighted_integral(rgrid *grid, REAL (*func)(void *arg, REAL x, REAL y, REAL z), void *farg);

/* Reductions: FP64 accumulation with a compensated final sum. */
REAL rgrid_integral(rgrid *grid);
REAL rgrid_integral_of_product(rgrid *a, rgrid *b);
/* h^3 sum over the points with x in [xl, xu], y in [yl, yu], z in [zl, zu]
   (x_i = (i - nx/2) step - x0; bounds inclusive) */
REAL rgrid_integral_region(rgrid *grid, REAL xl, REAL xu, REAL yl, REAL yu, REAL zl, REAL zu);
REAL_COMPLEX cgrid_integral_region(cgrid *grid, REAL xl, REAL xu, REAL yl, REAL yu, REAL zl, REAL zu);
REAL_COMPLEX cgrid_integral(cgrid *grid);
/* h^3 sum conj(a) b */
REAL_COMPLEX cgrid_integral_of_conjugate_product(cgrid *a, cgrid *b);
/* h^3 sum |psi|^2 Re V */
REAL grid_wf_potential_energy(wf *gwf, cgrid *potential);
REAL grid_wf_norm(wf *gwf);
void grid_wf_normalize(wf *gwf);
/* grid_wf_normalize without allocation or host synchronisation (no counterpart
   in the reference): for callers that cannot allocate, such as code under HIP
   graph capture or tight stepping loops.  psi *= sqrt(gwf->norm / N), with
   N = h^3 sum |psi|^2 summed in FP64 on the device, all on the library
   stream.  Its partial sums live in a fixed device array of 1026 doubles per
   GPU, held until the process ends and shared by every call on that GPU:
   calls that overlap on one GPU race.  Returns 0, or -1 (message on stderr)
   when there is no wave function or no usable device. */
int grid_wf_normalize_noalloc(wf *gwf);
/* Kinetic energy h^3/N sum_k ((k - k0)^2/2M)|psi_k|^2, plus h^3 sum |psi|^2 Re V
   when potential is not NULL. */
REAL grid_wf_energy(wf *gwf, cgrid *potential);

/* ------------------------------------------------ split-operator stepping -- */
/* psi <- e^{-iV dt/2} F^-1 e^{-i k^2 dt/2M} F e^{-iV dt/2} psi, complex V and dt. */
void grid_wf_propagate(wf *gwf, cgrid *potential, REAL_COMPLEX time);
/* gwfp = Step(gwf; potential) */
void grid_wf_propagate_predict(wf *gwf, wf *gwfp, cgrid *potential, REAL_COMPLEX time);
/* gwf = Step(gwf; potential), potential = (V_n + V_p)/2 formed by the caller */
void grid_wf_propagate_correct(wf *gwf, cgrid *potential, REAL_COMPLEX time);

/* ------------------------------------------------------------ timers -- */
typedef struct { double t0; } grid_timer;
void grid_timer_start(grid_timer *timer);
REAL grid_timer_wall_clock_time(grid_timer *timer);

/* ============================================================ OT functional */

/* Model flags, src/ot.h:37-60 (PLAIN, KC, BACKFLOW, the HD/HD2 high-density
   penalty with its backflow density switch, and the thermal models T0MK ...
   T3000MK are implemented). */
#define DFT_OT_PLAIN    0
#define DFT_OT_KC       1
#define DFT_OT_HD       2
#define DFT_OT_HD2      4
#define DFT_OT_BACKFLOW 8
#define DFT_OT_T0MK    16
#define DFT_OT_T400MK  32
#define DFT_OT_T600MK  64
#define DFT_OT_T800MK  128
#define DFT_OT_T1200MK 256
#define DFT_OT_T1400MK 512
#define DFT_OT_T1600MK 1024
#define DFT_OT_T1800MK 2048
#define DFT_OT_T2000MK 4096
#define DFT_OT_T2100MK 8192
#define DFT_OT_T2200MK 16384
#define DFT_OT_T2400MK 32768
#define DFT_OT_T2600MK 65536
#define DFT_OT_T2800MK 131072
#define DFT_OT_T3000MK 262144
#define DFT_GP         524288
#define DFT_DR         1048576
#define DFT_ZERO       2097152
#define DFT_GP2        4194304

#define DFT_MIN_SUBSTEPS 4   /* src/ot.h:142 */
#define DFT_MAX_SUBSTEPS 32  /* src/ot.h:143 */
#define DFT_EPS 1E-4         /* src/ot.h:149 */

typedef struct {
  REAL h, sigma, epsilon, cval;   /* src/dft.h:24-29 */
} dft_common_lj;

typedef struct {
  REAL g11, g12, g21, g22, a1, a2; /* src/ot.h:67-69 */
} dft_ot_bf;

struct dft_ot_device;

/* src/ot.h:90-133.  Kernel fields hold real half spectra on the device
   (the kernels are even, so their spectra are real). */
typedef struct dft_ot_functional_struct {
  INT model;
  REAL b, c2, c2_exp, c3, c3_exp, c4, rho_0s, alpha_s, l_g, mass, rho0, temp;
  dft_common_lj lj_params;
  dft_ot_bf bf_params;
  rgrid *lennard_jones, *spherical_avg, *gaussian_tf, *gaussian_x_tf, *gaussian_y_tf, *gaussian_z_tf, *backflow_pot;
  REAL beta, rhom, C, mu0, xi, rhobf;
  REAL div_epsilon, max_veloc, max_bfpot, c_bfpot;
  rgrid *workspace1, *workspace2, *workspace3, *workspace4, *workspace5, *workspace6, *workspace7, *workspace8,
      *workspace9;
  rgrid *density;
  /* ---- extensions (not in the reference struct) ---- */
  REAL kc_coef23;                 /* KC terms 2,3 coefficient: 1 (reference) or 1/rho_0s (exact) */
  struct dft_ot_device *device;   /* device state: plans, workspaces, scratch */
} dft_ot_functional;

#define DFT_HELIUM_MASS (4.002602 / GRID_AUTOAMU) /* src/ot.h:139 */

dft_ot_functional *dft_ot_alloc(INT model, wf *gwf, INT min_substeps, INT max_substeps);
void dft_ot_free(dft_ot_functional *otf);
void dft_ot_init_params(dft_ot_functional *otf, INT model);
void dft_ot_potential(dft_ot_functional *otf, cgrid *potential, wf *gwf);
void dft_ot_potential2(dft_ot_functional *otf, cgrid *potential, rgrid *density, rgrid *veloc_x, rgrid *veloc_y,
                       rgrid *veloc_z);
void dft_ot_backflow_potential(dft_ot_functional *otf, cgrid *potential, rgrid *density, rgrid *veloc_x,
                               rgrid *veloc_y, rgrid *veloc_z, rgrid *workspace1, rgrid *workspace2,
                               rgrid *workspace3, rgrid *workspace4, rgrid *workspace5, rgrid *workspace6);
/* src/ot.c:524: potential += V_LJ * density.  Unlike the reference, workspace1 does
 * not receive FFT(density): the workspaces are accepted and not written. */
void dft_ot_add_lennard_jones_potential(dft_ot_functional *otf, cgrid *potential, rgrid *density,
                                        rgrid *workspace1, rgrid *workspace2);
/* src/ot.c:547: potential += the local correlation terms of rhobar = Pi * rho.  rho_tf
 * (FFT(rho) in the reference) and the workspaces are accepted and unused. */
void dft_ot_add_local_correlation_potential(dft_ot_functional *otf, cgrid *potential, rgrid *rho, rgrid *rho_tf,
                                            rgrid *workspace1, rgrid *workspace2);
void dft_ot_energy_density(dft_ot_functional *otf, rgrid *energy_density, wf *gwf);
void dft_ot_energy_density_kc(dft_ot_functional *otf, rgrid *energy_density, wf *gwf, rgrid *density);
void dft_ot_energy_density_bf(dft_ot_functional *otf, rgrid *energy_density, wf *gwf, rgrid *density);

/* Ideal gas functions of the thermal models (host FP64), src/common.c:253
   (thermal wavelength cubed), 320-374 (g_{1/2}, g_{3/2}, g_{5/2}: 256-term
   series), 403-437 (z with g_{3/2}(z) = val, golden section to 1e-6),
   489-530 (Bose gas energy per volume and its density derivative; params is
   the dft_ot_functional), 452-480 (classical ideal gas). */
REAL dft_common_lwl3(REAL mass, REAL temp);
REAL dft_common_fit_g12(REAL z);
REAL dft_common_fit_g32(REAL z);
REAL dft_common_fit_g52(REAL z);
REAL dft_common_fit_z(REAL val);
REAL dft_common_bose_idealgas_energy(REAL rhop, void *params);
REAL dft_common_bose_idealgas_dEdRho(REAL rhop, void *params);
REAL dft_common_classical_idealgas_energy(REAL rhop, void *params);
REAL dft_common_classical_idealgas_dEdRho(REAL rhop, void *params);

/* Bulk closed forms, src/helium-ot-bulk.c (host FP64). */
REAL dft_ot_bulk_energy(dft_ot_functional *otf, REAL rho);
REAL dft_ot_bulk_dEdRho(dft_ot_functional *otf, REAL rho);
REAL dft_ot_bulk_density(dft_ot_functional *otf);
REAL dft_ot_bulk_chempot(dft_ot_functional *otf);
REAL dft_ot_bulk_chempot2(dft_ot_functional *otf);
REAL dft_ot_bulk_chempot3(dft_ot_functional *otf, REAL rho0);
REAL dft_ot_bulk_pressure(dft_ot_functional *otf, REAL rho);
REAL dft_ot_bulk_dPdRho(dft_ot_functional *otf, REAL rho);
REAL dft_ot_bulk_density_pressurized(dft_ot_functional *otf, REAL pressure);
REAL dft_ot_bulk_chempot_pressurized(dft_ot_functional *otf, REAL pressure);
REAL dft_ot_bulk_compressibility(dft_ot_functional *otf, REAL rho);
REAL dft_ot_bulk_sound_speed(dft_ot_functional *otf, REAL rho);
REAL dft_ot_bulk_dispersion(dft_ot_functional *otf, REAL *k, REAL rho0);
/* Dispersion relation from the dynamics of a density wave (src/helium-ot-bulk.c:274):
   energy (hartree) of the wave k (rounded to the grid, returned in *k) along
   axis dir; predict-correct steps when pred.  0 for k = 0; -1 on invalid
   input or when no half period is found. */
REAL dft_ot_dispersion(wf *gwf, dft_ot_functional *otf, REAL ts, REAL *k, REAL amp, char pred, char dir);
REAL dft_ot_bulk_istatic(dft_ot_functional *otf, REAL *k, REAL rho0);

/* Kernel functions, src/common.c:31-66, 131-138, 173-179, 193-200; src/ot-private.h:11-22. */
REAL dft_common_lj_func(REAL r2, REAL sig, REAL eps);
REAL dft_common_lennard_jones(void *arg, REAL x, REAL y, REAL z);
REAL dft_common_spherical_avg(void *arg, REAL x, REAL y, REAL z);
REAL dft_common_spherical_avg_k(void *arg, REAL kx, REAL ky, REAL kz);
REAL dft_common_gaussian(void *arg, REAL x, REAL y, REAL z);
REAL dft_ot_backflow_pot(void *arg, REAL x, REAL y, REAL z);

/* Map functions for initial states and test potentials (host, for rgrid_map /
   grid_wf_map): the complex gaussian (src/common.c:235-242), a plane-wave
   density modulation (1214-1224, parameters as src/dft.h:55-60), the
   Feynman-Onsager vortex-line potentials 1/(2 M r^2) with the core cut at
   r = 0.05 bohr (1240-1290, param: the mass), vortex lines of one and two
   quanta along x, y, z (src/initial.c:31-150) and a sharp bubble of radius
   *prm (154-161). */
typedef struct {
  REAL kx, ky, kz; /* wave vector */
  REAL a;          /* relative amplitude */
  REAL rho;        /* background density */
} dft_plane_wave;
REAL_COMPLEX dft_common_cgaussian(void *arg, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_common_planewave(void *arg, REAL x, REAL y, REAL z);
REAL dft_common_vortex_x(void *param, REAL x, REAL y, REAL z);
REAL dft_common_vortex_y(void *param, REAL x, REAL y, REAL z);
REAL dft_common_vortex_z(void *param, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_initial_vortex_x_n1(void *na, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_initial_vortex_y_n1(void *na, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_initial_vortex_z_n1(void *na, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_initial_vortex_x_n2(void *na, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_initial_vortex_y_n2(void *na, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_initial_vortex_z_n2(void *na, REAL x, REAL y, REAL z);
REAL_COMPLEX dft_initial_bubble(void *prm, REAL x, REAL y, REAL z);

/* ---- extensions ---- */
/* One split-operator step with V = V_OT[psi] + V_ext - mu (vext may be NULL);
   equal to the reference single-step driver sequence cgrid_zero /
   grid_real_to_complex_re / dft_ot_potential / cgrid_add(-mu) /
   grid_wf_propagate (doc/libdft-manual.tex:496-504), fused on the device. */
void dft_ot_propagate(dft_ot_functional *otf, wf *gwf, rgrid *vext, REAL mu, REAL_COMPLEX time);
/* Predict-correct step of examples/imag_time/imag_time.c:231-239 (gwfp gets the predictor). */
void dft_ot_propagate_pc(dft_ot_functional *otf, wf *gwf, wf *gwfp, rgrid *vext, REAL mu, REAL_COMPLEX time);
/* dft_ot_propagate / dft_ot_propagate_pc followed by grid_wf_normalize (psi *= sqrt(gwf->norm / N)),
   as examples/imag_time_rot/imag_time.c:198-208 does; the norm is summed in the step's closing
   z pass and applied on the device, with no host round trip.
   0 on success, -1 (with a libdft: message) for invalid arguments. */
int dft_ot_propagate_normalize(dft_ot_functional *otf, wf *gwf, rgrid *vext, REAL mu, REAL_COMPLEX time);
int dft_ot_propagate_pc_normalize(dft_ot_functional *otf, wf *gwf, wf *gwfp, rgrid *vext, REAL mu,
                                  REAL_COMPLEX time);
/* KC potential mode: 0 = reference formula (default), 1 = exact dE/drho. */
void dft_ot_set_kc_mode(dft_ot_functional *otf, int exact);
/* Energy of the functional, h^3 sum eps_OT, in FP64 (kinetic part excluded). */
REAL dft_ot_energy(dft_ot_functional *otf, wf *gwf);
/* Hand-written backward of E[rho] = h^3 sum eps (LJ, c2/c3, KC parts of
   src/ot-energy.c:38-191): grad = dE/drho, the exact derivative of the
   discrete energy.  For PLAIN it equals dft_ot_potential's output; for KC it
   differs from it (src/ot.c:660-675 omits 1/rho_0s on terms 2 and 3).  With a
   thermal model the ideal Bose gas adds c4 k_B T ln z below saturation. */
void dft_ot_energy_gradient(dft_ot_functional *otf, rgrid *grad, rgrid *density);
/* E[rho] of the same parts, FP64 reduction. */
REAL dft_ot_energy_rho(dft_ot_functional *otf, rgrid *density);

/* ---- adjoint (extensions; no counterpart in the reference) ---- */
/* Reverse pass of dft_ot_potential (PLAIN and KC models, with or without the
   HD/HD2 penalty) at the state gwf:
   rho_bar = (dV_OT/drho)^T v_bar, the transpose of the potential's Jacobian
   applied to v_bar.  Returns 0, or -1 (message on stderr) for a backflow
   model, mismatched grids or lack of device memory. */
int dft_ot_potential_vjp(dft_ot_functional *otf, rgrid *rho_bar, rgrid *v_bar, wf *gwf);
/* Reverse pass of dft_ot_potential with respect to the wave function, for every
   model (backflow's V depends on psi's phase through the velocity,
   src/ot.c:385-396, 840-991): psi_bar = (dV_OT/dpsi)^T v_bar, where Re v_bar is
   the adjoint of Re V and Im v_bar that of Im V, in the convention
   dJ = Re sum conj(psi_bar) dpsi.  psi_bar is overwritten.  Returns 0, or -1
   (message on stderr) for mismatched grids or lack of device memory. */
int dft_ot_potential_vjp_psi(dft_ot_functional *otf, cgrid *psi_bar, cgrid *v_bar, wf *gwf);
/* Gradient of the total energy E[psi_K] = E_kin + E_OT + h^3 sum rho V_ext
   after nsteps steps of dft_ot_propagate(otf, ., V_ext, mu, time) from gwf
   (unchanged), with V_ext(r) = amp exp(-|r - r0|^2 / (2 width^2)) on the grid
   r_i = (i - n/2) step - r0, with respect to (amp, width): grad[0..1].  An
   adjoint sweep that keeps every `every`-th state and recomputes each segment
   in the reverse pass (memory: nsteps/every + every states).  Every model
   (PLAIN, KC, BACKFLOW, HD/HD2 and its backflow switch, the thermal models);
   returns 0, or -1 (message on stderr) for invalid arguments or lack of device
   memory. */
int dft_ot_adjoint_well(dft_ot_functional *otf, wf *gwf, REAL amp, REAL width, REAL x0, REAL y0, REAL z0, REAL mu,
                        REAL_COMPLEX time, INT nsteps, INT every, REAL *energy, REAL *grad);
/* Released cavity: writes psi_0 = sqrt(rho0 (1 + tanh((r - radius)/width)) / 2)
   (r = |r_i|, r_i = (i - n/2) step) into gwf, propagates it nsteps steps of
   dft_ot_propagate(otf, ., no V_ext, mu, time) and returns the bubble radius
   R_b = (3 V_b / 4 pi)^(1/3), V_b = h^3 sum W(r) (1 - rho / rho0) with the window
   W = (1 - tanh((r - 1.5 radius) / (4 width))) / 2 (the missing volume of the
   whole box is the conserved N / rho0), of the final state and its derivative
   with respect to the initial cavity radius, from the same
   checkpointed adjoint sweep as dft_ot_adjoint_well.  Every model; returns 0,
   or -1 (message on stderr) for invalid arguments or lack of device memory. */
int dft_ot_adjoint_cavity(dft_ot_functional *otf, wf *gwf, REAL rho0, REAL radius, REAL width, REAL mu,
                          REAL_COMPLEX time, INT nsteps, INT every, REAL *bubble_radius, REAL *grad);
/* Progress lines of the adjoint sweeps on stderr (level > 0), one per checkpoint. */
void dft_set_verbose(int level);

/* ---- measurement (extensions; no counterpart in the reference) ---- */
/* Per-launch HIP-event timing of the library's kernels on its stream. */
void dft_profile_enable(int on);
void dft_profile_reset(void);
/* Aggregates recorded launches per kernel name; returns the number of names. */
int dft_profile_collect(void);
/* Entry i: name, launches, total milliseconds, algorithmic bytes (inputs + outputs once). */
int dft_profile_entry(int i, char *name, int len, INT *launches, REAL *ms, REAL *bytes);
/* Plain device copy between two grids (traffic-counter calibration). */
void grid_bench_copy(rgrid *dst, rgrid *src);

/* Atomic-unit conversions (CODATA 2018; libgrid's au.h). */
#define GRID_AUTOANG 0.529177210903
#define GRID_AUTOK 315775.02480407
#define GRID_AUTOFS 0.024188843265857
#define GRID_AUTOS 2.4188843265857E-17
#define GRID_HZTOCM1 3.3356409519815E-11
#define GRID_AUTOAMU (1.0 / 1822.888486209)
#define GRID_AUTOMPS 2187691.26364
#define GRID_AUTOBAR 2.9421015697E8
#define GRID_AUTOPA 2.9421015697E13
#define GRID_AUTOCM1 219474.6313632
#define GRID_AUTOEV 27.211386245988
#define GRID_AUTOM 5.29177210903E-11
#define GRID_AUTOJ 4.3597447222071E-18
#define GRID_AUTOKG 9.1093837015E-31

#ifdef __cplusplus
}
#endif

#endif /* LIBDFT_MI355X_H */
