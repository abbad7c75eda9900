/*
 * FP64 CPU oracle for the Orsay-Trento (OT) hot path of libdft.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker (and the timed CPU
 * baseline).  Nothing in the product library links, loads or calls it.
 *
 * It restates, in plain C99 + OpenMP on host memory, the reference functions
 *   src/ot.c:66-258        dft_ot_alloc (kernel spectra)
 *   src/ot.c:314-408       dft_ot_potential
 *   src/ot.c:524-797       LJ, local correlation and kinetic correlation (KC)
 *   src/ot.c:840-991       dft_ot_backflow_potential (BF)
 *   src/ot.c:1002-1271     dft_ot_init_params (T = 0 rows and common rows)
 *   src/ot-energy.c:38-288 dft_ot_energy_density (+ _kc, _bf)
 *   src/common.c:31-66,131-138,193-200, src/ot-private.h:11-22  kernel functions
 *   src/helium-ot-bulk.c:21-258, 379-463  bulk closed forms
 *   src/ot.c:164-170, 374-378, 1004-1007, src/ot-energy.c:96-99  HD/HD2 penalty
 * and the libgrid primitives the path uses (libgrid is not vendored; its
 * semantics are fixed as stated next to each function in ot_oracle.c).
 * The HD penalty's formula is libgrid's (grid_func4/5) and is restated from
 * PRB 72, 214522 (2005): parity with libgrid itself is unpinned for it; the
 * tests pin it to its own energy by finite differences.
 *
 * Layout: row-major [i][j][k], k fastest; x_i = (i - nx/2) * step.
 * All quantities in atomic units (hbar = 1).
 */
#ifndef OT_ORACLE_H
#define OT_ORACLE_H

#include <stdint.h>
#include <complex.h>

/* Model flags: same values as src/ot.h:37-60 (PLAIN/KC/BACKFLOW, and the
   high-density penalty HD/HD2 without backflow). */
#define ORC_OT_PLAIN    0
#define ORC_OT_KC       1
#define ORC_OT_HD       2
#define ORC_OT_HD2      4
#define ORC_OT_BACKFLOW 8
#define ORC_OT_T0MK     16
#define ORC_OT_T400MK   32      /* the thermal rows: one bit each, 32 ... 262144 */
#define ORC_GP          524288

/* Unit conversions: CODATA 2018 (libgrid's au.h, which the reference uses, is not available). */
#define ORC_AUTOANG 0.529177210903
#define ORC_AUTOK   315775.02480407
#define ORC_AUTOFS  0.024188843265857
#define ORC_AUTOAMU (1.0 / 1822.888486209)
#define ORC_AUTOMPS 2187691.26364
#define ORC_AUTOBAR 2.9421015697e8
#define ORC_HELIUM_MASS (4.002602 / ORC_AUTOAMU)   /* src/ot.h:139 DFT_HELIUM_MASS */
#define ORC_DFT_EPS 1e-4                           /* src/ot.h:149 DFT_EPS */

typedef struct {
  int64_t model;
  int nx, ny, nz;
  double step;
  /* parameters, src/ot.c:1026-1036, 1234-1270 (atomic units) */
  double b, c2, c2_exp, c3, c3_exp, c4, rho_0s, alpha_s, l_g, mass, rho0, temp;
  double lj_h, lj_sigma, lj_epsilon, lj_cval;
  double g11, g12, g21, g22, a1, a2;
  double div_epsilon, max_veloc, max_bfpot, c_bfpot;
  /* coefficient on KC terms 2 and 3: 1 = reference formula (src/ot.c:660-675),
     1/rho_0s = exact functional derivative of src/ot-energy.c:123-191 */
  double kc_coef23;
  /* full complex spectra (N entries), centred sampling: r = 0 at index n/2 */
  double complex *lj_tf, *sa_tf, *g_tf, *gx_tf, *gy_tf, *gz_tf, *bf_tf;
  /* high-density penalty (src/ot.c:1004-1011; zero without HD/HD2) and the
     spherical-average radius (h, or 1.065 h for HD2, src/ot.c:164-170) */
  double hd_beta, hd_rhom, hd_C, sa_h;
  /* backflow density switch of HD / HD2 with backflow (src/ot.c:1235-1236) */
  double bf_xi, bf_rhobf;
} orc_ot;

/* Parameters (a1) and kernel spectra (a2). min_sub/max_sub as in dft_ot_alloc. */
orc_ot *orc_ot_alloc(int64_t model, int nx, int ny, int nz, double step, int min_sub, int max_sub);
void orc_ot_free(orc_ot *otf);
void orc_ot_set_kc_mode(orc_ot *otf, int exact);

/* Adds V_OT[psi] to pot (complex, N entries). rho_out (N doubles) may be NULL. */
void orc_ot_potential(orc_ot *otf, double complex *pot, const double complex *psi, double wf_mass, double *rho_out);
/* Same, with density and velocity inputs (dft_ot_potential2, src/ot.c:431-510). vx..vz are overwritten. */
void orc_ot_potential2(orc_ot *otf, double complex *pot, const double *rho, double *vx, double *vy, double *vz);
/* Overwrites eps with the OT energy density (kinetic part excluded). */
void orc_ot_energy_density(orc_ot *otf, double *eps, const double complex *psi, double wf_mass);
/* Same with the density and the velocity given (the backflow part takes v as
   is; the other parts depend on rho only). */
void orc_ot_energy_density_rv(orc_ot *otf, double *eps, const double *rho, const double *vx, const double *vy,
                              const double *vz);
/* Tests only: plant a numbered mistake in one term (0 = none), so that the
   tests can show that the checks pinning the oracle catch it. */
void orc_set_plant(int which);

/* Backward of E[rho]: exact discrete dE/drho of the LJ, c2/c3 and KC energy. */
void orc_ot_energy_gradient(orc_ot *otf, double *grad, const double *rho);
/* E[rho] (LJ, c2/c3, KC parts; h^3 sum eps). */
double orc_ot_energy_rho(orc_ot *otf, const double *rho);

/* libgrid-equivalent primitives used on the path. */
void orc_fft3d(double complex *a, int nx, int ny, int nz, int sign);
void orc_wf_density(const double complex *psi, double *rho, int64_t n);
void orc_wf_velocity(const double complex *psi, double *vx, double *vy, double *vz,
                     int nx, int ny, int nz, double step, double mass, double eps);
void orc_gradient(const double *f, double *gx, double *gy, double *gz, int nx, int ny, int nz, double step);
/* One Strang step psi <- e^{-iV dt/2} F^-1 e^{-i k^2 dt / 2M} F e^{-iV dt/2} psi, complex dt. */
void orc_wf_propagate(double complex *psi, const double complex *pot, double complex dt,
                      int nx, int ny, int nz, double step, double mass);
double orc_wf_norm(const double complex *psi, int64_t n, double step);
double orc_wf_kinetic_energy(const double complex *psi, int nx, int ny, int nz, double step, double mass);
double orc_integral(const double *f, int64_t n, double step);

/* Adaptive cell-average map of one of the path's kernels (centred sampling). which:
   0 = LJ (src/common.c:53), 1 = spherical average (src/common.c:131),
   2 = gaussian F (src/common.c:193), 3 = backflow U_J (src/ot-private.h:11). */
void orc_kernel_map(orc_ot *otf, int which, double *out, int min_sub, int max_sub);

/* Bulk closed forms, src/helium-ot-bulk.c:21-258, 379-463. */
/* Ideal Bose gas of the thermal models (tests): energy per volume and its
   density derivative at density rho. */
double orc_ig_energy(const orc_ot *otf, double rho);
double orc_ig_dEdRho(const orc_ot *otf, double rho);
double orc_ig_dEdRho_energy(const orc_ot *otf, double rho);
double orc_bulk_energy(const orc_ot *otf, double rho);
double orc_bulk_dEdRho(const orc_ot *otf, double rho);
double orc_bulk_pressure(const orc_ot *otf, double rho);
double orc_bulk_dPdRho(const orc_ot *otf, double rho);
double orc_bulk_density_pressurized(const orc_ot *otf, double pressure);
double orc_bulk_chempot_pressurized(const orc_ot *otf, double pressure);
double orc_bulk_compressibility(const orc_ot *otf, double rho);
double orc_bulk_sound_speed(const orc_ot *otf, double rho);
double orc_bulk_istatic(const orc_ot *otf, double k, double rho0);
double orc_bulk_dispersion(const orc_ot *otf, double k, double rho0);

/* Whole-run drivers (CPU baseline and golden trajectories). */
/* Predict-correct step, examples/imag_time/imag_time.c:231-239. vext may be NULL.
   Returns nothing; psi updated in place. work: 2N complex scratch. */
void orc_pc_step(orc_ot *otf, double complex *psi, const double *vext, double mu0,
                 double complex dt, double wf_mass, double complex *work);
/* Single Strang step with potential (doc/libdft-manual.tex:496-504). */
void orc_single_step(orc_ot *otf, double complex *psi, const double *vext, double mu0,
                     double complex dt, double wf_mass, double complex *work);
/* Energy, norm and mu of psi: out[0] = E_total, out[1] = N, out[2] = E_kin, out[3] = E_OT,
   out[4] = E_ext, out[5] = mu = (E_kin + int rho Re V_OT + E_ext) / N. */
void orc_observables(orc_ot *otf, const double complex *psi, const double *vext, double wf_mass, double *out);

double orc_wall_time(void);
int orc_num_threads(void);

#endif
