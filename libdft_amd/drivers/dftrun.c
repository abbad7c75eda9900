/*
 * dftrun -- C driver for the OT workloads, written against the C ABI the way
 * the reference drivers are written against libdft + libgrid
 * (examples/imag_time/imag_time.c:175-265, examples/dynamics/dynamics.c:62-168).
 *
 *   dftrun <workload> [--steps K] [--report R] [--prec fp32|fp64] [--single]
 *
 * workloads:
 *   bulk64     plain OT, 64^3, h = 0.5 A, psi = sqrt(0.0218 A^-3)(1 + 0.01 U(-1,1)),
 *              imaginary time dt = 10 fs, renormalised to the initial norm every step
 *   droplet128 full OT (KC+BF), 128^3, h = 0.4 A, N = 1000 droplet with tanh profile,
 *              gaussian well -50 K exp(-r^2 / 2 (2 A)^2) at the centre, real time dt = 5 fs
 *   vortex256  full OT, 256^3, h = 0.3 A, bulk rho0 with a straight vortex line
 *              (phase atan2(y, x) along z), 1e-3 density noise, real time dt = 2 fs
 *   bubble512  full OT, 512^3, h = 0.4 A, bulk rho0 relaxed around a cavity carved by
 *              1000 K (1 - tanh((r - 17 A)/0.5 A))/2, then released: real time dt = 2 fs
 *
 * Uniform deviates: splitmix64, seed 20261202.  Each report line is JSON:
 * step, energy (K), norm, E/N (K), mu (K), and wall time.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "libdft_mi355x.h"

static uint64_t sm_state;
static double splitmix_uniform(void) {
  uint64_t z = (sm_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return -1.0 + 2.0 * (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
  const char *name;
  int n;
  double step_ang, dt_fs;
  int model, imag, steps, report;
} workload;

static const workload WL[] = {
    {"bulk64", 64, 0.5, 10.0, DFT_OT_PLAIN, 1, 200, 20},
    {"droplet128", 128, 0.4, 5.0, DFT_OT_KC | DFT_OT_BACKFLOW, 0, 1000, 100},
    {"vortex256", 256, 0.3, 2.0, DFT_OT_KC | DFT_OT_BACKFLOW, 0, 500, 50},
    {"bubble512", 512, 0.4, 2.0, DFT_OT_KC | DFT_OT_BACKFLOW, 0, 2000, 100},
};

static double coord(int i, int n, double h) { return (double)(i - n / 2) * h; }

/* Total energy (kinetic + OT + external), norm and mu = <H>/N. */
static void observables(dft_ot_functional *otf, wf *gwf, rgrid *vext, cgrid *pot, rgrid *rho, double *E,
                        double *N, double *mu) {
  double ekin = grid_wf_energy(gwf, NULL);
  double eot = dft_ot_energy(otf, gwf);
  double eext = 0.0;
  if (vext) {
    grid_wf_density(gwf, rho);
    eext = rgrid_integral_of_product(rho, vext);
  }
  cgrid_zero(pot);
  dft_ot_potential(otf, pot, gwf);
  double ev = grid_wf_potential_energy(gwf, pot);
  *N = grid_wf_norm(gwf);
  *E = ekin + eot + eext;
  *mu = (ekin + ev + eext) / *N;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s bulk64|droplet128|vortex256|bubble512 [--steps K] [--report R] "
                    "[--prec fp32|fp64] [--single]\n", argv[0]);
    return 1;
  }
  const workload *w = NULL;
  for (size_t i = 0; i < sizeof(WL) / sizeof(WL[0]); i++)
    if (!strcmp(argv[1], WL[i].name)) w = &WL[i];
  if (!w) {
    fprintf(stderr, "unknown workload %s\n", argv[1]);
    return 1;
  }
  int steps = w->steps, report = w->report, prec = GRID_FP64, single = 0;
  for (int a = 2; a < argc; a++) {
    if (!strcmp(argv[a], "--steps") && a + 1 < argc) steps = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--report") && a + 1 < argc) report = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--prec") && a + 1 < argc) prec = strcmp(argv[++a], "fp32") ? GRID_FP64 : GRID_FP32;
    else if (!strcmp(argv[a], "--single")) single = 1;
  }
  const int n = w->n;
  const double h = w->step_ang / GRID_AUTOANG;
  const int64_t N3 = (int64_t)n * n * n;

  grid_set_precision(prec);
  wf *gwf = grid_wf_alloc(n, n, n, h, DFT_HELIUM_MASS, WF_PERIODIC_BOUNDARY, WF_2ND_ORDER_FFT, "gwf");
  wf *gwfp = grid_wf_clone(gwf, "gwfp");
  dft_ot_functional *otf = dft_ot_alloc(w->model, gwf, DFT_MIN_SUBSTEPS, DFT_MAX_SUBSTEPS);
  if (!gwf || !gwfp || !otf) {
    fprintf(stderr, "allocation failed\n");
    return 1;
  }
  cgrid *pot = cgrid_clone(gwf->grid, "potential");
  rgrid *rho = rgrid_clone(otf->density, "rho");
  rgrid *vext = NULL;
  double mu0 = dft_ot_bulk_chempot_pressurized(otf, 0.0);

  /* initial state and external potential (host, uploaded once) */
  REAL_COMPLEX *psi = malloc(sizeof(REAL_COMPLEX) * (size_t)N3);
  REAL *ve = malloc(sizeof(REAL) * (size_t)N3);
  sm_state = 20261202ULL;
  const double a3 = GRID_AUTOANG * GRID_AUTOANG * GRID_AUTOANG;
  int have_vext = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        const int64_t c = ((int64_t)i * n + j) * n + k;
        const double x = coord(i, n, h), y = coord(j, n, h), z = coord(k, n, h), r = sqrt(x * x + y * y + z * z);
        ve[c] = 0.0;
        if (!strcmp(w->name, "bulk64")) {
          psi[c] = sqrt(0.0218 * a3) * (1.0 + 0.01 * splitmix_uniform());
        } else if (!strcmp(w->name, "droplet128")) {
          const double rho0 = 0.0218 * a3, R = cbrt(3.0 * 1000.0 / (4.0 * M_PI * rho0));
          psi[c] = sqrt(0.5 * rho0 * (1.0 - tanh((r - R) / (1.0 / GRID_AUTOANG))));
          const double wd = 2.0 / GRID_AUTOANG;
          ve[c] = (-50.0 / GRID_AUTOK) * exp(-r * r / (2.0 * wd * wd));
          have_vext = 1;
        } else if (!strcmp(w->name, "vortex256")) {
          psi[c] = sqrt(otf->rho0 * (1.0 + 1e-3 * splitmix_uniform())) * cexp(I * atan2(y, x));
        } else {
          psi[c] = sqrt(otf->rho0);
          ve[c] = (1000.0 / GRID_AUTOK) * 0.5 * (1.0 - tanh((r - 17.0 / GRID_AUTOANG) / (0.5 / GRID_AUTOANG)));
          have_vext = 1;
        }
      }
  cgrid_copy_from_host(gwf->grid, psi);
  if (have_vext) {
    vext = rgrid_clone(otf->density, "vext");
    rgrid_copy_from_host(vext, ve);
  }
  free(psi);
  free(ve);
  gwf->norm = grid_wf_norm(gwf);

  if (!strcmp(w->name, "bubble512")) {
    /* carve the cavity: 200 imaginary-time steps in the repulsive potential, then release it */
    for (int it = 0; it < 200; it++) dft_ot_propagate_pc(otf, gwf, gwfp, vext, mu0, -I * 10.0 / GRID_AUTOFS);
    rgrid_free(vext);
    vext = NULL;
  }

  const REAL_COMPLEX dt = w->imag ? -I * w->dt_fs / GRID_AUTOFS : w->dt_fs / GRID_AUTOFS;
  grid_timer timer;
  grid_timer_start(&timer);
  for (int it = 0; it <= steps; it++) {
    if (it % report == 0 || it == steps) {
      double E, Nn, mu;
      grid_synchronize();
      double wall = grid_timer_wall_clock_time(&timer);
      observables(otf, gwf, vext, pot, rho, &E, &Nn, &mu);
      printf("{\"workload\": \"%s\", \"prec\": \"%s\", \"step\": %d, \"energy_K\": %.15e, \"norm\": %.15e, "
             "\"energy_per_atom_K\": %.15e, \"mu_K\": %.15e, \"wall_s\": %.3f}\n",
             w->name, prec == GRID_FP32 ? "fp32" : "fp64", it, E * GRID_AUTOK, Nn, E / Nn * GRID_AUTOK,
             mu * GRID_AUTOK, wall);
      fflush(stdout);
      grid_timer_start(&timer);
    }
    if (it == steps) break;
    if (single) dft_ot_propagate(otf, gwf, vext, mu0, dt);
    else dft_ot_propagate_pc(otf, gwf, gwfp, vext, mu0, dt);
    if (w->imag) grid_wf_normalize(gwf);
  }
  grid_synchronize();
  if (vext) rgrid_free(vext);
  rgrid_free(rho);
  cgrid_free(pot);
  dft_ot_free(otf);
  grid_wf_free(gwfp);
  grid_wf_free(gwf);
  return 0;
}
