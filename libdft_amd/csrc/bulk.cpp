// Host FP64 parts of the API: bulk closed forms (src/helium-ot-bulk.c), with the
// ideal Bose gas term of the thermal models (c4 != 0), and the kernel functions
// (src/common.c, src/ot-private.h).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../include/libdft_mi355x.h"
#include "ideal_gas.hpp"

namespace {
const double *ig_table() {
  static const struct Tab {
    double t[3 * dft::ig::kTerms];
    Tab() { dft::ig::fill_table(t); }
  } tab;
  return tab.t;
}
dft::ig::Params ig_params(const dft_ot_functional *otf) {
  const double *tab = ig_table();
  dft::ig::Params p;
  p.tab = tab;
  p.ckt = otf->c4 * (otf->temp / GRID_AUTOK);
  p.l3 = dft::ig::lambda3(otf->mass, otf->temp, GRID_AUTOK);
  p.g1 = dft::ig::series(1.0, tab + dft::ig::kTerms);
  return p;
}
}  // namespace

extern "C" {

// The ideal gas functions of src/common.c:245-530 (the thermal models' gas term).
REAL dft_common_lwl3(REAL mass, REAL temp) { return dft::ig::lambda3(mass, temp, GRID_AUTOK); }
REAL dft_common_fit_g12(REAL z) { return dft::ig::series(z, ig_table()); }
REAL dft_common_fit_g32(REAL z) { return dft::ig::series(z, ig_table() + dft::ig::kTerms); }
REAL dft_common_fit_g52(REAL z) { return dft::ig::series(z, ig_table() + 2 * dft::ig::kTerms); }
REAL dft_common_fit_z(REAL val) {
  const double *p32 = ig_table() + dft::ig::kTerms;
  return dft::ig::fugacity(val, p32, dft::ig::series(1.0, p32));
}
REAL dft_common_bose_idealgas_energy(REAL rhop, void *params) {
  return dft::ig::energy(rhop, ig_params((const dft_ot_functional *)params));
}
REAL dft_common_bose_idealgas_dEdRho(REAL rhop, void *params) {
  return dft::ig::potential(rhop, ig_params((const dft_ot_functional *)params));
}
// classical ideal gas, src/common.c:452-480 (density floor 1e-14 inside the logarithm)
REAL dft_common_classical_idealgas_energy(REAL rhop, void *params) {
  const dft_ot_functional *otf = (const dft_ot_functional *)params;
  const double l3 = dft::ig::lambda3(otf->mass, otf->temp, GRID_AUTOK);
  return -rhop * (otf->temp / GRID_AUTOK) * (1.0 - std::log(rhop * l3 + 1e-14));
}
REAL dft_common_classical_idealgas_dEdRho(REAL rhop, void *params) {
  const dft_ot_functional *otf = (const dft_ot_functional *)params;
  const double l3 = dft::ig::lambda3(otf->mass, otf->temp, GRID_AUTOK);
  return (otf->temp / GRID_AUTOK) * std::log(rhop * l3 + 1e-14);
}

// src/helium-ot-bulk.c:21-33
REAL dft_ot_bulk_energy(dft_ot_functional *otf, REAL rho) {
  double e = 0.5 * otf->b * rho * rho + 0.5 * otf->c2 * rho * rho * rho + (1.0 / 3.0) * otf->c3 * rho * rho * rho * rho;
  if (otf->c4 != 0.0) e += dft::ig::energy(rho, ig_params(otf));
  return e;
}

// src/helium-ot-bulk.c:45-57
REAL dft_ot_bulk_dEdRho(dft_ot_functional *otf, REAL rho) {
  double d = otf->b * rho + 1.5 * otf->c2 * rho * rho + (4.0 / 3.0) * otf->c3 * rho * rho * rho;
  if (otf->c4 != 0.0) d += dft::ig::potential(rho, ig_params(otf));
  return d;
}

// src/helium-ot-bulk.c:143-146
REAL dft_ot_bulk_pressure(dft_ot_functional *otf, REAL rho) {
  return rho * dft_ot_bulk_dEdRho(otf, rho) - dft_ot_bulk_energy(otf, rho);
}

// src/helium-ot-bulk.c:157-172 (central difference, step 1e-7)
REAL dft_ot_bulk_dPdRho(dft_ot_functional *otf, REAL rho) {
  const double e = 1e-7;
  return (dft_ot_bulk_pressure(otf, rho + e) - dft_ot_bulk_pressure(otf, rho - e)) / (2.0 * e);
}

// src/helium-ot-bulk.c:185-211: Newton from rho = 1 until |P(rho) - P| < 1e-12.
REAL dft_ot_bulk_density_pressurized(dft_ot_functional *otf, REAL pressure) {
  double rho0 = 1.0, misP = dft_ot_bulk_pressure(otf, rho0) - pressure;
  for (int i = 0; i < 1000; ++i) {
    if (std::fabs(misP) < 1e-12) return rho0;
    rho0 -= misP / dft_ot_bulk_dPdRho(otf, rho0);
    misP = dft_ot_bulk_pressure(otf, rho0) - pressure;
  }
  fprintf(stderr, "libdft: Error in dft_ot_bulk_density_pressurized - Newton-Raphson did not converge.\n");
  abort();
}

REAL dft_ot_bulk_density(dft_ot_functional *otf) { return dft_ot_bulk_density_pressurized(otf, 0.0); }
REAL dft_ot_bulk_chempot(dft_ot_functional *otf) { return dft_ot_bulk_dEdRho(otf, dft_ot_bulk_density(otf)); }
REAL dft_ot_bulk_chempot2(dft_ot_functional *otf) { return dft_ot_bulk_dEdRho(otf, otf->rho0); }
REAL dft_ot_bulk_chempot3(dft_ot_functional *otf, REAL rho0) { return dft_ot_bulk_dEdRho(otf, rho0); }
REAL dft_ot_bulk_chempot_pressurized(dft_ot_functional *otf, REAL pressure) {
  return dft_ot_bulk_dEdRho(otf, dft_ot_bulk_density_pressurized(otf, pressure));
}
REAL dft_ot_bulk_compressibility(dft_ot_functional *otf, REAL rho) {
  return 1.0 / (rho * dft_ot_bulk_dPdRho(otf, rho));
}
REAL dft_ot_bulk_sound_speed(dft_ot_functional *otf, REAL rho) {
  return 1.0 / std::sqrt(otf->mass * rho * dft_ot_bulk_compressibility(otf, rho));
}

// Radial transforms used by the semi-analytic response, src/helium-ot-bulk.c:376-409.
static double ft_lj(const dft_ot_functional *otf, double k) {
  const double sigma = otf->lj_params.sigma, eps = otf->lj_params.epsilon, h = otf->lj_params.h, ks = k * sigma;
  double val = 0.0;
  for (double x = h / sigma; x < 1e3; x += 1e-2) val += std::sin(ks * x) * (std::pow(x, -11.0) - std::pow(x, -5.0));
  return val * 1e-2 * 16.0 * M_PI * eps * sigma * sigma * sigma / ks;
}
static double ft_pi(const dft_ot_functional *otf, double k) {
  const double kh = k * otf->lj_params.h;
  return (3.0 / (kh * kh * kh)) * (std::sin(kh) - kh * std::cos(kh));
}
static double ft_vj(const dft_ot_functional *otf, double k) {
  const dft_ot_bf &p = otf->bf_params;
  const double k2 = k * k, m = std::pow(M_PI, 1.5);
  const double e1 = std::exp(-k2 / (4.0 * p.a1)), e2 = std::exp(-k2 / (4.0 * p.a2));
  return (p.g11 * std::pow(M_PI / p.a1, 1.5) + p.g12 * (6.0 * p.a1 - k2) * m / (4.0 * std::pow(p.a1, 3.5))) * e1 +
         (p.g21 * std::pow(M_PI / p.a2, 1.5) + p.g22 * (6.0 * p.a2 - k2) * m / (4.0 * std::pow(p.a2, 3.5))) * e2;
}

// src/helium-ot-bulk.c:411-434
REAL dft_ot_bulk_dispersion(dft_ot_functional *otf, REAL *k, REAL rho0) {
  if (rho0 < 0.0) return 0.0;
  const double tk = *k;
  if (tk < 1e-3) return 0.0;
  const double lj = ft_lj(otf, tk), pi = ft_pi(otf, tk), tmp = tk * tk / otf->mass;
  double ikai = tmp / 4.0 + rho0 * lj + otf->c2 * (2.0 * pi + pi * pi) * rho0 * rho0 +
                2.0 * otf->c3 * (pi + pi * pi) * rho0 * rho0 * rho0;
  if (otf->model & DFT_OT_KC)
    ikai += -(tmp / 2.0) * otf->alpha_s * rho0 * (1.0 - rho0 / otf->rho_0s) * (1.0 - rho0 / otf->rho_0s) *
            std::exp(-tk * tk * otf->l_g * otf->l_g / 4.0);
  if (otf->model & DFT_OT_BACKFLOW)
    return std::sqrt(tmp * ikai * (1.0 - rho0 * (ft_vj(otf, 0.0) - ft_vj(otf, tk))));
  return std::sqrt(tmp * ikai);
}

// src/helium-ot-bulk.c:445-463
REAL dft_ot_bulk_istatic(dft_ot_functional *otf, REAL *k, REAL rho0) {
  if (rho0 < 0.0) return 0.0;
  const double tk = *k;
  if (tk < 1e-2) return 1e-2;
  const double lj = ft_lj(otf, tk), pi = ft_pi(otf, tk);
  double ikai = tk * tk / (4.0 * otf->mass) + rho0 * lj + otf->c2 * (2.0 * pi + pi * pi) * rho0 * rho0 +
                2.0 * otf->c3 * (pi + pi * pi) * rho0 * rho0 * rho0;
  if (otf->model & DFT_OT_KC)
    ikai += -(1.0 / (2.0 * otf->mass)) * otf->alpha_s * rho0 * (1.0 - rho0 / otf->rho_0s) *
            (1.0 - rho0 / otf->rho_0s) * std::exp(-tk * tk * otf->l_g * otf->l_g / 4.0);
  return ikai;
}

// src/common.c:31-39
REAL dft_common_lj_func(REAL r2, REAL sig, REAL eps) {
  r2 = sig * sig / r2;
  r2 = r2 * r2 * r2;
  return 4.0 * eps * r2 * (r2 - 1.0);
}

// src/common.c:53-66
REAL dft_common_lennard_jones(void *arg, REAL x, REAL y, REAL z) {
  const dft_common_lj *p = (const dft_common_lj *)arg;
  const double r2 = x * x + y * y + z * z;
  if (r2 <= p->h * p->h) return p->cval;
  return dft_common_lj_func(r2, p->sigma, p->epsilon);
}

// src/common.c:131-138
REAL dft_common_spherical_avg(void *arg, REAL x, REAL y, REAL z) {
  const double h = *(const double *)arg;
  return (x * x + y * y + z * z <= h * h) ? 3.0 / (4.0 * M_PI * h * h * h) : 0.0;
}

// src/common.c:173-179
REAL dft_common_spherical_avg_k(void *arg, REAL kx, REAL ky, REAL kz) {
  const double hk = *(const double *)arg * std::sqrt(kx * kx + ky * ky + kz * kz);
  if (hk < 1e-5) return 1.0 - 0.1 * hk * hk;
  return 3.0 * (std::sin(hk) - hk * std::cos(hk)) / (hk * hk * hk);
}

// src/common.c:193-200
REAL dft_common_gaussian(void *arg, REAL x, REAL y, REAL z) {
  const double iw = *(const double *)arg;
  double norm = 0.5 * (2.0 / std::sqrt(M_PI)) * iw;
  norm = norm * norm * norm;
  return norm * std::exp(-(x * x + y * y + z * z) * iw * iw);
}

// ---- map functions for initial states and test potentials
namespace {
REAL_COMPLEX cval(double re, double im) {
  REAL_COMPLEX v;
  __real__ v = re;
  __imag__ v = im;
  return v;
}
constexpr double kCore = 0.05;  // vortex core cut (bohr), src/common.c:1238
}  // namespace

// src/common.c:235-242 (the gaussian of dft_common_gaussian as a complex value)
REAL_COMPLEX dft_common_cgaussian(void *arg, REAL x, REAL y, REAL z) { return cval(dft_common_gaussian(arg, x, y, z), 0.0); }

// src/common.c:1214-1224: sqrt(rho) (1 + a cos(k.r))
REAL_COMPLEX dft_common_planewave(void *arg, REAL x, REAL y, REAL z) {
  const dft_plane_wave *p = (const dft_plane_wave *)arg;
  const double psi = std::sqrt(p->rho);
  return cval(psi + p->a * psi * std::cos(p->kx * x + p->ky * y + p->kz * z), 0.0);
}

// src/common.c:1240-1290: 1 / (2 M r_perp^2), r_perp at least the core cut
static double vortex_pot(const void *param, double rp2) {
  if (rp2 < kCore * kCore) rp2 = kCore * kCore;
  return 1.0 / (2.0 * *(const double *)param * rp2);
}
REAL dft_common_vortex_x(void *param, REAL, REAL y, REAL z) { return vortex_pot(param, y * y + z * z); }
REAL dft_common_vortex_y(void *param, REAL x, REAL, REAL z) { return vortex_pot(param, x * x + z * z); }
REAL dft_common_vortex_z(void *param, REAL x, REAL y, REAL) { return vortex_pot(param, x * x + y * y); }

// src/initial.c:31-89: (u + i v) / |(u, v)|, 0 inside the core
static REAL_COMPLEX vortex1(double u, double v) {
  const double d = std::sqrt(u * u + v * v);
  return d < kCore ? cval(0.0, 0.0) : cval(u / d, v / d);
}
REAL_COMPLEX dft_initial_vortex_x_n1(void *, REAL, REAL y, REAL z) { return vortex1(y, z); }
REAL_COMPLEX dft_initial_vortex_y_n1(void *, REAL x, REAL, REAL z) { return vortex1(x, z); }
REAL_COMPLEX dft_initial_vortex_z_n1(void *, REAL x, REAL y, REAL) { return vortex1(x, y); }

// src/initial.c:91-150: (u + i v)^2 / (u^2 + v^2), 0 where the core test holds;
// the reference tests sqrt(x^2 + y^2) for every axis, and so does this
static REAL_COMPLEX vortex2(double u, double v, double dx, double dy) {
  if (std::sqrt(dx * dx + dy * dy) < kCore) return cval(0.0, 0.0);
  const double u2 = u * u, v2 = v * v;
  return cval((u2 - v2) / (u2 + v2), 2 * u * v / (u2 + v2));
}
REAL_COMPLEX dft_initial_vortex_x_n2(void *, REAL x, REAL y, REAL z) { return vortex2(y, z, x, y); }
REAL_COMPLEX dft_initial_vortex_y_n2(void *, REAL x, REAL y, REAL z) { return vortex2(x, z, x, y); }
REAL_COMPLEX dft_initial_vortex_z_n2(void *, REAL x, REAL y, REAL) { return vortex2(x, y, x, y); }

// src/initial.c:154-161: 0 inside the radius *prm, 1 outside
REAL_COMPLEX dft_initial_bubble(void *prm, REAL x, REAL y, REAL z) {
  return cval(std::sqrt(x * x + y * y + z * z) < *(const double *)prm ? 0.0 : 1.0, 0.0);
}

// src/ot-private.h:11-22
REAL dft_ot_backflow_pot(void *arg, REAL x, REAL y, REAL z) {
  const dft_ot_bf *p = (const dft_ot_bf *)arg;
  const double r2 = x * x + y * y + z * z;
  return (p->g11 + p->g12 * r2) * std::exp(-p->a1 * r2) + (p->g21 + p->g22 * r2) * std::exp(-p->a2 * r2);
}

}  // extern "C"
