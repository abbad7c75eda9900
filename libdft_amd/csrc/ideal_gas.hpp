// Ideal Bose gas term of the thermal OT models (T400MK ... T3000MK), shared by
// the host bulk closed forms and the device kernels.  The reference states it in
// src/common.c and uses it for the bulk closed forms (src/helium-ot-bulk.c:30-31,
// 54-55); the potential and energy pass the density through libgrid's
// grid_func6a/6b with the same mass, temperature and c4 (src/ot.c:804-808,
// src/ot-energy.c:101-104), here taken to be these same functions:
//   lambda^3 = (2 pi / (M k_B T))^{3/2}                       (common.c:253-260)
//   g_s(z) = sum_{k=1}^{256} z^k / k^s                       (common.c:293-306)
//   z(x): the golden-section minimum of (x - g_{3/2}(z))^2 on [0, 1] down to a
//         bracket of 1e-6; 1 when x >= g_{3/2}(1)             (common.c:400-437)
//   energy per volume c4 k_B T (rho ln z - g_{5/2}(z) / lambda^3)   (489-497)
//   potential c4 k_B T (ln z + x / g_{1/2}(z) - g_{3/2}(z) / g_{1/2}(z)),
//   x = rho lambda^3, or -c4 k_B (T / lambda^3) g_{3/2}(1) when x >= g_{3/2}(1)
//   (510-530; the last two terms of the first form correct ln z for the bracket
//   the search stops at).
// The series divides by k^s from a table of pow(k, s) built on the host, so that
// host and device sum the same terms.
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define DFT_IG_FN __host__ __device__ inline
#else
#define DFT_IG_FN inline
#endif

namespace dft {
namespace ig {

constexpr int kTerms = 256;

// pow(k, s) for k = 1..256 and s = 1/2, 3/2, 5/2 (rows 0, 1, 2)
inline void fill_table(double *t) {
  const double s[3] = {0.5, 1.5, 2.5};
  for (int r = 0; r < 3; ++r)
    for (int k = 1; k <= kTerms; ++k) t[r * kTerms + k - 1] = std::pow((double)k, s[r]);
}

DFT_IG_FN double series(double z, const double *pk) {
  double v = 0.0, zk = 1.0;
  for (int k = 0; k < kTerms; ++k) {
    zk *= z;
    v += zk / pk[k];
  }
  return v;
}

// z with g_{3/2}(z) = x (g1 = g_{3/2}(1))
DFT_IG_FN double fugacity(double x, const double *p32, double g1) {
  if (x >= g1) return 1.0;
  const double gr = (std::sqrt(5.0) + 1.0) / 2.0;
  double a = 0.0, b = 1.0, c = b - (b - a) / gr, d = a + (b - a) / gr;
  while (std::fabs(c - d) > 1e-6) {
    const double tc = x - series(c, p32), td = x - series(d, p32);
    if (tc * tc < td * td) b = d;
    else a = c;
    c = b - (b - a) / gr;
    d = a + (b - a) / gr;
  }
  return (b + a) / 2.0;
}

// Parameters of one functional: c4 k_B T, lambda^3, g_{3/2}(1) and the table.
struct Params {
  double ckt, l3, g1;
  const double *tab;
};

DFT_IG_FN double energy(double rho, const Params &p) {
  const double z = fugacity(rho * p.l3, p.tab + kTerms, p.g1);
  return p.ckt * (rho * std::log(z) - series(z, p.tab + 2 * kTerms) / p.l3);
}

DFT_IG_FN double potential(double rho, const Params &p) {
  const double x = rho * p.l3;
  if (x >= p.g1) return -p.ckt / p.l3 * p.g1;
  const double z = fugacity(x, p.tab + kTerms, p.g1);
  const double g12 = series(z, p.tab), g32 = series(z, p.tab + kTerms);
  return p.ckt * (std::log(z) + x / g12 - g32 / g12);
}

// Derivatives for the reverse passes.  The search returns a z that is constant
// between the points where its branch decisions change (steps of ~1e-6 in z),
// so, between them, d(energy)/d rho = c4 k_B T ln z and d(potential)/d rho =
// c4 k_B T lambda^3 / g_{1/2}(z); both are 0 at saturation.  (The steps
// themselves move the energy and the potential by about (1e-6 / z)^2 of their
// values: below 1e-10 for z > 0.1.)
DFT_IG_FN double energy_derivative(double rho, const Params &p) {
  const double x = rho * p.l3;
  if (x >= p.g1) return 0.0;
  return p.ckt * std::log(fugacity(x, p.tab + kTerms, p.g1));
}
DFT_IG_FN double potential_derivative(double rho, const Params &p) {
  const double x = rho * p.l3;
  if (x >= p.g1) return 0.0;
  return p.ckt * p.l3 / series(fugacity(x, p.tab + kTerms, p.g1), p.tab);
}

// lambda^3 of mass M (a.u.) at T kelvin (hbar = 1, k_B = 1 / autok)
inline double lambda3(double mass, double temp, double autok) {
  const double l = std::sqrt(2.0 * M_PI / (mass * (temp / autok)));
  return l * l * l;
}

}  // namespace ig
}  // namespace dft
