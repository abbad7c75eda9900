// Pointwise helpers shared by the OT kernels.
#pragma once
#include "common.hpp"

namespace dft {

// --------------------------------------------------------- pointwise math --
template <typename T> struct PW {
  int nx, ny, nz;
  T inv2h;                 // 1 / 2h
  T invM;                  // 1 / wavefunction mass (velocity)
  T eps_v;                 // density epsilon in the velocity
  T c2, c3;                // local correlation
  T inv_rho0s;             // 1 / rho_0s
  T ckc;                   // alpha_s / (2 mass)
  T k23;                   // KC terms 2,3 coefficient
  T mass;                  // functional mass
  T div_eps;               // backflow density epsilon
  T cbf;                   // backflow weight
  T vmax, bfmax;           // clamps
  int clamp_v, clamp_bf;
  T hdC, hdB, hdRm;        // high-density penalty C, beta, rho_m (hdC = 0: off)
  T bfXi, bfRho;           // backflow density switch xi, rho_bf
  int bfhd;                // switch on (HD or HD2 with backflow)
};

struct Nbr {
  int64_t p[3], m[3];
};
__device__ inline Nbr neighbours(int64_t c, int nx, int ny, int nz) {
  const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
  const int64_t sx = (int64_t)ny * nz;
  Nbr r;
  r.p[0] = c + (i == nx - 1 ? -(int64_t)(nx - 1) * sx : sx);
  r.m[0] = c + (i == 0 ? (int64_t)(nx - 1) * sx : -sx);
  r.p[1] = c + (j == ny - 1 ? -(int64_t)(ny - 1) * nz : (int64_t)nz);
  r.m[1] = c + (j == 0 ? (int64_t)(ny - 1) * nz : -(int64_t)nz);
  r.p[2] = c + (k == nz - 1 ? -(int64_t)(nz - 1) : 1);
  r.m[2] = c + (k == 0 ? (int64_t)(nz - 1) : -1);
  return r;
}

// High-density penalty of Ancilotto, Barranco, Caupin, Mayol and Pi, PRB 72,
// 214522 (2005), the term the reference adds for DFT_OT_HD / DFT_OT_HD2
// (src/ot.c:374-378, 814-819; src/ot-energy.c:96-99):
//   eps = C rho (1 + tanh(beta (rho - rho_m))),
//   V = d eps / d rho = C s + C beta rho s t,
//   dV/d rho = 2 C beta s t (1 - beta rho tanh),
// with s = 1 + tanh = 2 / (1 + e^{-2x}) and t = 1 - tanh = 2 / (1 + e^{2x}),
// x = beta (rho - rho_m).  s and t are formed directly, so that the penalty
// stays accurate where it is tiny (s ~ 1e-12 at the liquid's density) and
// an overflowing exponential gives 0, not NaN.
template <typename T> struct HdST {
  T s, t;
};
template <typename T> __device__ inline HdST<T> hd_st(T r, const PW<T> &pw) {
  const T x = pw.hdB * (r - pw.hdRm);
  return {(T)2 / ((T)1 + exp((T)-2 * x)), (T)2 / ((T)1 + exp((T)2 * x))};
}
template <typename T> __device__ inline T hd_eps(T r, const PW<T> &pw) {
  return pw.hdC * r * hd_st(r, pw).s;
}
template <typename T> __device__ inline T hd_pot(T r, const PW<T> &pw) {
  const HdST<T> q = hd_st(r, pw);
  return pw.hdC * q.s * ((T)1 + pw.hdB * r * q.t);
}
template <typename T> __device__ inline T hd_dpot(T r, const PW<T> &pw) {
  const HdST<T> q = hd_st(r, pw);
  return (T)2 * pw.hdC * pw.hdB * q.s * q.t * ((T)1 - pw.hdB * r * (q.s - (T)1));
}

// Backflow density switch of HD / HD2 with backflow (src/ot.c:843-982,
// src/ot-energy.c:221-227): the backflow takes rho~ = G(rho) rho wherever it
// uses the density (A, B, C, the energy weight, the gradient half of
// Im V_BF), multiplies Re V_BF by d(G rho)/d rho and the divergence half of
// Im V_BF by G.  G is libgrid's (absent) grid_func1, restated from the
// published form (Ancilotto et al., Int. Rev. Phys. Chem. 36, 621 (2017)):
//   G = (1 - tanh(x)) / 2 = 1 / (1 + e^{2x}),  x = xi (rho - rho_bf),
//   G' = -2 xi e / (1 + e)^2 with e = e^{-2|x|},
// formed from e so that no exponential overflows; G = 1 exactly at liquid
// densities (x ~ -100).
template <typename T> __device__ inline T bf_e(T r, const PW<T> &pw) {
  return exp((T)-2 * fabs(pw.bfXi * (r - pw.bfRho)));
}
template <typename T> __device__ inline T bf_g(T r, const PW<T> &pw) {
  const T e = bf_e(r, pw);
  return r > pw.bfRho ? e / ((T)1 + e) : (T)1 / ((T)1 + e);
}
template <typename T> __device__ inline T bf_rt(T r, const PW<T> &pw) { return bf_g(r, pw) * r; }
// d(G rho)/d rho = G + rho G'
template <typename T> __device__ inline T bf_dg(T r, const PW<T> &pw) {
  const T e = bf_e(r, pw), q = (T)1 + e;
  return (r > pw.bfRho ? e / q : (T)1 / q) - (T)2 * pw.bfXi * r * e / (q * q);
}

template <typename T> __device__ inline T clampv(T v, T lim) {
  if (v > lim) v = lim;
  if (v < -lim) v = -lim;
  return v;
}


}  // namespace dft
