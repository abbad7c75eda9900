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

template <typename T> __device__ inline T clampv(T v, T lim) {
  if (v > lim) v = lim;
  if (v < -lim) v = -lim;
  return v;
}


}  // namespace dft
