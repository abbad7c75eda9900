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

template <typename T> __device__ inline T clampv(T v, T lim) {
  if (v > lim) v = lim;
  if (v < -lim) v = -lim;
  return v;
}


}  // namespace dft
