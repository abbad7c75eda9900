// Fused stage kernels of the OT potential and the split-operator step.
//
// Every z pass works on tiles of z-line pairs (adjacent y lines (x, 2p) and
// (x, 2p+1)), which are contiguous in memory.  The pointwise stages run as the
// prologue or epilogue of these passes, so their inputs and outputs never
// make an extra trip to HBM.  Stencils in x and y read neighbouring lines
// from global memory (they are in L2/MALL: adjacent planes are processed
// together).
//
//   zA    psi -> rho (stored) and, with backflow, q_d = rho v_d, q4 = rho |v|^2;
//         z r2c of {rho, q_x, q_y, q_z, q4}
//   xmid  x forward FFT, k-space products with real kernel spectra and the
//         central-difference symbol i sin(k h)/h, x inverse FFT (a small
//         program: several outputs per input, one multi-term output)
//   zAB   z c2r of {Pi*rho, F*rho, U*rho, U*q_d, U*q4}; local terms, G_d, BF real
//         part, W_d; z r2c of {Q, G_x, G_y, G_z}
//   zBC   z c2r of {J_d, T1}; H, KC terms 1 and 3, BF divergence; z r2c of {H}
//   zC    z c2r of the summed linear terms; V += (API), or phase = e^{-iV dt/2}
//         and the first z pass of the propagator (step)
#pragma once
#include "fft.hpp"
#include "ot_common.hpp"

namespace dft {
namespace fused {

using fft::Shape;

// Line-pair tiles for z passes holding LINES LDS lines per pair.
template <typename T, int N, int LINES> struct ZCfg {
  static constexpr int csize = 2 * sizeof(T);
  static constexpr int PITCH = Shape<N>::PITCH;
  static constexpr int fit(int t) {
    return (t > 1 && t * LINES * PITCH * csize + 2 * N * csize > 80 * 1024) ? fit(t / 2) : t;
  }
  static constexpr int TILE = fit(8);
  static constexpr int BS = 256;
};

// x-middle tiles: columns per block, three LDS buffers.
template <typename T, int N> struct MidCfg {
  static constexpr int csize = 2 * sizeof(T);
  static constexpr int PITCH = Shape<N>::PITCH;
  static constexpr int fit(int t) {
    return (t > 1 && (3 * t * PITCH * csize + 2 * N * csize > 80 * 1024 || t * Shape<N>::TPL > 256)) ? fit(t / 2) : t;
  }
  static constexpr int TILE = fit(16);
  static constexpr int BS = TILE * Shape<N>::TPL;
};

// All block threads transform NL lines (line l at L + l*PITCH), BS threads.
template <typename T, int N, int DIR, int BS>
__device__ inline void lines_fft(cplx<T> *L, const cplx<T> *tw, int NL) {
  constexpr int TPL = Shape<N>::TPL, PITCH = Shape<N>::PITCH, LPB = BS / TPL;
  for (int l0 = 0; l0 < NL; l0 += LPB) {
    const int line = l0 + (int)threadIdx.x % LPB;
    fft::line_fft<T, N, DIR>(L + line * PITCH, tw, (int)threadIdx.x / LPB, line < NL);
  }
}

// Half spectrum of line `which` of a packed pair from Z = FFT(a + ib).
template <typename C> __device__ inline C split(C z, C zm, int which) {
  using T = decltype(z.x);
  return which == 0 ? cmk<C>((T)0.5 * (z.x + zm.x), (T)0.5 * (z.y - zm.y))
                    : cmk<C>((T)0.5 * (z.y + zm.y), (T)0.5 * (zm.x - z.x));
}

// Z(m) = A(m) + i B(m) with the Hermitian extension of the half spectra A, B.
template <typename C> __device__ inline C merge(const C *A, const C *B, int m, int N) {
  C a, b;
  if (m <= N / 2) {
    a = A[m];
    b = B[m];
  } else {
    a = cconj(A[N - m]);
    b = cconj(B[N - m]);
  }
  return cmk<C>(a.x - b.y, a.y + b.x);
}

template <typename C> __device__ inline auto &comp(C &c, int which) { return ((decltype(c.x) *)&c)[which]; }

// v_d = Im(psi* (psi_{+d} - psi_{-d})/2h) / (M (rho + eps)), clamped when enabled.
template <typename T>
__device__ inline void velocity_at(const cplx<T> *__restrict__ psi, int64_t c, const PW<T> &pw, T r, T v[3]) {
  const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
  const cplx<T> s = psi[c];
  const T den = pw.invM / (r + pw.eps_v);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const cplx<T> a = psi[nb.p[d]], b = psi[nb.m[d]];
    const T dx = (a.x - b.x) * pw.inv2h, dy = (a.y - b.y) * pw.inv2h;
    v[d] = (s.x * dy - s.y * dx) * den;
    if (pw.clamp_v) v[d] = clampv(v[d], pw.vmax);
  }
}

struct Ptrs8 {
  void *p[8];
};

// Write NF split half spectra (fields f at lines f*TILE + pr) for np pairs from p0.
template <typename T, int N, int TILE>
__device__ inline void store_split(const cplx<T> *L, const Ptrs8 &outs, int NF, int64_t p0, int np, int P) {
  constexpr int PITCH = Shape<N>::PITCH;
  for (int f = 0; f < NF; ++f) {
    cplx<T> *out = (cplx<T> *)outs.p[f] + p0 * 2 * P;
    for (int e = threadIdx.x; e < np * 2 * P; e += 256) {
      const int pr = e / (2 * P), w = e - pr * 2 * P, which = w / P, k = w - which * P;
      const cplx<T> *ln = L + (f * TILE + pr) * PITCH;
      out[e] = split(ln[k], ln[(N - k) & (N - 1)], which);
    }
  }
}

// Load NIN half-spectrum inputs (rows 2p, 2p+1) as packed complex lines f*TILE + pr.
template <typename T, int N, int TILE>
__device__ inline void load_merge(cplx<T> *L, const Ptrs8 &ins, int NIN, int64_t p0, int np, int P) {
  constexpr int PITCH = Shape<N>::PITCH;
  for (int i = 0; i < NIN; ++i) {
    const cplx<T> *in = (const cplx<T> *)ins.p[i];
    for (int m = threadIdx.x; m < np * N; m += 256) {
      const int pr = m / N, idx = m - pr * N;
      const cplx<T> *A = in + (p0 + pr) * 2 * P;
      L[(i * TILE + pr) * PITCH + idx] = merge(A, A + P, idx, N);
    }
  }
}

// -------------------------------------------------------------------- zA --
template <typename T, int N, bool BF>
__global__ void __launch_bounds__(256)
    k_zA(const cplx<T> *__restrict__ psi, T *__restrict__ rho, Ptrs8 outs, int64_t npairs, int P, PW<T> pw,
         const cplx<T> *__restrict__ g_tw) {
  constexpr int NF = BF ? 5 : 1;
  using Cfg = ZCfg<T, N, NF>;
  constexpr int TILE = Cfg::TILE, PITCH = Cfg::PITCH;
  __shared__ cplx<T> L[NF * TILE * PITCH];
  __shared__ cplx<T> s_tw[2 * N];
  const int64_t p0 = (int64_t)blockIdx.x * TILE;
  const int np = (int)((npairs - p0) < TILE ? (npairs - p0) : TILE);
  fft::load_twiddles<T, N>(s_tw, g_tw);
  for (int e = threadIdx.x; e < np * 2 * N; e += 256) {
    const int pr = e / (2 * N), w = e - pr * 2 * N, which = w / N, z = w - which * N;
    const int64_t c = p0 * 2 * N + e;
    const cplx<T> s = psi[c];
    const T r = s.x * s.x + s.y * s.y;
    rho[c] = r;
    comp(L[(0 * TILE + pr) * PITCH + z], which) = r;
    if constexpr (BF) {
      T v[3];
      velocity_at(psi, c, pw, r, v);
      comp(L[(1 * TILE + pr) * PITCH + z], which) = v[0] * r;
      comp(L[(2 * TILE + pr) * PITCH + z], which) = v[1] * r;
      comp(L[(3 * TILE + pr) * PITCH + z], which) = v[2] * r;
      comp(L[(4 * TILE + pr) * PITCH + z], which) = (v[2] * v[2] + v[1] * v[1] + v[0] * v[0]) * r;
    }
  }
  __syncthreads();
  lines_fft<T, N, -1, 256>(L, s_tw, NF * TILE);
  store_split<T, N, TILE>(L, outs, NF, p0, np, P);
}

// ------------------------------------------------------------------ xmid --
struct MidTerm {
  int in, kern, deriv, pad;
  double scale;
};
struct MidOut {
  int nterms, pad;
  MidTerm t[3];
  void *out;
};
constexpr int kMidMaxIn = 6, kMidMaxOut = 8;
struct MidProg {
  int nin, nout, multi, pad;
  const void *in[kMidMaxIn];
  MidOut o[kMidMaxOut];
  const void *kern[4];
  const void *sinx, *siny, *sinz;
};

template <typename T>
__device__ inline cplx<T> mid_coef(const MidTerm &t, const MidProg &pg, int64_t gi, int el, int o, int kz, cplx<T> a) {
  T s = (T)t.scale;
  if (t.kern >= 0) s *= ((const T *)pg.kern[t.kern])[gi];
  if (t.deriv >= 0) {
    s *= t.deriv == 0 ? ((const T *)pg.sinx)[el] : (t.deriv == 1 ? ((const T *)pg.siny)[o] : ((const T *)pg.sinz)[kz]);
    return cmk<cplx<T>>(-s * a.y, s * a.x);
  }
  return cmk<cplx<T>>(s * a.x, s * a.y);
}

// Spectra are [nx][ny][P] after the y pass; the x line (ky = o, kz = c) has stride ny*P.
template <typename T, int N>
__global__ void __launch_bounds__((MidCfg<T, N>::BS))
    k_xmid(MidProg pg, int ny, int P, int ntile, const cplx<T> *__restrict__ g_tw) {
  using Cfg = MidCfg<T, N>;
  constexpr int TILE = Cfg::TILE, PITCH = Cfg::PITCH, BS = Cfg::BS;
  __shared__ cplx<T> A[TILE * PITCH];
  __shared__ cplx<T> B[TILE * PITCH];
  __shared__ cplx<T> Cc[TILE * PITCH];
  __shared__ cplx<T> s_tw[2 * N];
  // consecutive column tiles of one ky on one XCD, back to back (shared L2 lines)
  int b = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
  const int o = b / ntile, ct = b - o * ntile, c0 = ct * TILE;
  const int ncol = (P - c0) < TILE ? (P - c0) : TILE;
  const int64_t es = (int64_t)ny * P, base = (int64_t)o * P + c0;
  fft::load_twiddles<T, N>(s_tw, g_tw);
  if (pg.multi >= 0)
    for (int e = threadIdx.x; e < TILE * PITCH; e += BS) Cc[e] = cmk<cplx<T>>(0, 0);
  for (int i = 0; i < pg.nin; ++i) {
    const cplx<T> *in = (const cplx<T> *)pg.in[i];
    for (int e = threadIdx.x; e < TILE * N; e += BS) {
      const int col = e % TILE, el = e / TILE;
      if (col < ncol) A[col * PITCH + el] = in[base + el * es + col];
    }
    __syncthreads();
    lines_fft<T, N, -1, BS>(A, s_tw, TILE);
    if (pg.multi >= 0) {
      const MidOut &mo = pg.o[pg.multi];
      for (int t = 0; t < mo.nterms; ++t) {
        if (mo.t[t].in != i) continue;
        for (int e = threadIdx.x; e < TILE * N; e += BS) {
          const int col = e % TILE, el = e / TILE;
          if (col >= ncol) continue;
          const cplx<T> v = mid_coef<T>(mo.t[t], pg, base + el * es + col, el, o, c0 + col, A[col * PITCH + el]);
          Cc[col * PITCH + el] = cadd(Cc[col * PITCH + el], v);
        }
      }
    }
    for (int m = 0; m < pg.nout; ++m) {
      if (m == pg.multi || pg.o[m].t[0].in != i) continue;
      const MidTerm tm = pg.o[m].t[0];
      for (int e = threadIdx.x; e < TILE * N; e += BS) {
        const int col = e % TILE, el = e / TILE;
        if (col < ncol)
          B[col * PITCH + el] = mid_coef<T>(tm, pg, base + el * es + col, el, o, c0 + col, A[col * PITCH + el]);
      }
      __syncthreads();
      lines_fft<T, N, +1, BS>(B, s_tw, TILE);
      cplx<T> *out = (cplx<T> *)pg.o[m].out;
      for (int e = threadIdx.x; e < TILE * N; e += BS) {
        const int col = e % TILE, el = e / TILE;
        if (col < ncol) out[base + el * es + col] = B[col * PITCH + el];
      }
      __syncthreads();
    }
    __syncthreads();
  }
  if (pg.multi >= 0) {
    __syncthreads();
    lines_fft<T, N, +1, BS>(Cc, s_tw, TILE);
    cplx<T> *out = (cplx<T> *)pg.o[pg.multi].out;
    for (int e = threadIdx.x; e < TILE * N; e += BS) {
      const int col = e % TILE, el = e / TILE;
      if (col < ncol) out[base + el * es + col] = Cc[col * PITCH + el];
    }
  }
}

// ------------------------------------------------------------------- zAB --
struct ZABArgs {
  Ptrs8 in;       // rhobar, [rhot], [A, Bx, By, Bz, C]
  Ptrs8 out;      // Q, [G_x, G_y, G_z]
  const void *rho, *psi;
  void *rhot, *vre, *vim, *wx, *wy, *wz;
};

template <typename T, int N, bool KC, bool BF>
__global__ void __launch_bounds__(256) k_zAB(ZABArgs a, int64_t npairs, int P, PW<T> pw, const cplx<T> *__restrict__ g_tw) {
  constexpr int NIN = 1 + (KC ? 1 : 0) + (BF ? 5 : 0), NOUT = 1 + (KC ? 3 : 0);
  constexpr int NL = NIN > NOUT ? NIN : NOUT;
  using Cfg = ZCfg<T, N, NL>;
  constexpr int TILE = Cfg::TILE, PITCH = Cfg::PITCH;
  constexpr int IA = KC ? 2 : 1;  // first backflow input
  __shared__ cplx<T> L[NL * TILE * PITCH];
  __shared__ cplx<T> s_tw[2 * N];
  const int64_t p0 = (int64_t)blockIdx.x * TILE;
  const int np = (int)((npairs - p0) < TILE ? (npairs - p0) : TILE);
  fft::load_twiddles<T, N>(s_tw, g_tw);
  load_merge<T, N, TILE>(L, a.in, NIN, p0, np, P);
  __syncthreads();
  lines_fft<T, N, +1, 256>(L, s_tw, NIN * TILE);
  const T *rho = (const T *)a.rho;
  for (int e = threadIdx.x; e < np * 2 * N; e += 256) {
    const int pr = e / (2 * N), w = e - pr * 2 * N, which = w / N, z = w - which * N;
    const int64_t c = p0 * 2 * N + e;
    auto in_at = [&](int i) -> T { return comp(L[(i * TILE + pr) * PITCH + z], which); };
    const T r = rho[c];
    const T rb = in_at(0);
    T vre = rb * rb * (pw.c2 * (T)0.5) + rb * rb * rb * (pw.c3 / (T)3);
    const T Q = (rb * pw.c2 + rb * rb * pw.c3) * r;
    T g[3] = {0, 0, 0}, G[3] = {0, 0, 0};
    if constexpr (KC || BF) {
      const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
#pragma unroll
      for (int d = 0; d < 3; ++d) g[d] = (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h;
    }
    if constexpr (KC) {
      const T rt = in_at(1);
      const T rst = rt * (-pw.inv_rho0s) + (T)1;
#pragma unroll
      for (int d = 0; d < 3; ++d) G[d] = g[d] * rst;
      ((T *)a.rhot)[c] = rt;
    }
    if constexpr (BF) {
      T v[3];
      velocity_at((const cplx<T> *)a.psi, c, pw, r, v);
      const T A = in_at(IA), Bx = in_at(IA + 1), By = in_at(IA + 2), Bz = in_at(IA + 3), Cc = in_at(IA + 4);
      T t = (v[2] * v[2] + v[1] * v[1] + v[0] * v[0]) * A;
      t += (T)-2 * v[0] * Bx;
      t += (T)-2 * v[1] * By;
      t += (T)-2 * v[2] * Bz;
      t = t + Cc;
      T re = t * ((T)-0.5 * pw.mass);
      if (pw.clamp_bf) re = clampv(re, pw.bfmax);
      vre += re * pw.cbf;
      const T Wx = v[0] * A - Bx, Wy = v[1] * A - By, Wz = v[2] * A - Bz;
      ((T *)a.wx)[c] = Wx;
      ((T *)a.wy)[c] = Wy;
      ((T *)a.wz)[c] = Wz;
      const T inv = (T)1 / (r + pw.div_eps);
      T im = (T)0.5 * (g[0] * inv) * Wx;
      im += (T)0.5 * (g[1] * inv) * Wy;
      im += (T)0.5 * (g[2] * inv) * Wz;
      ((T *)a.vim)[c] = im;
    }
    ((T *)a.vre)[c] = vre;
    // outputs replace inputs at this (pair, z, line) only: every input of this point was read above
    comp(L[(0 * TILE + pr) * PITCH + z], which) = Q;
    if constexpr (KC) {
      comp(L[(1 * TILE + pr) * PITCH + z], which) = G[0];
      comp(L[(2 * TILE + pr) * PITCH + z], which) = G[1];
      comp(L[(3 * TILE + pr) * PITCH + z], which) = G[2];
    }
  }
  __syncthreads();
  lines_fft<T, N, -1, 256>(L, s_tw, NOUT * TILE);
  store_split<T, N, TILE>(L, a.out, NOUT, p0, np, P);
}

// ------------------------------------------------------------------- zBC --
struct ZBCArgs {
  Ptrs8 in;   // J_x, J_y, J_z, T1
  Ptrs8 out;  // H
  const void *rho, *rhot, *wx, *wy, *wz;
  void *vre, *vim;
};

template <typename T, int N, bool BF>
__global__ void __launch_bounds__(256) k_zBC(ZBCArgs a, int64_t npairs, int P, PW<T> pw, const cplx<T> *__restrict__ g_tw) {
  constexpr int NIN = 4;
  using Cfg = ZCfg<T, N, NIN>;
  constexpr int TILE = Cfg::TILE, PITCH = Cfg::PITCH;
  __shared__ cplx<T> L[NIN * TILE * PITCH];
  __shared__ cplx<T> s_tw[2 * N];
  const int64_t p0 = (int64_t)blockIdx.x * TILE;
  const int np = (int)((npairs - p0) < TILE ? (npairs - p0) : TILE);
  fft::load_twiddles<T, N>(s_tw, g_tw);
  load_merge<T, N, TILE>(L, a.in, NIN, p0, np, P);
  __syncthreads();
  lines_fft<T, N, +1, 256>(L, s_tw, NIN * TILE);
  const T *rho = (const T *)a.rho, *rhot = (const T *)a.rhot;
  for (int e = threadIdx.x; e < np * 2 * N; e += 256) {
    const int pr = e / (2 * N), w = e - pr * 2 * N, which = w / N, z = w - which * N;
    const int64_t c = p0 * 2 * N + e;
    auto in_at = [&](int i) -> T { return comp(L[(i * TILE + pr) * PITCH + z], which); };
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    const T rst = rhot[c] * (-pw.inv_rho0s) + (T)1;
    T H = 0, t3 = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T g = (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h;
      const T Pd = (rhot[nb.p[d]] - rhot[nb.m[d]]) * pw.inv2h;  // (dF_d * rho) = D_d (F * rho)
      const T J = in_at(d);
      H += g * J;
      t3 += Pd * J;
    }
    ((T *)a.vre)[c] += in_at(3) * rst * pw.ckc - t3 * (pw.ckc * pw.k23);
    if constexpr (BF) {
      const T *W[3] = {(const T *)a.wx, (const T *)a.wy, (const T *)a.wz};
      T im = ((T *)a.vim)[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) im += (T)0.5 * ((W[d][nb.p[d]] - W[d][nb.m[d]]) * pw.inv2h);
      if (pw.clamp_bf) im = clampv(im, pw.bfmax);
      ((T *)a.vim)[c] = im * pw.cbf;
    }
    comp(L[(0 * TILE + pr) * PITCH + z], which) = H;
  }
  __syncthreads();
  lines_fft<T, N, -1, 256>(L, s_tw, TILE);
  store_split<T, N, TILE>(L, a.out, 1, p0, np, P);
}

// -------------------------------------------------------------------- zC --
struct ZCArgs {
  Ptrs8 in;              // summed linear terms (half spectrum)
  const void *vre, *vim;  // vim null without backflow
  void *V;               // API mode: potential accumulated here
  // step mode
  const void *vext;      // real, may be null
  double mu;
  const void *vavg;      // complex, may be null: V_eff = (vavg + V_eff)/2
  void *vsave;           // complex, may be null: stores V_eff before averaging
  void *phase;           // complex out: e^{-i V_eff dt/2}
  const void *psi_src;
  void *psi_dst;         // z forward transform of psi_src * phase
  double dtr, dti;
};

template <typename T, int N, bool STEP>
__global__ void __launch_bounds__(256) k_zC(ZCArgs a, int64_t npairs, int P, const cplx<T> *__restrict__ g_tw) {
  constexpr int NL = STEP ? 3 : 1;
  using Cfg = ZCfg<T, N, NL>;
  constexpr int TILE = Cfg::TILE, PITCH = Cfg::PITCH;
  __shared__ cplx<T> L[NL * TILE * PITCH];
  __shared__ cplx<T> s_tw[2 * N];
  const int64_t p0 = (int64_t)blockIdx.x * TILE;
  const int np = (int)((npairs - p0) < TILE ? (npairs - p0) : TILE);
  fft::load_twiddles<T, N>(s_tw, g_tw);
  load_merge<T, N, TILE>(L, a.in, 1, p0, np, P);
  __syncthreads();
  lines_fft<T, N, +1, 256>(L, s_tw, TILE);
  const T *vre = (const T *)a.vre, *vim = (const T *)a.vim;
  for (int e = threadIdx.x; e < np * 2 * N; e += 256) {
    const int pr = e / (2 * N), w = e - pr * 2 * N, which = w / N, z = w - which * N;
    const int64_t c = p0 * 2 * N + e;
    const T re = vre[c] + comp(L[pr * PITCH + z], which);
    const T im = vim ? vim[c] : (T)0;
    if constexpr (!STEP) {
      cplx<T> *V = (cplx<T> *)a.V;
      cplx<T> v = V[c];
      v.x += re;
      v.y += im;
      V[c] = v;
    } else {
      cplx<T> ve = cmk<cplx<T>>(re - (T)a.mu + (a.vext ? ((const T *)a.vext)[c] : (T)0), im);
      if (a.vsave) ((cplx<T> *)a.vsave)[c] = ve;
      if (a.vavg) {
        const cplx<T> va = ((const cplx<T> *)a.vavg)[c];
        ve = cmk<cplx<T>>((va.x + ve.x) * (T)0.5, (va.y + ve.y) * (T)0.5);
      }
      const cplx<T> ph = fft::vphase<T>(ve, (T)a.dtr, (T)a.dti);
      ((cplx<T> *)a.phase)[c] = ph;
      L[((1 + which) * TILE + pr) * PITCH + z] = cmul(((const cplx<T> *)a.psi_src)[c], ph);
    }
  }
  if constexpr (STEP) {
    __syncthreads();
    lines_fft<T, N, -1, 256>(L + TILE * PITCH, s_tw, 2 * TILE);
    cplx<T> *dst = (cplx<T> *)a.psi_dst + p0 * 2 * N;
    for (int e = threadIdx.x; e < np * 2 * N; e += 256) {
      const int pr = e / (2 * N), w = e - pr * 2 * N, which = w / N, z = w - which * N;
      dst[e] = L[((1 + which) * TILE + pr) * PITCH + z];
    }
  }
}

}  // namespace fused
}  // namespace dft
