// FFT pass kernels on the register-resident line transforms of rfft.hpp.
// Included by fft.hip (and, for the pair helpers, ot_fused.hpp).
#pragma once
#include "rfft.hpp"

namespace dft {
namespace fft {

using rf::ColCfg;
using rf::gld;
using rf::gld_once;
using rf::gst;
using rf::lde;
using rf::lde_once;
using rf::ste;
using rf::LineCfg;
using rf::Regs;
using rf::RP;

// Half spectrum of line `which` of a packed pair from Z = FFT(a + ib):
// A(k) = (Z(k) + Z*(N-k))/2, B(k) = (Z(k) - Z*(N-k))/2i.
template <typename C> __device__ __forceinline__ C split(C z, C zm, int which) {
  using T = decltype(z.x);
  return which == 0 ? cmk<C>((T)0.5 * (z.x + zm.x), (T)0.5 * (z.y - zm.y))
                    : cmk<C>((T)0.5 * (z.y + zm.y), (T)0.5 * (zm.x - z.x));
}

// Split the spectrum of a packed pair held in v (decimated layout of plan
// RP<N, EM>) and store the half spectra (k = 0..N/2) of rows 2p and 2p+1 of
// the half-spectrum field `out` (row length P; skipped when !act).  X is the
// pair's LDS exchange buffer.
template <typename T, int N, int EM, class SYNC>
__device__ __forceinline__ void split_store(const Regs<T, N, EM> &v, cplx<T> *X, int t, void *out,
                                            uint32_t p, int P, bool act) {
  using C = cplx<T>;
  constexpr int E = RP<N, EM>::E, TPL = RP<N, EM>::TPL;
#pragma unroll
  for (int i = 0; i < E; ++i) X[rf::LayOf<T, N, EM, SYNC>::pad(t + TPL * i)] = v[i];
  SYNC::sync();
  const uint32_t ea = p * 2u * (uint32_t)P + (uint32_t)t, eb = ea + (uint32_t)P;  // rows 2p, 2p+1 at point t
  // points k = t + TPL*i < N/2 for i < E/2; k = N/2 is point i = E/2 of thread 0
  C zm[E / 2 + 1];
#pragma unroll
  for (int i = 0; i <= E / 2; ++i) zm[i] = X[rf::LayOf<T, N, EM, SYNC>::pad((N - t - TPL * i) & (N - 1))];
  SYNC::sync();
  if (act) {
#pragma unroll
    for (int i = 0; i < E / 2; ++i) {
      ste<C>(out, ea, split(v[i], zm[i], 0), TPL * i);
      ste<C>(out, eb, split(v[i], zm[i], 1), TPL * i);
    }
    if (t == 0) {
      ste<C>(out, ea, split(v[E / 2], zm[E / 2], 0), N / 2);
      ste<C>(out, eb, split(v[E / 2], zm[E / 2], 1), N / 2);
    }
  }
}

// Packed pair Z(m) = A(m) + i B(m) from rows 2p, 2p+1 of a half-spectrum field
// (Hermitian extension for m > N/2), in two parts: load_raw issues the loads of
// both rows, merge_raw forms Z once they have arrived, so a kernel can issue
// the next stage's loads before the current stage's transform.
template <typename T, int N, int EM> struct RawPair {
  cplx<T> a[RP<N, EM>::E], b[RP<N, EM>::E];
};
template <typename T, int N, int EM>
__device__ __forceinline__ void load_raw(RawPair<T, N, EM> &r, const void *in, uint32_t p, int P, int t) {
  using C = cplx<T>;
  constexpr int E = RP<N, EM>::E, TPL = RP<N, EM>::TPL;
  // point m = t + TPL*i: row index m for i < E/2, N - m = (N - t) - TPL*i for i >= E/2
  const uint32_t ra = p * 2u * (uint32_t)P + rf::opaque_zero(), rb = ra + (uint32_t)P;
  const uint32_t fa = ra + (uint32_t)t, fb = rb + (uint32_t)t;
  const uint32_t ma = ra + (uint32_t)(N - t), mb = rb + (uint32_t)(N - t);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    if (i < E / 2) {
      r.a[i] = lde_once<C>(in, fa, TPL * i);
      r.b[i] = lde_once<C>(in, fb, TPL * i);
    } else {
      r.a[i] = lde_once<C>(in, ma, -TPL * i);
      r.b[i] = lde_once<C>(in, mb, -TPL * i);
    }
  }
}
template <typename T, int N, int EM>
__device__ __forceinline__ void merge_raw(Regs<T, N, EM> &v, const RawPair<T, N, EM> &r, int t) {
  using C = cplx<T>;
  constexpr int E = RP<N, EM>::E;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    C a = r.a[i], b = r.b[i];
    if (i > E / 2 || (i == E / 2 && t > 0)) {  // m > N/2: conjugate of the mirrored point
      a.y = -a.y;
      b.y = -b.y;
    }
    v[i] = cmk<C>(a.x - b.y, a.y + b.x);
  }
}
template <typename T, int N, int EM>
__device__ __forceinline__ void load_merge(Regs<T, N, EM> &v, const void *in, uint32_t p, int P, int t) {
  RawPair<T, N, EM> r;
  load_raw<T, N, EM>(r, in, p, P, t);
  merge_raw<T, N, EM>(v, r, t);
}

// ------------------------------------------------ z pass: real -> half spec --
// Pairs of adjacent real z-lines (2p, 2p+1) -> half spectra rows 2p, 2p+1.
template <typename T, int N>
__global__ void __launch_bounds__((LineCfg<T, N>::BS)) k_zr2c(Batch b, int64_t npairs, int P, const cplx<T> *__restrict__ g_tw) {
  using C = cplx<T>;
  using G = LineCfg<T, N>;
  constexpr int E = G::E, TPL = G::TPL, LP = G::LP;
  __shared__ C s_x[G::LINES * LP];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  const void *in = b.in[blockIdx.y];
  void *out = b.out[blockIdx.y];
  const int ln = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const uint32_t p0 = (uint32_t)blockIdx.x * G::LINES + ln;
  const bool act = p0 < npairs;
  const uint32_t p = act ? p0 : 0;  // out-of-range lanes read pair 0 and store nothing
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  const uint32_t e0 = p * 2u * N + (uint32_t)t;
  C v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = cmk<C>(lde_once<T>(in, e0, TPL * i), lde_once<T>(in, e0, N + TPL * i));
  __syncthreads();
  rf::rfft<T, N, G::EM, -1, rf::WaveSync>(v, s_x + ln * LP, s_tw, t);
  split_store<T, N, G::EM, rf::WaveSync>(v, s_x + ln * LP, t, out, p, P, act);
}

// ------------------------------------------------ z pass: half spec -> real --
template <typename T, int N>
__global__ void __launch_bounds__((LineCfg<T, N>::BS))
    k_zc2r(Batch b, int64_t npairs, int P, const cplx<T> *__restrict__ g_tw, T scale) {
  using C = cplx<T>;
  using G = LineCfg<T, N>;
  constexpr int E = G::E, TPL = G::TPL, LP = G::LP;
  __shared__ C s_x[G::LINES * LP];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  const void *in = b.in[blockIdx.y];
  void *out = b.out[blockIdx.y];
  const int ln = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const uint32_t p0 = (uint32_t)blockIdx.x * G::LINES + ln;
  const bool act = p0 < npairs;
  const uint32_t p = act ? p0 : 0;
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  C v[E];
  load_merge<T, N, G::EM>(v, in, p, P, t);
  __syncthreads();
  rf::rfft<T, N, G::EM, +1, rf::WaveSync>(v, s_x + ln * LP, s_tw, t);
  if (act) {
    const uint32_t e0 = p * 2u * N + (uint32_t)t;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      ste<T>(out, e0, scale * v[i].x, TPL * i);
      ste<T>(out, e0, scale * v[i].y, N + TPL * i);
    }
  }
}

// ------------------------------------------- strided c2c pass (y or x axis) --
// Geometry: DFT_STR_EM points per thread, at least DFT_STR_COLS columns per
// block (N = 512 keeps the column default).
#ifndef DFT_STR_EM
#define DFT_STR_EM 16
#endif
#ifndef DFT_STR_COLS
#define DFT_STR_COLS 8
#endif
template <typename T, int N>
using StrCfg = typename std::conditional<(N >= 512), ColCfg<T, N>, ColCfg<T, N, DFT_STR_EM, DFT_STR_COLS>>::type;

// Line (o, c): element e at o*os + e*es + c, c in [0, P).  blockIdx.x = o*ntile + tile.
template <typename T, int N, int DIR>
__global__ void __launch_bounds__((StrCfg<T, N>::BS))
    k_strided(Batch b, int P, int64_t es, int64_t os, int ntile, const cplx<T> *__restrict__ g_tw, T scale) {
  using C = cplx<T>;
  using G = StrCfg<T, N>;
  constexpr int E = G::E, TPL = G::TPL, COLS = G::COLS, LP = G::LP;
  __shared__ C s_x[G::S > 1 ? COLS * LP : 1];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  const void *in = b.in[blockIdx.y];
  void *out = b.out[blockIdx.y];
  const int o = blockIdx.x / ntile, c0 = (blockIdx.x - o * ntile) * COLS;
  const int col = threadIdx.x % COLS, t = threadIdx.x / COLS;
  const bool act = c0 + col < P;
  const int cc = act ? c0 + col : P - 1;  // out-of-range lanes read the last column and store nothing
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  const uint32_t ob = (uint32_t)((o * os + cc + t * es) * (int64_t)sizeof(C));
  const uint32_t st = (uint32_t)(TPL * es * (int64_t)sizeof(C));
  C v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = gld<C>((const char *)in + (size_t)i * st, ob);
  __syncthreads();
  rf::rfft<T, N, G::EM, DIR, rf::BlockSync>(v, s_x + col * LP, s_tw, t);
  if (act) {
#pragma unroll
    for (int i = 0; i < E; ++i) gst<C>((char *)out + (size_t)i * st, ob, cscale(v[i], scale));
  }
}

// ------------------------------- inverse x pass with fused k-space combine --
// Spectra are [nx][ny][PP] (P values per row); x line (ky, kz) at ky*PP + kx*ny*PP + kz.
template <typename T, int N>
__global__ void __launch_bounds__((ColCfg<T, N>::BS))
    k_xcombine_inv(Batch b, CombineSet cs, KernTab kt, int ny, int P, int PP, int ntile,
                   const cplx<T> *__restrict__ g_tw) {
  using C = cplx<T>;
  using G = ColCfg<T, N>;
  constexpr int E = G::E, TPL = G::TPL, COLS = G::COLS, LP = G::LP;
  __shared__ C s_x[G::S > 1 ? COLS * LP : 1];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  const Combine cb = cs.c[blockIdx.y];
  void *out = b.out[blockIdx.y];
  const int o = blockIdx.x / ntile, c0 = (blockIdx.x - o * ntile) * COLS;
  const int col = threadIdx.x % COLS, t = threadIdx.x / COLS;
  const bool act = c0 + col < P;
  const int cc = act ? c0 + col : P - 1;
  const uint32_t es = (uint32_t)ny * PP;
  const uint32_t e0 = (uint32_t)o * PP + cc + t * es;  // element index of point t
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  const T *sx = (const T *)kt.sinx, *sy = (const T *)kt.siny, *sz = (const T *)kt.sinz;
  C v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = cmk<C>(0, 0);
  const size_t stC = (size_t)TPL * es * sizeof(C), stR = (size_t)TPL * es * sizeof(T);
  for (int q = 0; q < cb.nterms; ++q) {
    const Term tm = cb.t[q];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int el = t + TPL * i;
      C a = gld<C>((const char *)b.in[tm.in] + i * stC, e0 * (uint32_t)sizeof(C));
      T s = (T)tm.scale;
      if (tm.kern >= 0) s *= gld<T>((const char *)kt.k[tm.kern] + i * stR, e0 * (uint32_t)sizeof(T));
      if (tm.deriv >= 0) {
        s *= tm.deriv == 0 ? sx[el] : (tm.deriv == 1 ? sy[o] : sz[cc]);
        a = cmk<C>(-a.y, a.x);  // times i
      }
      v[i].x += s * a.x;
      v[i].y += s * a.y;
    }
  }
  __syncthreads();
  rf::rfft<T, N, G::EM, +1, rf::BlockSync>(v, s_x + col * LP, s_tw, t);
  if (act) {
#pragma unroll
    for (int i = 0; i < E; ++i) gst<C>((char *)out + i * stC, e0 * (uint32_t)sizeof(C), v[i]);
  }
}

// ---------------------------------------- propagator passes (complex psi) --
// Sum of an FP64 value over the block (BS threads) into partial[blockIdx.x].
template <int BS> __device__ __forceinline__ void block_sum(double acc, double *partial) {
  static_assert(BS == 64 || BS == 128 || BS == 256, "block of 1, 2 or 4 wavefronts");
  __shared__ double s_red[BS / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = s_red[0];
    if constexpr (BS == 128) t += s_red[1];
    if constexpr (BS == 256) t = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
    partial[blockIdx.x] = t;
  }
}

#ifndef DFT_ZINV_PF
#define DFT_ZINV_PF 1
#endif
#ifndef DFT_ZINV_PF64
#define DFT_ZINV_PF64 1
#endif
// z pass over a full complex field, multiplying by e^{-iV dt/2} (V = pot,
// complex) before (DIR<0) or after (DIR>0, together with `scale`) the
// transform.  After the inverse pass each block writes sum |psi|^2 to
// partial[block] (FP64) when `partial` is given.
template <typename T, int N, int DIR>
__global__ void __launch_bounds__((LineCfg<T, N>::BS))
    k_zc2c_pot(const cplx<T> *src, cplx<T> *dst, const cplx<T> *__restrict__ pot, int64_t nlines,
               const cplx<T> *__restrict__ g_tw, double dtr, double dti, T scale, double *__restrict__ partial) {
  using C = cplx<T>;
  using G = LineCfg<T, N>;
  constexpr int E = G::E, TPL = G::TPL, LP = G::LP;
  __shared__ C s_x[G::S > 1 ? G::LINES * LP : 1];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  const int ln = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const uint32_t l0 = (uint32_t)blockIdx.x * G::LINES + ln;
  const bool act = l0 < nlines;
  const uint32_t l = act ? l0 : 0;
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  const uint32_t e0 = l * (uint32_t)N + (uint32_t)t;
  C v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = lde_once<C>(src, e0, TPL * i);
  if (DIR < 0 && pot) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      v[i] = apply_vphase<T>(v[i], lde<C>(pot, e0, TPL * i), dtr, dti);
    }
  }
  // inverse pass: the potential is read before the transform (DFT_ZINV_PF,
  // DFT_ZINV_PF64), so that its loads overlap the field's
  constexpr bool PF = DIR > 0 && (sizeof(T) == 4 ? DFT_ZINV_PF : DFT_ZINV_PF64);
  C pv[PF ? E : 1];
  if constexpr (PF) {
    if (pot)
#pragma unroll
      for (int i = 0; i < E; ++i) pv[i] = lde<C>(pot, e0, TPL * i);
  }
  __syncthreads();
  rf::rfft<T, N, G::EM, DIR, rf::WaveSync>(v, s_x + ln * LP, s_tw, t);
  double acc = 0.0;
  if (DIR > 0) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      v[i] = cscale(v[i], scale);
      if (pot) v[i] = apply_vphase<T>(v[i], PF ? pv[PF ? i : 0] : lde<C>(pot, e0, TPL * i), dtr, dti);
      acc += (double)v[i].x * (double)v[i].x + (double)v[i].y * (double)v[i].y;
    }
    if (!act) acc = 0.0;
  }
  if (act) {
#pragma unroll
    for (int i = 0; i < E; ++i) ste<C>(dst, e0, v[i], TPL * i);
  }
  if (DIR > 0 && partial) block_sum<G::BS>(acc, partial);
}

// Total of the per-block partial sums partial[0 .. nb), in every thread of a
// 256-thread block.
__device__ __forceinline__ double block_total_256(const double *partial, int nb) {
  __shared__ double s_tot[4];
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) t += partial[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_tot[threadIdx.x >> 6] = t;
  __syncthreads();
  return (s_tot[0] + s_tot[1]) + (s_tot[2] + s_tot[3]);
}

// psi * f for a real FP64 factor f: in FP32 applied as its rounded value plus
// its residual.
template <typename T> __device__ __forceinline__ cplx<T> scale_by(cplx<T> v, double f) {
  if constexpr (sizeof(T) == 4) {
    const float h = (float)f, l = (float)(f - (double)h);
    return cmk<cplx<T>>(__builtin_fmaf(v.x, h, v.x * l), __builtin_fmaf(v.y, h, v.y * l));
  } else {
    return cmk<cplx<T>>(v.x * f, v.y * f);
  }
}

// Renormalisation fused into a step's closing z pass (grid_wf_normalize): after
// k_zc2c_pot<+1> wrote its per-block sums of |psi|^2, one block turns them into
// psi's factor sqrt(target / (h3 sum)), stored after them, and k_scale_psi
// applies it; no host round trip.
static __global__ void k_norm_factor(double *partial, int nb, double h3, double target) {
  const double tot = block_total_256(partial, nb) * h3;
  if (threadIdx.x == 0) partial[nb] = tot > 0.0 ? sqrt(target / tot) : 1.0;
}
template <typename T> __global__ void k_scale_psi(cplx<T> *psi, const double *f, int64_t n) {
  const double s = *f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    psi[i] = scale_by<T>(psi[i], s);
}

// Fused x pass of the propagator: forward x FFT, kinetic phase
// exp(-i k^2 dt / 2M), inverse x FFT.  Field is [nx][ny][nz] complex.
// The phase factorises as X(kx) YZ(ky, kz); both factors are evaluated in FP64
// (X once per block, in LDS), multiplied in FP64, and in FP32 applied as the
// rounded value plus its residual, so that its modulus is exact: a phase of
// the same mode is applied every step, and a rounded one would change that
// mode's amplitude by the same factor every time.
// Geometry: at least DFT_XKIN_COLS columns per block (N = 512 keeps the
// column default).
#ifndef DFT_XKIN_COLS
#define DFT_XKIN_COLS 8
#endif
// waves per SIMD asked of the register allocator (0: no request): at 3 the
// FP32 pass fits 168 VGPRs with 28 bytes per lane of spills and runs 70 -> 58 us
// at 256^3 (178 VGPRs and 2 waves without); more waves do not fit its LDS.
#ifndef DFT_XKIN_MINB
#define DFT_XKIN_MINB 3
#endif
// the same request for the FP64 pass (0: none)
#ifndef DFT_XKIN_MINB64
#define DFT_XKIN_MINB64 0
#endif
template <typename T, int N>
using KinCfg = typename std::conditional<(N >= 512), ColCfg<T, N>, ColCfg<T, N, 16, DFT_XKIN_COLS>>::type;

template <typename T, int N>
__global__ void __launch_bounds__((KinCfg<T, N>::BS), sizeof(T) == 4 ? DFT_XKIN_MINB : DFT_XKIN_MINB64)
    k_xkin(cplx<T> *psi, int ny, int nz, int ntile, const cplx<T> *__restrict__ g_tw, double cx, double cy, double cz,
           double dtr, double dti, double ox, double oy, double oz) {
  using C = cplx<T>;
  using G = KinCfg<T, N>;
  constexpr int E = G::E, TPL = G::TPL, COLS = G::COLS, LP = G::LP;
  __shared__ C s_x[G::S > 1 ? COLS * LP : 1];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  __shared__ double2 s_kx[N];
  const int o = blockIdx.x / ntile, c0 = (blockIdx.x - o * ntile) * COLS;
  const int col = threadIdx.x % COLS, t = threadIdx.x / COLS;
  const bool act = c0 + col < nz;
  const int cc = act ? c0 + col : nz - 1;
  const uint32_t es = (uint32_t)ny * nz;
  const uint32_t ob = ((uint32_t)o * nz + cc + t * es) * (uint32_t)sizeof(C);
  const uint32_t st = (uint32_t)TPL * es * (uint32_t)sizeof(C);
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  C v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = gld<C>((const char *)psi + (size_t)i * st, ob);
  // exp(-i K dt) = exp(K dti) (cos(K dtr) - i sin(K dtr)) for the x part
  for (int e = threadIdx.x; e < N; e += G::BS) {
    const double mx = (double)(e < N / 2 ? e : e - N) - ox;
    const double K = cx * (mx * mx);
    double s, c;
    sincos(K * dtr, &s, &c);
    const double m = exp(K * dti);
    s_kx[e] = make_double2(m * c, -m * s);
  }
  __syncthreads();
  rf::rfft<T, N, G::EM, -1, rf::BlockSync>(v, s_x + col * LP, s_tw, t);
  const double my = (double)(o < ny / 2 ? o : o - ny) - oy;
  const double mz = (double)(cc < nz / 2 ? cc : cc - nz) - oz;
  const double Kyz = cy * (my * my) + cz * (mz * mz);
  double sy, cyz;
  sincos(Kyz * dtr, &sy, &cyz);
  const double myz = exp(Kyz * dti);
  const double yr = myz * cyz, yi = -myz * sy;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const double2 x = s_kx[t + TPL * i];
    const double pr = x.x * yr - x.y * yi, pi = x.x * yi + x.y * yr;
    if constexpr (sizeof(T) == 4) {
      const float hr = (float)pr, hi = (float)pi;
      v[i] = cmul_hl_phase(v[i], make_float2(hr, hi), make_float2((float)(pr - hr), (float)(pi - hi)));
    } else {
      v[i] = cmul(v[i], cmk<C>(pr, pi));
    }
  }
  rf::rfft<T, N, G::EM, +1, rf::BlockSync>(v, s_x + col * LP, s_tw, t);
  if (act) {
#pragma unroll
    for (int i = 0; i < E; ++i) gst<C>((char *)psi + (size_t)i * st, ob, v[i]);
  }
}

}  // namespace fft
}  // namespace dft
