// Fused stage kernels of the OT potential and the split-operator step, on the
// register-resident line transforms of rfft.hpp.
//
// Every z pass works on pairs of adjacent z lines (x, 2p) and (x, 2p+1),
// contiguous in memory and packed as one complex line a + ib.  TPL lanes of
// one wavefront hold a pair, E points each (z = t + TPL*i), so the pointwise
// stages run in registers as the prologue or epilogue of the z transforms,
// field by field: a stage that consumes several spectra transforms them one
// after another and accumulates; a stage that produces several transforms and
// stores each as soon as it is formed.  Stencils in x and y read neighbouring
// lines from global memory (L2/MALL resident: adjacent planes are processed
// together).
//
//   zA    psi -> rho (stored) and, with backflow, q_d = rho v_d, q4 = rho |v|^2;
//         z r2c of {rho, q_x, q_y, q_z, q4} (and of rho~ = G rho, with rho~ in
//         place of rho in q_d and q4, under the HD / HD2 backflow switch)
//   xmid  x forward FFT, k-space products with real kernel spectra and the
//         central-difference symbol i sin(k h)/h, x inverse FFT (a small
//         program: several outputs per input, one multi-term output)
//   zAB   z c2r of {U*rho, U*q_d, U*q4} -> BF real part, W_d, BF imaginary part;
//         z c2r of Pi*rho -> local terms, Q, z r2c of Q;
//         z c2r of F*rho -> rho_st, G_d, z r2c of G_d
//   zBC   z c2r of {J_d, T1} -> H, KC terms 1 and 3, BF divergence; z r2c of H
//   zC    z c2r of the summed linear terms; V += (API), or phase = e^{-iV dt/2}
//         and the first z pass of the propagator (step)
#pragma once
#include "fft_kernels.hpp"
#include "ot_common.hpp"

namespace dft {
namespace fused {

using rf::ColCfg;
using rf::LineCfg;
using rf::RP;

struct Ptrs8 {
  void *p[8];
};

// Addresses of a pair's points and of their periodic neighbours.  Line a is
// (ix, j), line b is (ix, j + 1); thread t holds points z = t + TPL*i of both.
// Every access is (element index of point t of some line) + a constant delta,
// so all points of a thread share a handful of per-thread offsets.
template <typename T, int N, int TPL> struct PairPos {
  uint32_t a0;         // point t of line a (line b is a0 + N)
  uint32_t xp0, xm0;   // point t of lines (ix +- 1, j)
  uint32_t ym0, yp1;   // point t of lines (ix, j - 1) and (ix, j + 2)
  int t;
  PairPos() = default;
  __device__ PairPos(uint32_t p, int t_, const PW<T> &pw) : t(t_) {
    const uint32_t l = 2 * p;
    const int j = (int)(l & (uint32_t)(pw.ny - 1)), ix = (int)(l >> __builtin_ctz(pw.ny));
    a0 = l * N + t;
    xp0 = ((uint32_t)((ix + 1) & (pw.nx - 1)) * pw.ny + j) * N + t;
    xm0 = ((uint32_t)((ix - 1) & (pw.nx - 1)) * pw.ny + j) * N + t;
    ym0 = ((uint32_t)ix * pw.ny + ((j - 1) & (pw.ny - 1))) * N + t;
    yp1 = ((uint32_t)ix * pw.ny + ((j + 2) & (pw.ny - 1))) * N + t;
  }
  static constexpr int d(int w, int i) { return w * N + TPL * i; }
  // The same positions with offsets that the loads of the current stage
  // depend on (see rf::opaque_zero).
  __device__ PairPos here() const {
    PairPos r = *this;
    const uint32_t z = rf::opaque_zero();
    r.a0 += z;
    r.xp0 += z;
    r.xm0 += z;
    r.ym0 += z;
    r.yp1 += z;
    return r;
  }
  // field value at point (w, i)
  template <typename X> __device__ X at(const void *f, int w, int i) const { return rf::lde<X>(f, a0, d(w, i)); }
  // the same for a field this kernel reads only here
  template <typename X> __device__ X at_once(const void *f, int w, int i) const {
    return rf::lde_once<X>(f, a0, d(w, i));
  }
  template <typename X> __device__ void put(void *f, int w, int i, X v) const { rf::ste<X>(f, a0, v, d(w, i)); }
  // neighbour of point (w, i) one step along axis D in direction S (+1 / -1)
  template <typename X, int D, int S> __device__ X nb(const void *f, int w, int i) const {
    constexpr int E = N / TPL;
    if constexpr (D == 0) {
      return rf::lde<X>(f, S > 0 ? xp0 : xm0, d(w, i));
    } else if constexpr (D == 1) {
      if (S > 0) return w ? rf::lde<X>(f, yp1, TPL * i) : rf::lde<X>(f, a0, d(1, i));
      return w ? rf::lde<X>(f, a0, d(0, i)) : rf::lde<X>(f, ym0, TPL * i);
    } else {
      // z +- 1; only the last point of the line wraps forward, the first backward
      if (S > 0) {
        if (i < E - 1) return rf::lde<X>(f, a0, d(w, i) + 1);
        return rf::lde<X>(f, t == TPL - 1 ? a0 - t : a0 + TPL * i + 1, w * N);
      }
      if (i > 0) return rf::lde<X>(f, a0, d(w, i) - 1);
      return rf::lde<X>(f, t == 0 ? a0 + (N - 1) : a0 - 1, w * N);
    }
  }
};

// v_d = Im(psi* (psi_{+d} - psi_{-d})/2h) / (M (rho + eps)), clamped when
// enabled; f = 1 / (M (rho + eps)) is computed once per point for all axes.
template <typename T> __device__ __forceinline__ T velocity_factor(const PW<T> &pw, T r) {
  return pw.invM / (r + pw.eps_v);
}
// the same from psi at the point (s) and at its two neighbours (a: +d, b: -d)
template <typename T>
__device__ __forceinline__ T velocity_sab(const PW<T> &pw, T f, cplx<T> s, cplx<T> a, cplx<T> b) {
  const T dx = (a.x - b.x) * pw.inv2h, dy = (a.y - b.y) * pw.inv2h;
  T v = (s.x * dy - s.y * dx) * f;
  if (pw.clamp_v) v = clampv(v, pw.vmax);
  return v;
}
template <typename T, int D, class Q>
__device__ __forceinline__ T velocity_d(const void *psi, const Q &q, int w, int i, const PW<T> &pw, T f,
                                        cplx<T> s) {
  return velocity_sab(pw, f, s, q.template nb<cplx<T>, D, 1>(psi, w, i), q.template nb<cplx<T>, D, -1>(psi, w, i));
}
template <typename T, int D, class Q>
__device__ __forceinline__ T velocity_d(const void *psi, const Q &q, int w, int i, const PW<T> &pw, T f) {
  return velocity_d<T, D>(psi, q, w, i, pw, f, q.template at<cplx<T>>(psi, w, i));
}

template <typename T, int D, class Q>
__device__ __forceinline__ T grad_d(const void *f, const Q &q, int w, int i, const PW<T> &pw) {
  return (q.template nb<T, D, 1>(f, w, i) - q.template nb<T, D, -1>(f, w, i)) * pw.inv2h;
}
// central difference of rho~ = G(rho) rho (backflow density switch)
template <typename T, int D, class Q>
__device__ __forceinline__ T grad_rt_d(const void *rho, const Q &q, int w, int i, const PW<T> &pw) {
  return (bf_rt(q.template nb<T, D, 1>(rho, w, i), pw) - bf_rt(q.template nb<T, D, -1>(rho, w, i), pw)) * pw.inv2h;
}

// Pair kernels use the plan with the fewest points per thread, 4: a line is
// one wavefront up to N = 256 and two at N = 512 (exchanges then take the
// block barrier).
template <int N> constexpr int pair_em() { return 4; }
template <int N>
using PairSync = std::conditional_t<(rf::RP<N, pair_em<N>()>::TPL <= 64), rf::WaveSync, rf::LineBlockSync>;

// Threads per block of the pair kernels.
template <typename T, int N> constexpr int PairBS = LineCfg<T, N, pair_em<N>()>::BS;

// Common prologue of the pair kernels.
#define FUSED_PAIR_SETUP()                                                   \
  using C = cplx<T>;                                                         \
  using G = LineCfg<T, N, pair_em<N>()>;                                     \
  constexpr int E = G::E, TPL = G::TPL, LP = G::LP;                          \
  __shared__ C s_x[G::LINES * LP];                                           \
  __shared__ C s_tw[rf::tw_len<T, N>()];                                     \
  const int ln = threadIdx.x / TPL, t = threadIdx.x % TPL;                   \
  const uint32_t p0 = (uint32_t)blockIdx.x * G::LINES + ln;                  \
  const bool act = p0 < npairs;                                              \
  /* out-of-range lanes compute on pair 0 and store nothing */               \
  const uint32_t p = act ? p0 : 0;                                           \
  C *X = s_x + ln * LP;                                                      \
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);                                 \
  __syncthreads();                                                           \
  const PairPos<T, N, TPL> q(p, t, pw)

// r2c of the pair held in v and store of both half spectra to field `out`.
template <typename T, int N>
__device__ __forceinline__ void r2c_store(rf::Regs<T, N, pair_em<N>()> &v, cplx<T> *X, const cplx<T> *tw, int t,
                                          void *out, uint32_t p, int P, bool act) {
  rf::rfft<T, N, pair_em<N>(), -1, PairSync<N>>(v, X, tw, t);
  fft::split_store<T, N, pair_em<N>(), PairSync<N>>(v, X, t, out, p, P, act);
}

// Load + c2r of both lines of the pair from half-spectrum field `in`.
template <typename T, int N>
__device__ __forceinline__ void load_c2r(rf::Regs<T, N, pair_em<N>()> &v, cplx<T> *X, const cplx<T> *tw, int t,
                                         const void *in, uint32_t p, int P) {
  fft::load_merge<T, N, pair_em<N>()>(v, in, p, P, t);
  rf::rfft<T, N, pair_em<N>(), +1, PairSync<N>>(v, X, tw, t);
}

template <typename C> __device__ __forceinline__ auto &part(C &c, int w) { return w ? c.y : c.x; }

// -------------------------------------------------------------------- zA --
// waves per SIMD asked of the register allocator for the FP32 zA and zC
// kernels (0: no request)
#ifndef DFT_ZA_MINB
#define DFT_ZA_MINB 5
#endif
#ifndef DFT_ZC_MINB
#define DFT_ZC_MINB 0
#endif
// FP32 zA with backflow keeps psi of its points in registers for the three
// velocity axes and q4 instead of reading psi and rho again (zA 158 -> 136 us at
// 256^3, 5 waves per SIMD either way)
#ifndef DFT_ZA_KEEP
#define DFT_ZA_KEEP 1
#endif
// ... and takes the z neighbours of those points from the line's other lanes
// (zA 135 -> 122 us at 256^3, bit-identical; at 4 waves per SIMD, where it
// needs no spills)
#ifndef DFT_ZA_ZSHFL
#define DFT_ZA_ZSHFL 1
#endif
#ifndef DFT_ZA_MINB_SH
#define DFT_ZA_MINB_SH 4
#endif
// the same for FP64 (132 -> 192 VGPRs, 3 -> 2 waves per SIMD; zA 257 -> 243 us
// at 256^3)
#ifndef DFT_ZA_KEEP64
#define DFT_ZA_KEEP64 1
#endif
template <typename T, int N> constexpr int za_minb() {
  return sizeof(T) != 4 ? 0 : (DFT_ZA_KEEP && DFT_ZA_ZSHFL && N <= 256 ? DFT_ZA_MINB_SH : DFT_ZA_MINB);
}
// HB: the backflow density switch of HD / HD2 (q_d = rho~ v_d, q4 = rho~ |v|^2
// and a sixth output, the transform of rho~ = G rho)
template <typename T, int N, bool BF, bool HB>
__device__ __forceinline__ void zA_body(const cplx<T> *psi, T *rho, Ptrs8 outs, int64_t npairs, int P, PW<T> pw,
                                        const cplx<T> *__restrict__ g_tw) {
  FUSED_PAIR_SETUP();
  // with backflow, psi of the points is kept for the three axes (KEEP: up to
  // N = 256; at 512 FP32 it would spill) or read again
  constexpr bool KEEP = BF && DFT_ZA_KEEP && (sizeof(T) == 4 || DFT_ZA_KEEP64) && N <= 256;
  C v[E], sp[2][KEEP ? E : 1];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const C sa = q.template at<C>(psi, 0, i), sb = q.template at<C>(psi, 1, i);
    if constexpr (KEEP) {
      sp[0][i] = sa;
      sp[1][i] = sb;
    }
    v[i] = cmk<C>(sa.x * sa.x + sa.y * sa.y, sb.x * sb.x + sb.y * sb.y);
  }
  if (act) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      q.template put<T>(rho, 0, i, v[i].x);
      q.template put<T>(rho, 1, i, v[i].y);
    }
  }
  if constexpr (BF) {
    // velocity factor per point, kept for the three axes
    T fv[2][E];
    T rt[2][HB ? E : 1];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      fv[0][i] = velocity_factor(pw, v[i].x);
      fv[1][i] = velocity_factor(pw, v[i].y);
      if constexpr (HB) {
        rt[0][i] = bf_rt(v[i].x, pw);
        rt[1][i] = bf_rt(v[i].y, pw);
      }
    }
    r2c_store<T, N>(v, X, s_tw, t, outs.p[0], p, P, act);
    T v2[2][E];
#pragma unroll
    for (int i = 0; i < E; ++i) v2[0][i] = v2[1][i] = 0;
    auto qd = [&](auto dtag) {
      constexpr int D = decltype(dtag)::value;
      const auto qq = q.here();
      // z neighbours of the kept points from the line's other lanes: point
      // t + TPL i has z + 1 at lane t + 1 (index i), or at lane 0 (index i + 1)
      // for t = TPL - 1; z - 1 likewise (wrapping around the line)
      constexpr bool ZSH = KEEP && D == 2 && DFT_ZA_ZSHFL && TPL <= 64;
      C zp[2][ZSH ? E : 1], zm[2][ZSH ? E : 1];
      if constexpr (ZSH) {
        const int up = (t + 1) & (TPL - 1), dn = (t + TPL - 1) & (TPL - 1);
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          C su[E], sd[E];
#pragma unroll
          for (int i = 0; i < E; ++i) {
            su[i] = cmk<C>(__shfl(sp[w][i].x, up, TPL), __shfl(sp[w][i].y, up, TPL));
            sd[i] = cmk<C>(__shfl(sp[w][i].x, dn, TPL), __shfl(sp[w][i].y, dn, TPL));
          }
#pragma unroll
          for (int i = 0; i < E; ++i) {
            zp[w][i] = t == TPL - 1 ? su[(i + 1) % E] : su[i];
            zm[w][i] = t == 0 ? sd[(i + E - 1) % E] : sd[i];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < E; ++i) {
        T o[2];
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          C s;
          if constexpr (KEEP) s = sp[w][i];
          else s = qq.template at<C>(psi, w, i);
          T r = s.x * s.x + s.y * s.y;
          if constexpr (HB) r = rt[w][i];
          T vd;
          if constexpr (ZSH) vd = velocity_sab(pw, fv[w][i], s, zp[w][i], zm[w][i]);
          else vd = velocity_d<T, D>(psi, qq, w, i, pw, fv[w][i], s);
          v2[w][i] += vd * vd;
          o[w] = vd * r;
        }
        v[i] = cmk<C>(o[0], o[1]);
      }
      r2c_store<T, N>(v, X, s_tw, t, outs.p[1 + D], p, P, act);
    };
    qd(std::integral_constant<int, 0>{});
    qd(std::integral_constant<int, 1>{});
    qd(std::integral_constant<int, 2>{});
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if constexpr (HB) v[i] = cmk<C>(v2[0][i] * rt[0][i], v2[1][i] * rt[1][i]);
      else if constexpr (KEEP)
        v[i] = cmk<C>(v2[0][i] * (sp[0][i].x * sp[0][i].x + sp[0][i].y * sp[0][i].y),
                      v2[1][i] * (sp[1][i].x * sp[1][i].x + sp[1][i].y * sp[1][i].y));
      else v[i] = cmk<C>(v2[0][i] * q.template at<T>(rho, 0, i), v2[1][i] * q.template at<T>(rho, 1, i));
    }
    r2c_store<T, N>(v, X, s_tw, t, outs.p[4], p, P, act);
    if constexpr (HB) {
#pragma unroll
      for (int i = 0; i < E; ++i) v[i] = cmk<C>(rt[0][i], rt[1][i]);
      r2c_store<T, N>(v, X, s_tw, t, outs.p[5], p, P, act);
    }
  } else {
    r2c_store<T, N>(v, X, s_tw, t, outs.p[0], p, P, act);
  }
}
template <typename T, int N, bool BF>
__global__ void __launch_bounds__((PairBS<T, N>), (za_minb<T, N>()))
    k_zA(const cplx<T> *psi, T *rho, Ptrs8 outs, int64_t npairs, int P, PW<T> pw,
         const cplx<T> *__restrict__ g_tw) {
  zA_body<T, N, BF, false>(psi, rho, outs, npairs, P, pw, g_tw);
}
template <typename T, int N>
__global__ void __launch_bounds__((PairBS<T, N>), 0)
    k_zA_bfhd(const cplx<T> *psi, T *rho, Ptrs8 outs, int64_t npairs, int P, PW<T> pw,
              const cplx<T> *__restrict__ g_tw) {
  zA_body<T, N, true, true>(psi, rho, outs, npairs, P, pw, g_tw);
}

// ------------------------------------------------------------------ xmid --
struct MidTerm {
  int in, kern, deriv, pad;  // kern: slot of MidProg::kern (or -1); deriv: axis of i sin(k h)/h (or -1)
  double scale;
};
struct MidOut {
  int nterms, pad;
  MidTerm t[3];
  void *out;
};
constexpr int kMidMaxIn = 6, kMidMaxOut = 8, kMidMaxKern = 3;
struct MidProg {
  int nin, nout, multi, nkern;
  const void *in[kMidMaxIn];
  MidOut o[kMidMaxOut];
  const void *kern[kMidMaxKern];  // the distinct kernel spectra the terms use
  const void *sinx, *siny, *sinz;
};

// x-middle geometry: DFT_XMID_EM points per thread (several spectra are live
// at once), at least DFT_XMID_COLS columns per block.
#ifndef DFT_XMID_EM
#define DFT_XMID_EM 8
#endif
#ifndef DFT_XMID_COLS
#define DFT_XMID_COLS 8
#endif
template <typename T, int N> using MidCfg = ColCfg<T, N, DFT_XMID_EM, DFT_XMID_COLS>;
// minimum resident blocks per CU asked of the register allocator (0: none)
#ifndef DFT_XMID_ORDER
#define DFT_XMID_ORDER 0
#endif
#ifndef DFT_XMID_MINB
#define DFT_XMID_MINB 0
#endif
template <typename T> constexpr int mid_minb() { return sizeof(T) == 4 ? DFT_XMID_MINB : 0; }

// Blocks of the x-middle pass: ny * ntile tiles of COLS consecutive kz < P-1
// at one ky, then ceil(ny / COLS) tiles of the Nyquist column kz = P-1 over
// COLS consecutive ky.
template <int COLS> __host__ __device__ inline int mid_tiles(int P) { return (P - 1 + COLS - 1) / COLS; }
template <int COLS> __host__ __device__ inline int mid_blocks(int ny, int P) {
  return ny * mid_tiles<COLS>(P) + (ny + COLS - 1) / COLS;
}

// Spectra are [nx][ny][PP] (P values per row) after the y pass; the x line
// (ky = o, kz) has stride ny*PP.  Each distinct kernel spectrum is read once
// per point into registers (NK of them) and shared by the terms that use it.
// The program is read from device memory: indexed at run time, a by-value
// kernel argument of this size would be copied to scratch by every thread.
template <typename T, int N, int NK, bool DERIV>
__global__ void __launch_bounds__((MidCfg<T, N>::BS), mid_minb<T>()) k_xmid(const MidProg *__restrict__ pgp, int ny, int P, int PP,
                                                           const cplx<T> *__restrict__ g_tw) {
  const MidProg &pg = *pgp;
  using C = cplx<T>;
  using G = MidCfg<T, N>;
  constexpr int E = G::E, TPL = G::TPL, COLS = G::COLS, LP = G::LP;
  __shared__ C s_x[G::S > 1 ? COLS * LP : 1];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  // DFT_XMID_ORDER 0: consecutive tiles (one ky) on one XCD, back to back
  // (shared L2 lines); 1: ky fastest, so that the blocks in flight spread over
  // whole x-planes
  const int b = DFT_XMID_ORDER == 0 ? rf::xcd_block() : (int)blockIdx.x;
  const int col = threadIdx.x % COLS, t = threadIdx.x / COLS;
  const int ntile = mid_tiles<COLS>(P);
  int o, kz;
  bool act;
  if (b < ny * ntile) {
    const int tile = DFT_XMID_ORDER == 0 ? b % ntile : b / ny;
    o = DFT_XMID_ORDER == 0 ? b / ntile : b % ny;
    const int c = tile * COLS + col;
    act = c < P - 1;
    kz = act ? c : P - 2;  // out-of-range lanes read an in-range column and store nothing
  } else {
    const int r = (b - ny * ntile) * COLS + col;
    act = r < ny;
    o = act ? r : ny - 1;
    kz = P - 1;
  }
  const uint32_t es = (uint32_t)ny * PP;
  const uint32_t e0 = (uint32_t)o * PP + kz + t * es;  // element index of point t
  const uint32_t obC = e0 * (uint32_t)sizeof(C), obR = e0 * (uint32_t)sizeof(T);
  const size_t stC = (size_t)TPL * es * sizeof(C), stR = (size_t)TPL * es * sizeof(T);  // per point i
  C *X = s_x + col * LP;
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  // the first input and the kernels are in flight while the twiddles land
  C nxt[E];
#pragma unroll
  for (int i = 0; i < E; ++i) nxt[i] = rf::gld<C>((const char *)pg.in[0] + i * stC, obC);
  T kv[NK][E];
#pragma unroll
  for (int k = 0; k < NK; ++k)
#pragma unroll
    for (int i = 0; i < E; ++i) kv[k][i] = rf::gld<T>((const char *)pg.kern[k] + i * stR, obR);
  T sx[DERIV ? E : 1], sy = 0, sz = 0;
  if constexpr (DERIV) {
#pragma unroll
    for (int i = 0; i < E; ++i) sx[i] = ((const T *)pg.sinx)[t + TPL * i];
    sy = ((const T *)pg.siny)[o];
    sz = ((const T *)pg.sinz)[kz];
  }
  // coefficient scale * K(k) [* i sin(k_d h)/h] times a, point i
  auto coef = [&](const MidTerm &tm, int i, C a) {
    T s = (T)tm.scale;
    if (NK > 0 && tm.kern >= 0) {
      T k = kv[0][i];
#pragma unroll
      for (int j = 1; j < NK; ++j)
        if (tm.kern == j) k = kv[j][i];
      s *= k;
    }
    if (DERIV && tm.deriv >= 0) {
      s *= tm.deriv == 0 ? sx[DERIV ? i : 0] : (tm.deriv == 1 ? sy : sz);
      return cmk<C>(-s * a.y, s * a.x);
    }
    return cmk<C>(s * a.x, s * a.y);
  };
  __syncthreads();
  C acc[E];
  if (pg.multi >= 0)
#pragma unroll
    for (int i = 0; i < E; ++i) acc[i] = cmk<C>(0, 0);
  // the next input is loaded while the current one is transformed
  for (int ii = 0; ii < pg.nin; ++ii) {
    C v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = nxt[i];
    if (ii + 1 < pg.nin)
#pragma unroll
      for (int i = 0; i < E; ++i) nxt[i] = rf::gld<C>((const char *)pg.in[ii + 1] + i * stC, obC);
    rf::rfft<T, N, G::EM, -1, rf::BlockSync>(v, X, s_tw, t);
    if (pg.multi >= 0) {
      const MidOut &mo = pg.o[pg.multi];
      for (int q = 0; q < mo.nterms; ++q) {
        if (mo.t[q].in != ii) continue;
        const MidTerm tm = mo.t[q];
#pragma unroll
        for (int i = 0; i < E; ++i) acc[i] = cadd(acc[i], coef(tm, i, v[i]));
      }
    }
    for (int m = 0; m < pg.nout; ++m) {
      if (m == pg.multi || pg.o[m].t[0].in != ii) continue;
      const MidTerm tm = pg.o[m].t[0];
      C u[E];
#pragma unroll
      for (int i = 0; i < E; ++i) u[i] = coef(tm, i, v[i]);
      rf::rfft<T, N, G::EM, +1, rf::BlockSync>(u, X, s_tw, t);
      if (act)
#pragma unroll
        for (int i = 0; i < E; ++i) rf::gst<C>((char *)pg.o[m].out + i * stC, obC, u[i]);
    }
  }
  if (pg.multi >= 0) {
    rf::rfft<T, N, G::EM, +1, rf::BlockSync>(acc, X, s_tw, t);
    if (act)
#pragma unroll
      for (int i = 0; i < E; ++i) rf::gst<C>((char *)pg.o[pg.multi].out + i * stC, obC, acc[i]);
  }
}

// ----------------------------------------------------------------- xconv --
// The x-middle program one output per block: out_m = F_x^-1 sum_t c_t F_x in_t
// over the terms of output m.  An input shared by several outputs is read once
// per output; the outputs of one column tile are consecutive logical blocks
// on one XCD, so those re-reads mostly hit that XCD's L2.  Holding one line
// (16 points per thread) plus an accumulator keeps the register count low
// enough for four waves per SIMD.  Single-term outputs take 32 columns per
// block (256-byte row segments in FP32), the accumulating form 16.
// DFT_XCONV1_EM / DFT_XCONV1_COLS set that geometry for the single-term form,
// DFT_XCONV_EM / DFT_XCONV_COLS for the accumulating one; DFT_XCONV_MINB asks
// the register allocator for that many waves per SIMD (0: no request).
// Without the next-term prefetch below, 4 fit the accumulating form in 128
// VGPRs with 60 bytes per lane of spills: 165 -> 144 us at 256^3 FP32 (152
// VGPRs, 3 waves, without the request); with the prefetch, 2.
#ifndef DFT_XCONV1_EM
#define DFT_XCONV1_EM 16
#endif
#ifndef DFT_XCONV1_COLS
#define DFT_XCONV1_COLS 32
#endif
#ifndef DFT_XCONV_EM
#define DFT_XCONV_EM 16
#endif
#ifndef DFT_XCONV_COLS
#define DFT_XCONV_COLS 16
#endif
#ifndef DFT_XCONV_MINB
#define DFT_XCONV_MINB 2
#endif
// FP32 accumulating outputs read the next term's input and kernel before the
// current term's transform (+32 VGPRs: with it they run at 2 waves per SIMD
// without spills, 146 -> 139 us at 256^3; at 3 they spill and take 161 us)
#ifndef DFT_XCONV_PF
#define DFT_XCONV_PF 1
#endif
// the same request for the FP64 instantiations (0: none)
#ifndef DFT_XCONV_MINB64
#define DFT_XCONV_MINB64 2
#endif
// (N = 512 keeps 16 points per thread and 16 columns.  FP64 single-term
// outputs take DFT_XCONV1_COLS64 columns: 16 measured faster than 32, which
// fill the LDS with one block per CU; the accumulating form takes at most 32.)
#ifndef DFT_XCONV1_COLS64
#define DFT_XCONV1_COLS64 16
#endif
template <typename T, int N, bool ONE>
using XConvCfg = ColCfg<T, N, N >= 512 ? 16 : (ONE ? DFT_XCONV1_EM : DFT_XCONV_EM),
                        N >= 512 ? 16
                                 : (ONE ? (sizeof(T) == 8 ? DFT_XCONV1_COLS64 : DFT_XCONV1_COLS)
                                        : (sizeof(T) == 8 && DFT_XCONV_COLS > 32 ? 32 : DFT_XCONV_COLS)),
                        !(ONE && N == 128)>;

template <typename T, int N, bool DERIV, bool ONE>
__global__ void __launch_bounds__((XConvCfg<T, N, ONE>::BS), sizeof(T) == 4 ? DFT_XCONV_MINB : DFT_XCONV_MINB64) k_xconv(const MidProg *__restrict__ pgp, int m0, int nl, int ny,
                                                              int P, int PP, const cplx<T> *__restrict__ g_tw) {
  const MidProg &pg = *pgp;
  using C = cplx<T>;
  using G = XConvCfg<T, N, ONE>;
  constexpr int E = G::E, TPL = G::TPL, COLS = G::COLS, LP = G::LP;
  __shared__ C s_x[G::S > 1 ? COLS * LP : 1];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  // outputs m0 .. m0 + nl - 1 of the program
  const int L = rf::xcd_block();
  const int b = L / nl, m = m0 + (L - b * nl);
  const int col = threadIdx.x % COLS, t = threadIdx.x / COLS;
  const int ntile = mid_tiles<COLS>(P);
  int o, kz;
  bool act;
  if (b < ny * ntile) {
    o = b / ntile;
    const int c = (b - o * ntile) * COLS + col;
    act = c < P - 1;
    kz = act ? c : P - 2;  // out-of-range lanes read an in-range column and store nothing
  } else {
    const int r = (b - ny * ntile) * COLS + col;
    act = r < ny;
    o = act ? r : ny - 1;
    kz = P - 1;
  }
  const uint32_t es = (uint32_t)ny * PP;
  const uint32_t e0 = (uint32_t)o * PP + kz + t * es;  // element index of point t
  const uint32_t obC = e0 * (uint32_t)sizeof(C), obR = e0 * (uint32_t)sizeof(T);
  const size_t stC = (size_t)TPL * es * sizeof(C), stR = (size_t)TPL * es * sizeof(T);  // per point i
  C *X = s_x + col * LP;
  const MidOut &mo = pg.o[m];
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  if constexpr (ONE) {
    // single-term outputs: transform, scale and transform back in place
    const MidTerm tm = mo.t[0];
    C v[E];
    T k[E];
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = rf::gld<C>((const char *)pg.in[tm.in] + i * stC, obC);
#pragma unroll
    for (int i = 0; i < E; ++i) k[i] = rf::gld<T>((const char *)pg.kern[tm.kern] + i * stR, obR);
    __syncthreads();
    rf::rfft<T, N, G::EM, -1, rf::BlockSync>(v, X, s_tw, t);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const T sc = (T)tm.scale * k[i];
      v[i] = cmk<C>(sc * v[i].x, sc * v[i].y);
    }
    rf::rfft<T, N, G::EM, +1, rf::BlockSync>(v, X, s_tw, t);
    if (act)
#pragma unroll
      for (int i = 0; i < E; ++i) rf::gst<C>((char *)mo.out + i * stC, obC, v[i]);
    return;
  }
  C acc[E];
#pragma unroll
  for (int i = 0; i < E; ++i) acc[i] = cmk<C>(0, 0);
  // term q's input and kernel; with DFT_XCONV_PF (FP32) term q + 1's are read
  // before term q's transform
  constexpr bool PF = sizeof(T) == 4 && DFT_XCONV_PF;
  C nv[PF ? E : 1];
  T nk[PF ? E : 1];
  auto fetch = [&](int qq, C *fv, T *fk) {
    const MidTerm tq = mo.t[qq];
#pragma unroll
    for (int i = 0; i < E; ++i) fv[i] = rf::gld<C>((const char *)pg.in[tq.in] + i * stC, obC);
    if (tq.kern >= 0) {
#pragma unroll
      for (int i = 0; i < E; ++i) fk[i] = rf::gld<T>((const char *)pg.kern[tq.kern] + i * stR, obR);
    } else {
#pragma unroll
      for (int i = 0; i < E; ++i) fk[i] = (T)1;
    }
  };
  if constexpr (PF) fetch(0, nv, nk);
  for (int q = 0; q < mo.nterms; ++q) {
    const MidTerm tm = mo.t[q];
    C v[E];
    T k[E];
    if constexpr (PF) {
#pragma unroll
      for (int i = 0; i < E; ++i) v[i] = nv[i], k[i] = nk[i];
      if (q + 1 < mo.nterms) fetch(q + 1, nv, nk);
    } else {
      fetch(q, v, k);
    }
    if (q == 0) __syncthreads();  // twiddles
    rf::rfft<T, N, G::EM, -1, rf::BlockSync>(v, X, s_tw, t);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      T sc = (T)tm.scale * k[i];
      if (DERIV && tm.deriv >= 0) {
        sc *= tm.deriv == 0 ? ((const T *)pg.sinx)[t + TPL * i]
                            : (tm.deriv == 1 ? ((const T *)pg.siny)[o] : ((const T *)pg.sinz)[kz]);
        acc[i] = cadd(acc[i], cmk<C>(-sc * v[i].y, sc * v[i].x));
      } else {
        acc[i] = cadd(acc[i], cmk<C>(sc * v[i].x, sc * v[i].y));
      }
    }
  }
  rf::rfft<T, N, G::EM, +1, rf::BlockSync>(acc, X, s_tw, t);
  if (act)
#pragma unroll
    for (int i = 0; i < E; ++i) rf::gst<C>((char *)mo.out + i * stC, obC, acc[i]);
}

// ------------------------------------------------------------------- zAB --
// minimum resident blocks per CU asked of the register allocator for the FP32
// zAB / zBC kernels (0: none)
#ifndef DFT_ZAB_MINB
#define DFT_ZAB_MINB 0
#endif
#ifndef DFT_ZBC_MINB
#define DFT_ZBC_MINB 2
#endif
// the same requests for the single FP64 kernels (0: none), used when the
// split forms below are off.  FP64 zBC at 2 waves per SIMD spills 240 bytes per
// lane and still beats 1 wave (256 VGPRs + 56 AGPRs): 192.5 -> 198 steps/s at
// 256^3; zAB at 2 waves (412 bytes of spills) is slower.
#ifndef DFT_ZAB_MINB64
#define DFT_ZAB_MINB64 0
#endif
#ifndef DFT_ZBC_MINB64
#define DFT_ZBC_MINB64 2
#endif
// FP64 zAB (backflow) and zBC as two kernels each (1) or one (0), and the
// waves per SIMD asked for the halves (0: none).  In FP64 neither half spills:
// 200 -> 229 steps/s at 256^3 with both.
#ifndef DFT_ZAB_SPLIT64
#define DFT_ZAB_SPLIT64 1
#endif
#ifndef DFT_ZBC_SPLIT64
#define DFT_ZBC_SPLIT64 1
#endif
#ifndef DFT_ZAB_SPLIT_MINB64
#define DFT_ZAB_SPLIT_MINB64 0
#endif
#ifndef DFT_ZBC_SPLIT_MINB64
#define DFT_ZBC_SPLIT_MINB64 0
#endif
// (the zAB split in FP32, where neither half gets past 2 waves per SIMD without
// spills, measured slower: zAB 375 -> 386 us at 256^3)
#ifndef DFT_ZAB_SPLIT32
#define DFT_ZAB_SPLIT32 0
#endif
// the same for the FP32 kernels with the HD / HD2 backflow switch (one wave per
// SIMD in one kernel)
#ifndef DFT_ZAB_SPLIT32_HB
#define DFT_ZAB_SPLIT32_HB 1
#endif
// Early loads in zAB / zBC, so that memory round trips overlap the transforms
// (bit masks, FP32; FP64 holds twice the registers and spills at its
// occupancy, so it uses them only with DFT_PREFETCH64=1):
//  1  a stage's stencils (velocity, density gradients) before its spectrum's transform;
//  2  the density (zAB) / the accumulated fields (zBC) before the last transform;
//  4  zAB: the KC density gradient along the next axis before the current r2c;
//  8  the next stage's spectrum before the current stage's transform;
// zC (bit 1): the fields of each line of the step before the preceding transform.
#ifndef DFT_ZAB_PF
#define DFT_ZAB_PF 15
#endif
#ifndef DFT_ZBC_PF
#define DFT_ZBC_PF 11
#endif
#ifndef DFT_ZC_PF
#define DFT_ZC_PF 1
#endif
#ifndef DFT_PREFETCH64
#define DFT_PREFETCH64 0
#endif
template <typename T, int MASK> constexpr int prefetch_bits() { return (sizeof(T) == 4 || DFT_PREFETCH64) ? MASK : 0; }

// Pipelined c2r stages: the spectrum of the current stage was loaded into
// `raw` earlier; the next stage's loads (if any) are issued before this
// stage's transform.
template <typename T, int N>
__device__ __forceinline__ void c2r_pipe(rf::Regs<T, N, pair_em<N>()> &v, fft::RawPair<T, N, pair_em<N>()> &raw,
                                         cplx<T> *X, const cplx<T> *tw, int t, const void *next, uint32_t p, int P) {
  fft::merge_raw<T, N, pair_em<N>()>(v, raw, t);
  if (next) fft::load_raw<T, N, pair_em<N>()>(raw, next, p, P, t);
  rf::rfft<T, N, pair_em<N>(), +1, PairSync<N>>(v, X, tw, t);
}
struct ZABArgs {
  Ptrs8 in;       // rhobar, [rhot], [A, Bx, By, Bz, C]
  Ptrs8 out;      // Q, [G_x, G_y, G_z]
  const void *rho, *psi;
  void *rhot, *vre, *vim, *wx, *wy, *wz;
};

// PART 0: the whole stage; with backflow it may run as two kernels with fewer
// live fields, PART 1 (the backflow terms, V_re stored) then PART 2 (the local
// and KC terms, added to the stored V_re).
// HB: the backflow density switch of HD / HD2 (A, B_d, C come from rho~; Re
// V_BF is multiplied by d(G rho)/d rho, the gradient half of Im V_BF takes rho~).
template <typename T, int N, bool KC, bool BF, int PART, bool HB>
__device__ __forceinline__ void zAB_body(ZABArgs a, int64_t npairs, int P, PW<T> pw,
                                         const cplx<T> *__restrict__ g_tw) {
  static_assert(PART == 0 || BF, "only the backflow stage splits");
  FUSED_PAIR_SETUP();
  // (with the switch the stencils' exponentials would be held across the transform)
  constexpr int PF = prefetch_bits<T, HB ? (DFT_ZAB_PF & ~1) : DFT_ZAB_PF>();
  constexpr int IA = KC ? 2 : 1;  // first backflow input
  constexpr bool DO_BF = BF && PART != 2, DO_LOCAL = PART != 1;
  const void *rho = a.rho;
  C v[E];
  T vre[2][E];
  fft::RawPair<T, N, pair_em<N>()> raw;
  // c2r of stage input `cur`, then (with bit 8) the loads of `next`
  auto c2r = [&](const void *cur, const void *next) {
    if constexpr ((PF & 8) != 0) c2r_pipe<T, N>(v, raw, X, s_tw, t, next, p, P);
    else load_c2r<T, N>(v, X, s_tw, t, cur, p, P);
  };
  if constexpr ((PF & 8) != 0) fft::load_raw<T, N, pair_em<N>()>(raw, a.in.p[DO_BF ? IA : 0], p, P, t);
  if constexpr (PART == 2) {
    const auto qv = q.here();
#pragma unroll
    for (int i = 0; i < E; ++i)
#pragma unroll
      for (int w = 0; w < 2; ++w) vre[w][i] = qv.template at_once<T>(a.vre, w, i);
  }
  if constexpr (DO_BF) {
    // A = U*rho; per axis: B_d = U*q_d -> W_d = v_d A - B_d (stored),
    // BF real part sum_d (v_d^2 A - 2 v_d B_d) + C, imaginary part sum_d (g_d/(rho+eps)) W_d / 2
    T A[2][E], tr[2][E], im[2][E];
    c2r(a.in.p[IA], a.in.p[IA + 1]);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      A[0][i] = v[i].x;
      A[1][i] = v[i].y;
    }
    const void *psi = a.psi;
    // velocity factor and 1 / (rho + eps) per point, kept for the three axes
    T fv[2][E], fg[2][E], dg[2][HB ? E : 1];
    {
      const auto qr = q.here();
#pragma unroll
      for (int i = 0; i < E; ++i)
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          const T r = qr.template at<T>(rho, w, i);
          fv[w][i] = velocity_factor(pw, r);
          fg[w][i] = (T)1 / (r + pw.div_eps);
          if constexpr (HB) dg[w][i] = bf_dg(r, pw);
        }
    }
    auto axis = [&](auto dtag) {
      constexpr int D = decltype(dtag)::value;
      void *wd = D == 0 ? a.wx : (D == 1 ? a.wy : a.wz);
      // stencil positions taken after the transform unless prefetching, so
      // that the compiler does not hoist those loads into it
      std::remove_const_t<decltype(q)> qq;
      if constexpr ((PF & 1) != 0) qq = q.here();
      // velocity and (D_d rho)/(rho + eps) at point (w, i)
      auto vg = [&](int w, int i, T &vd, T &gi) {
        vd = velocity_d<T, D>(psi, qq, w, i, pw, fv[w][i]);
        if constexpr (HB) gi = grad_rt_d<T, D>(rho, qq, w, i, pw) * fg[w][i];
        else gi = grad_d<T, D>(rho, qq, w, i, pw) * fg[w][i];
      };
      T vds[2][E], gis[2][E];
      if constexpr ((PF & 1) != 0)
#pragma unroll
        for (int i = 0; i < E; ++i)
#pragma unroll
          for (int w = 0; w < 2; ++w) vg(w, i, vds[w][i], gis[w][i]);
      c2r(a.in.p[IA + 1 + D], a.in.p[IA + 2 + D]);
      if constexpr ((PF & 1) == 0) qq = q.here();
      T Wv[2][E];
#pragma unroll
      for (int i = 0; i < E; ++i) {
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          const T Bd = part(v[i], w);
          T vd, gi;
          if constexpr ((PF & 1) != 0) vd = vds[w][i], gi = gis[w][i];
          else vg(w, i, vd, gi);
          const T W = vd * A[w][i] - Bd;
          Wv[w][i] = W;
          const T ti = vd * vd * A[w][i] + (T)-2 * vd * Bd;
          const T ii = (T)0.5 * gi * W;
          if constexpr (D == 0) {
            tr[w][i] = ti;
            im[w][i] = ii;
          } else {
            tr[w][i] += ti;
            im[w][i] += ii;
          }
        }
      }
      if (act)
#pragma unroll
        for (int i = 0; i < E; ++i) {
          q.template put<T>(wd, 0, i, Wv[0][i]);
          q.template put<T>(wd, 1, i, Wv[1][i]);
        }
    };
    axis(std::integral_constant<int, 0>{});
    axis(std::integral_constant<int, 1>{});
    axis(std::integral_constant<int, 2>{});
    c2r(a.in.p[IA + 4], DO_LOCAL ? a.in.p[0] : nullptr);
#pragma unroll
    for (int i = 0; i < E; ++i) {
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        T re = tr[w][i] + part(v[i], w);
        if constexpr (HB) re = re * dg[w][i];
        re = re * ((T)-0.5 * pw.mass);
        if (pw.clamp_bf) re = clampv(re, pw.bfmax);
        vre[w][i] = re * pw.cbf;
      }
    }
    if (act)
#pragma unroll
      for (int i = 0; i < E; ++i) {
        q.template put<T>(a.vim, 0, i, im[0][i]);
        q.template put<T>(a.vim, 1, i, im[1][i]);
      }
  }
  if constexpr (DO_LOCAL) {
    // local correlation: rhobar = Pi*rho
    std::remove_const_t<decltype(q)> ql;
    T rl[2][E];
    auto dens = [&] {
      ql = q.here();
#pragma unroll
      for (int i = 0; i < E; ++i)
#pragma unroll
        for (int w = 0; w < 2; ++w) rl[w][i] = ql.template at<T>(rho, w, i);
    };
    if constexpr ((PF & 2) != 0) dens();
    c2r(a.in.p[0], KC ? a.in.p[1] : nullptr);
    if constexpr ((PF & 2) == 0) dens();
#pragma unroll
    for (int i = 0; i < E; ++i) {
      T Q[2];
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        const T rb = part(v[i], w), r = rl[w][i];
        T loc = rb * rb * (pw.c2 * (T)0.5) + rb * rb * rb * (pw.c3 / (T)3);
        if (pw.hdC != (T)0) loc += hd_pot(r, pw);
        if constexpr (BF) vre[w][i] += loc;
        else vre[w][i] = loc;
        Q[w] = (rb * pw.c2 + rb * rb * pw.c3) * r;
      }
      v[i] = cmk<C>(Q[0], Q[1]);
    }
    r2c_store<T, N>(v, X, s_tw, t, a.out.p[0], p, P, act);
    if constexpr (KC) {
      // rhot = F*rho (stored), rho_st = 1 - rhot/rho_0s, G_d = (D_d rho) rho_st
      T gn[2][E];
      auto grad_next = [&](auto dtag) {
        constexpr int D = decltype(dtag)::value;
        const auto qq = q.here();
#pragma unroll
        for (int i = 0; i < E; ++i)
#pragma unroll
          for (int w = 0; w < 2; ++w) gn[w][i] = grad_d<T, D>(rho, qq, w, i, pw);
      };
      if constexpr ((PF & 4) != 0) grad_next(std::integral_constant<int, 0>{});
      c2r(a.in.p[1], nullptr);
      T rst[2][E];
#pragma unroll
      for (int i = 0; i < E; ++i) {
#pragma unroll
        for (int w = 0; w < 2; ++w) rst[w][i] = part(v[i], w) * (-pw.inv_rho0s) + (T)1;
      }
      if (act)
#pragma unroll
        for (int i = 0; i < E; ++i) {
          q.template put<T>(a.rhot, 0, i, v[i].x);
          q.template put<T>(a.rhot, 1, i, v[i].y);
        }
      auto gd = [&](auto dtag) {
        constexpr int D = decltype(dtag)::value;
        if constexpr ((PF & 4) == 0) grad_next(dtag);
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = cmk<C>(gn[0][i] * rst[0][i], gn[1][i] * rst[1][i]);
        if constexpr ((PF & 4) != 0 && D < 2) grad_next(std::integral_constant<int, D + 1>{});
        r2c_store<T, N>(v, X, s_tw, t, a.out.p[1 + D], p, P, act);
      };
      gd(std::integral_constant<int, 0>{});
      gd(std::integral_constant<int, 1>{});
      gd(std::integral_constant<int, 2>{});
    }
  }
  if (act) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      q.template put<T>(a.vre, 0, i, vre[0][i]);
      q.template put<T>(a.vre, 1, i, vre[1][i]);
    }
  }
}
template <typename T, int N, bool KC, bool BF, int PART = 0>
__global__ void __launch_bounds__((PairBS<T, N>), sizeof(T) == 4 ? DFT_ZAB_MINB : (PART ? DFT_ZAB_SPLIT_MINB64 : DFT_ZAB_MINB64))
    k_zAB(ZABArgs a, int64_t npairs, int P, PW<T> pw, const cplx<T> *__restrict__ g_tw) {
  zAB_body<T, N, KC, BF, PART, false>(a, npairs, P, pw, g_tw);
}
template <typename T, int N, bool KC, int PART = 0>
__global__ void __launch_bounds__((PairBS<T, N>), sizeof(T) == 4 ? DFT_ZAB_MINB : (PART ? DFT_ZAB_SPLIT_MINB64 : DFT_ZAB_MINB64))
    k_zAB_bfhd(ZABArgs a, int64_t npairs, int P, PW<T> pw, const cplx<T> *__restrict__ g_tw) {
  zAB_body<T, N, KC, true, PART, true>(a, npairs, P, pw, g_tw);
}

// ------------------------------------------------------------------- zBC --
struct ZBCArgs {
  Ptrs8 in;   // J_x, J_y, J_z, T1
  Ptrs8 out;  // H
  const void *rho, *rhot, *wx, *wy, *wz;
  void *vre, *vim;
  void *hbuf, *tbuf;  // split form: H and the term-3 sum between the two kernels (real fields)
};

// Every stage first issues the loads that do not depend on its spectrum (the
// stencils of rho and rho~, the fields accumulated into), so that their
// latency overlaps the spectrum's loads and transform.
// PART 0: the whole stage; PART 1 (the three J_d stages, H and the term-3
// sum stored) then PART 2 (T1, the tail and the r2c of H): the split form FP64
// runs with fewer live fields.
// HB: the backflow density switch of HD / HD2 (the divergence half of Im V_BF
// is multiplied by G(rho)).
template <typename T, int N, bool BF, int PART, bool HB>
__device__ __forceinline__ void zBC_body(ZBCArgs a, int64_t npairs, int P, PW<T> pw,
                                         const cplx<T> *__restrict__ g_tw) {
  FUSED_PAIR_SETUP();
  constexpr int PF = prefetch_bits<T, DFT_ZBC_PF>();
  const void *rho = a.rho, *rhot = a.rhot;
  C v[E];
  T H[2][PART == 2 ? 1 : E], t3[2][E];
  fft::RawPair<T, N, pair_em<N>()> raw;
  auto c2r = [&](const void *cur, const void *next) {
    if constexpr ((PF & 8) != 0) c2r_pipe<T, N>(v, raw, X, s_tw, t, next, p, P);
    else load_c2r<T, N>(v, X, s_tw, t, cur, p, P);
  };
  if constexpr ((PF & 8) != 0) fft::load_raw<T, N, pair_em<N>()>(raw, a.in.p[PART == 2 ? 3 : 0], p, P, t);
  auto axis = [&](auto dtag) {
    constexpr int D = decltype(dtag)::value;
    T g[2][E], Pd[2][E];
    auto stencil = [&] {
      const auto qq = q.here();
#pragma unroll
      for (int i = 0; i < E; ++i)
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          g[w][i] = grad_d<T, D>(rho, qq, w, i, pw);
          Pd[w][i] = grad_d<T, D>(rhot, qq, w, i, pw);  // (dF_d * rho) = D_d (F * rho)
        }
    };
    if constexpr ((PF & 1) != 0) stencil();
    c2r(a.in.p[D], a.in.p[D + 1]);
    if constexpr ((PF & 1) == 0) stencil();
#pragma unroll
    for (int i = 0; i < E; ++i) {
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        const T J = part(v[i], w);
        if constexpr (D == 0) {
          H[w][i] = g[w][i] * J;
          t3[w][i] = Pd[w][i] * J;
        } else {
          H[w][i] += g[w][i] * J;
          t3[w][i] += Pd[w][i] * J;
        }
      }
    }
  };
  if constexpr (PART != 2) {
    axis(std::integral_constant<int, 0>{});
    axis(std::integral_constant<int, 1>{});
    axis(std::integral_constant<int, 2>{});
  }
  if constexpr (PART == 1) {
    // first kernel of the split form: H and the term-3 sum to memory
    if (act) {
      const auto qs = q.here();
#pragma unroll
      for (int i = 0; i < E; ++i)
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          qs.template put<T>(a.hbuf, w, i, H[w][i]);
          qs.template put<T>(a.tbuf, w, i, t3[w][i]);
        }
    }
    return;
  } else if constexpr (PART == 2) {
    const auto qs = q.here();
#pragma unroll
    for (int i = 0; i < E; ++i)
#pragma unroll
      for (int w = 0; w < 2; ++w) t3[w][i] = qs.template at_once<T>(a.tbuf, w, i);
  }
  // rho_st, V_re so far, and the backflow imaginary part (which does not
  // involve T1); with prefetching, all before T1's transform
  std::remove_const_t<decltype(q)> qt;
  auto rst_at = [&](int w, int i) { return qt.template at<T>(rhot, w, i) * (-pw.inv_rho0s) + (T)1; };
  auto vre_at = [&](int w, int i) { return qt.template at_once<T>(a.vre, w, i); };
  auto kc_at = [&](T J, T rst, int w, int i) { return J * rst * pw.ckc - t3[w][i] * (pw.ckc * pw.k23); };
  auto bf_im = [&](int w, int i) {
    if constexpr (BF) {
      T im = qt.template at_once<T>(a.vim, w, i);
      if constexpr (HB) {
        const T gw = bf_g(qt.template at<T>(rho, w, i), pw);
        im += (T)0.5 * (((qt.template nb<T, 0, 1>(a.wx, w, i) - qt.template nb<T, 0, -1>(a.wx, w, i)) * pw.inv2h) * gw);
        im += (T)0.5 * (((qt.template nb<T, 1, 1>(a.wy, w, i) - qt.template nb<T, 1, -1>(a.wy, w, i)) * pw.inv2h) * gw);
        im += (T)0.5 * (((qt.template nb<T, 2, 1>(a.wz, w, i) - qt.template nb<T, 2, -1>(a.wz, w, i)) * pw.inv2h) * gw);
      } else {
        im += (T)0.5 * ((qt.template nb<T, 0, 1>(a.wx, w, i) - qt.template nb<T, 0, -1>(a.wx, w, i)) * pw.inv2h);
        im += (T)0.5 * ((qt.template nb<T, 1, 1>(a.wy, w, i) - qt.template nb<T, 1, -1>(a.wy, w, i)) * pw.inv2h);
        im += (T)0.5 * ((qt.template nb<T, 2, 1>(a.wz, w, i) - qt.template nb<T, 2, -1>(a.wz, w, i)) * pw.inv2h);
      }
      if (pw.clamp_bf) im = clampv(im, pw.bfmax);
      if (act) qt.template put<T>(a.vim, w, i, im * pw.cbf);
    }
  };
  if constexpr ((PF & 2) != 0) {
    qt = q.here();
    T rst[2][E], base[2][E];
#pragma unroll
    for (int i = 0; i < E; ++i)
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        rst[w][i] = rst_at(w, i);
        base[w][i] = vre_at(w, i);
        bf_im(w, i);
      }
    c2r(a.in.p[3], nullptr);
    if (act)
#pragma unroll
      for (int i = 0; i < E; ++i)
#pragma unroll
        for (int w = 0; w < 2; ++w) qt.template put<T>(a.vre, w, i, base[w][i] + kc_at(part(v[i], w), rst[w][i], w, i));
  } else {
    c2r(a.in.p[3], nullptr);
    qt = q.here();
    if (act)
#pragma unroll
      for (int i = 0; i < E; ++i)
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          qt.template put<T>(a.vre, w, i, vre_at(w, i) + kc_at(part(v[i], w), rst_at(w, i), w, i));
          bf_im(w, i);
        }
  }
  if constexpr (PART == 2) {
    const auto qs = q.here();
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = cmk<C>(qs.template at_once<T>(a.hbuf, 0, i), qs.template at_once<T>(a.hbuf, 1, i));
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = cmk<C>(H[0][i], H[1][i]);
  }
  r2c_store<T, N>(v, X, s_tw, t, a.out.p[0], p, P, act);
}
template <typename T, int N, bool BF, int PART = 0>
__global__ void __launch_bounds__((PairBS<T, N>), sizeof(T) == 4 ? DFT_ZBC_MINB : (PART ? DFT_ZBC_SPLIT_MINB64 : DFT_ZBC_MINB64))
    k_zBC(ZBCArgs a, int64_t npairs, int P, PW<T> pw, const cplx<T> *__restrict__ g_tw) {
  zBC_body<T, N, BF, PART, false>(a, npairs, P, pw, g_tw);
}
template <typename T, int N, int PART = 0>
__global__ void __launch_bounds__((PairBS<T, N>), sizeof(T) == 4 ? DFT_ZBC_MINB : (PART ? DFT_ZBC_SPLIT_MINB64 : DFT_ZBC_MINB64))
    k_zBC_bfhd(ZBCArgs a, int64_t npairs, int P, PW<T> pw, const cplx<T> *__restrict__ g_tw) {
  zBC_body<T, N, true, PART, true>(a, npairs, P, pw, g_tw);
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
  void *phase;           // complex out: V_eff of the step (the inverse z pass applies e^{-i V_eff dt/2})
  const void *psi_src;
  void *psi_dst;         // z forward transform of psi_src * phase
  double dtr, dti;
};

template <typename T, int N, bool STEP>
__global__ void __launch_bounds__((PairBS<T, N>), sizeof(T) == 4 ? DFT_ZC_MINB : 0) k_zC(ZCArgs a, int64_t npairs, int P, const cplx<T> *__restrict__ g_tw) {
  using C = cplx<T>;
  using G = LineCfg<T, N, pair_em<N>()>;
  constexpr int E = G::E, TPL = G::TPL, LP = G::LP;
  __shared__ C s_x[G::LINES * LP];
  __shared__ C s_tw[rf::tw_len<T, N>()];
  const int ln = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const uint32_t p0 = (uint32_t)blockIdx.x * G::LINES + ln;
  const bool act = p0 < npairs;
  const uint32_t p = act ? p0 : 0;
  C *X = s_x + ln * LP;
  rf::load_tw<T, N, G::EM, G::BS>(s_tw, g_tw);
  __syncthreads();
  constexpr bool PF = prefetch_bits<T, DFT_ZC_PF>() != 0;
  C v[E];
  auto c2r_in = [&] { load_c2r<T, N>(v, X, s_tw, t, a.in.p[0], p, P); };
  const uint32_t cb = p * 2u * N + (uint32_t)t;
  if constexpr (!STEP) {
    c2r_in();
    const uint32_t c0 = cb + rf::opaque_zero();
    if (act) {
#pragma unroll
      for (int i = 0; i < E; ++i)
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          const int d = w * N + TPL * i;
          C u = rf::lde<C>(a.V, c0, d);
          u.x += rf::lde<T>(a.vre, c0, d) + part(v[i], w);
          u.y += a.vim ? rf::lde<T>(a.vim, c0, d) : (T)0;
          rf::ste<C>(a.V, c0, u, d);
        }
    }
  } else {
    // the fields of line w; with prefetching, line 0's are read before the
    // transform above and line 1's before line 0's transform
    T bre[E], bim[E], bex[E];
    C bav[E], bps[E];
    auto fetch = [&](int w) {
      const uint32_t c0 = cb + rf::opaque_zero();
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int d = w * N + TPL * i;
        bre[i] = rf::lde_once<T>(a.vre, c0, d);
        bim[i] = a.vim ? rf::lde_once<T>(a.vim, c0, d) : (T)0;
        bex[i] = a.vext ? rf::lde_once<T>(a.vext, c0, d) : (T)0;
        if (a.vavg) bav[i] = rf::lde_once<C>(a.vavg, c0, d);
        bps[i] = rf::lde_once<C>(a.psi_src, c0, d);
      }
    };
    if constexpr (PF) fetch(0);
    c2r_in();
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      if constexpr (!PF) fetch(w);
      const uint32_t c0 = cb + rf::opaque_zero();
      C u[E], ph[E], vs[E];  // ph: the potential of this half step (re-applied by the inverse z pass)
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const T re = bre[i] + part(v[i], w);
        const T im = bim[i];
        C ve = cmk<C>(re - (T)a.mu + bex[i], im);
        if (a.vsave) vs[i] = ve;
        if (a.vavg) {
          const C va = bav[i];
          ve = cmk<C>((va.x + ve.x) * (T)0.5, (va.y + ve.y) * (T)0.5);
        }
        ph[i] = ve;
        u[i] = fft::apply_vphase<T>(bps[i], ve, a.dtr, a.dti);
      }
      if (act) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
          rf::ste<C>(a.phase, c0, ph[i], w * N + TPL * i);
          if (a.vsave) rf::ste<C>(a.vsave, c0, vs[i], w * N + TPL * i);
        }
      }
      if constexpr (PF) {
        if (w == 0) fetch(1);
      }
      rf::rfft<T, N, pair_em<N>(), -1, PairSync<N>>(u, X, s_tw, t);
      if (act)
#pragma unroll
        for (int i = 0; i < E; ++i) rf::ste<C>(a.psi_dst, c0, u[i], w * N + TPL * i);
    }
  }
}

}  // namespace fused
}  // namespace dft
