// Batched 3-D FFT passes for gfx950, hand-written (no rocFFT / hipFFT).
//
// A 3-D transform is three passes, one per axis.  Each pass stages a tile of
// whole lines in LDS and runs a Stockham autosort FFT on them.  Each stage is
// a radix-R butterfly held in VGPRs.  Global <-> LDS copies use a coalesced
// thread mapping.  The butterflies use a line-fastest mapping over an odd
// LDS pitch, so lanes of one wavefront hit different banks.
//
//  * z pass, real -> half spectrum: two adjacent real z-lines are packed
//    as one complex line (a + ib), transformed, and split with
//    A(k) = (Z(k) + Z*(N-k))/2 and B(k) = (Z(k) - Z*(N-k))/2i.  The inverse
//    pass does the reverse, so real fields cost one complex line FFT per
//    two lines.
//  * y / x passes work on column tiles of the half spectrum
//    [nx][ny][nz/2+1] (or of a full complex field for the propagator).
//  * The inverse x pass can fuse a k-space combination on load:
//    out(k) = sum_t s_t K_t(k) D_t(k) in_t(k), with K_t a real kernel
//    spectrum and D_t = i sin(k_d h)/h, the symbol of the periodic central
//    difference.
// Twiddles are computed in FP64 on the host and rounded once to the field
// precision.  Forward transforms are unnormalised (e^{-i}).
#pragma once
#include "common.hpp"

// Instantiate CALL with `N_` bound to a supported line length n.
#define DFT_DISPATCH_N(n, CALL)                                   \
  switch (n) {                                                    \
    case 4: { constexpr int N_ = 4; CALL; } break;                \
    case 8: { constexpr int N_ = 8; CALL; } break;                \
    case 16: { constexpr int N_ = 16; CALL; } break;              \
    case 32: { constexpr int N_ = 32; CALL; } break;              \
    case 64: { constexpr int N_ = 64; CALL; } break;              \
    case 128: { constexpr int N_ = 128; CALL; } break;            \
    case 256: { constexpr int N_ = 256; CALL; } break;            \
    case 512: { constexpr int N_ = 512; CALL; } break;            \
    default:                                                      \
      fprintf(stderr, "libdft: unsupported FFT length %d\n", n);  \
      abort();                                                    \
  }

namespace dft {
namespace fft {

// ------------------------------------------------------------ radix plans --
template <int N> struct Plan;
template <> struct Plan<4> { static constexpr int S = 1, R0 = 4, R1 = 1, R2 = 1; };
template <> struct Plan<8> { static constexpr int S = 1, R0 = 8, R1 = 1, R2 = 1; };
template <> struct Plan<16> { static constexpr int S = 1, R0 = 16, R1 = 1, R2 = 1; };
template <> struct Plan<32> { static constexpr int S = 2, R0 = 8, R1 = 4, R2 = 1; };
template <> struct Plan<64> { static constexpr int S = 2, R0 = 8, R1 = 8, R2 = 1; };
template <> struct Plan<128> { static constexpr int S = 2, R0 = 16, R1 = 8, R2 = 1; };
template <> struct Plan<256> { static constexpr int S = 2, R0 = 16, R1 = 16, R2 = 1; };
template <> struct Plan<512> { static constexpr int S = 3, R0 = 8, R1 = 8, R2 = 8; };
template <> struct Plan<1024> { static constexpr int S = 3, R0 = 16, R1 = 8, R2 = 8; };

// Threads cooperating on one line, LDS pitch (complex elements) and lines per tile.
template <int N> struct Shape {
  static constexpr int TPL = N / 16 < 1 ? 1 : (N / 16 > 16 ? 16 : N / 16);
  static constexpr int PITCH = N + 4;
};
// Lines per block: 256 threads when the tile fits in `budget` bytes of LDS.
template <typename T, int N, int LINES_FACTOR = 1> struct Tile {
  static constexpr int csize = 2 * sizeof(T);
  static constexpr int budget = 80 * 1024;
  static constexpr int want = 256 / Shape<N>::TPL;
  static constexpr int fit(int t) {
    return (t > 1 && t * LINES_FACTOR * Shape<N>::PITCH * csize + 2 * N * csize > budget) ? fit(t / 2) : t;
  }
  static constexpr int TILE = fit(want);
  static constexpr int THREADS = TILE * Shape<N>::TPL;
};

// ------------------------------------------------- in-register small DFTs --
__host__ __device__ constexpr int brev(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b)
    if (i & (1 << b)) r |= 1 << (bits - 1 - b);
  return r;
}
__host__ __device__ constexpr int lg2(int n) { return n <= 1 ? 0 : 1 + lg2(n / 2); }

// cos and sin of 2 pi m / 16 (m = 0..7), the constants of the in-register DFTs.
constexpr double kC16[8] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                            0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674};
constexpr double kS16[8] = {0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674,
                            1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977};

// v * (hi + lo) for a twiddle carried as an FP32 pair: the rounded value plus
// its rounding residual.  A twiddle rounded to FP32 alone has a modulus error
// of up to ~6e-8 that is the same on every call, so each Fourier mode picks up
// a fixed gain error per transform; over many split-operator steps that adds
// up linearly in the norm and kinetic energy.  With the residual the modulus is
// exact to ~1e-14 and only the (unbiased) rounding of the products remains.
__device__ inline float2 cmul_hl(float2 v, float2 hi, float2 lo) {
  const float cr = v.x * lo.x - v.y * lo.y;
  const float ci = v.x * lo.y + v.y * lo.x;
  return make_float2(__builtin_fmaf(v.x, hi.x, __builtin_fmaf(-v.y, hi.y, cr)),
                     __builtin_fmaf(v.x, hi.y, __builtin_fmaf(v.y, hi.x, ci)));
}

// v * exp(DIR * 2 pi i m / 16), DIR = -1 forward, +1 inverse.
template <int M, int DIR, typename C> __device__ inline C tw16(C a) {
  using T = decltype(a.x);
  if constexpr (M == 0) {
    return a;
  } else if constexpr (M == 4) {
    // multiply by DIR * i
    return DIR > 0 ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x);
  } else if constexpr (sizeof(T) == 4) {
    constexpr float c = (float)kC16[M], s = (float)kS16[M];
    constexpr float cl = (float)(kC16[M] - (double)c), sl = (float)(kS16[M] - (double)s);
    return cmul_hl(a, cmk<C>(c, DIR * s), cmk<C>(cl, DIR * sl));
  } else {
    const T c = (T)kC16[M], s = (T)(DIR * kS16[M]);
    return cmk<C>(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

template <int R, int DIR, int LEN, int J, typename C> __device__ inline void bfly_group(C *v) {
  if constexpr (J < LEN / 2) {
#pragma unroll
    for (int s = 0; s < R; s += LEN) {
      C u = v[s + J];
      C t = tw16<J *(16 / LEN), DIR>(v[s + J + LEN / 2]);
      v[s + J] = cadd(u, t);
      v[s + J + LEN / 2] = csub(u, t);
    }
    bfly_group<R, DIR, LEN, J + 1>(v);
  }
}
template <int R, int DIR, int LEN, typename C> __device__ inline void bfly_levels(C *v) {
  if constexpr (LEN <= R) {
    bfly_group<R, DIR, LEN, 0>(v);
    bfly_levels<R, DIR, LEN * 2>(v);
  }
}
// In-place DFT of R = 2..16 values in registers (natural order in and out).
template <int R, int DIR, typename C> __device__ inline void dft(C *v) {
  constexpr int B = lg2(R);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int r = brev(i, B);
    if (r > i) {
      C t = v[i];
      v[i] = v[r];
      v[r] = t;
    }
  }
  bfly_levels<R, DIR, 2>(v);
}

// ---------------------------------------------------- Stockham LDS stages --
// One radix-R stage over a line in LDS.  Every thread of the block must call
// it (it contains barriers); `active` masks threads whose line is padding.
template <typename T, int N, int R, int Ns, int DIR>
__device__ inline void stage(cplx<T> *L, const cplx<T> *tw, int tl, bool active) {
  using C = cplx<T>;
  constexpr int TPL = Shape<N>::TPL;
  constexpr int NB = N / R;
  constexpr int PER = (NB + TPL - 1) / TPL;
  C v[PER][R];
  if (active) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int j = tl + p * TPL;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = L[j + r * NB];
      if constexpr (Ns > 1) {
        const int k = j & (Ns - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          // table: forward twiddles in [0, N); for FP32 their residuals in [N, 2N)
          const int ti = k * r * (N / (Ns * R));
          C w = tw[ti];
          if (DIR > 0) w.y = -w.y;
          if constexpr (sizeof(T) == 4) {
            C wl = tw[N + ti];
            if (DIR > 0) wl.y = -wl.y;
            v[p][r] = cmul_hl(v[p][r], w, wl);
          } else {
            v[p][r] = cmul(v[p][r], w);
          }
        }
      }
      dft<R, DIR>(v[p]);
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int j = tl + p * TPL;
      const int idxD = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
      for (int r = 0; r < R; ++r) L[idxD + r * Ns] = v[p][r];
    }
  }
  __syncthreads();
}

// Full line FFT (all stages).  Data in natural order before and after.
template <typename T, int N, int DIR>
__device__ inline void line_fft(cplx<T> *L, const cplx<T> *tw, int tl, bool active) {
  using P = Plan<N>;
  stage<T, N, P::R0, 1, DIR>(L, tw, tl, active);
  if constexpr (P::S > 1) stage<T, N, P::R1, P::R0, DIR>(L, tw, tl, active);
  if constexpr (P::S > 2) stage<T, N, P::R2, P::R0 * P::R1, DIR>(L, tw, tl, active);
}

template <typename T, int N> __device__ inline void load_twiddles(cplx<T> *s_tw, const cplx<T> *g_tw) {
  for (int i = threadIdx.x; i < 2 * N; i += blockDim.x) s_tw[i] = g_tw[i];
}

// e^{-i V dt/2} for complex V and complex dt.
template <typename T>
__device__ inline cplx<T> vphase(cplx<T> v, T dtr, T dti) {
  const T a = (T)0.5 * (v.x * dti + v.y * dtr);
  const T ph = (T)-0.5 * (v.x * dtr - v.y * dti);
  T s, c;
  dsincos(ph, &s, &c);
  const T m = dexp(a);
  return cmk<cplx<T>>(m * c, m * s);
}

// -------------------------------------------------------- pass parameters --
constexpr int kMaxBatch = 16;

struct Batch {
  const void *in[kMaxBatch];
  void *out[kMaxBatch];
};

// k-space combination for the inverse x pass.
struct Term {
  int in;       // index into Batch::in
  int kern;     // index into KernTab::k, -1 = none
  int deriv;    // 0,1,2 = multiply by i sin(k_d h)/h, -1 = none
  double scale;
};
struct Combine {
  int nterms;
  Term t[4];
};
struct CombineSet {
  Combine c[kMaxBatch];
};
struct KernTab {
  const void *k[8];    // real half-spectrum arrays [nx][ny][P]
  const void *sinx, *siny, *sinz;  // sin(2 pi m / n) / h tables
};

}  // namespace fft
}  // namespace dft
