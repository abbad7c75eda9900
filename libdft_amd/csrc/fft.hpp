// Batched 3-D FFT building blocks for gfx950, hand-written (no rocFFT / hipFFT).
//
// A 3-D transform is three passes, one per axis; each pass transforms whole
// lines held in registers (rfft.hpp) with a Stockham autosort radix plan per
// length (below) and in-register radix-2..16 DFTs.
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
// Twiddles are computed in FP64 on the host; FP32 carries each as a rounded
// value plus its residual.  Forward transforms are unnormalised (e^{-i}).
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

// cmul_hl for factors of any angle (the phase factors of the propagator).
// The product with the larger coefficient goes outermost: nested inside, a
// product that is exact (a coefficient of 1, as for an angle below ~2.4e-4,
// or of another short mantissa) leaves the residual term below its last bit,
// where the rounding drops it, always shrinking the factor's angle and
// raising its modulus to 1/cos: a norm gain of up to 1 - cos(angle) per use.
// The FFT twiddles are not of that kind and keep the cheaper cmul_hl.
__device__ inline float2 cmul_hl_phase(float2 v, float2 hi, float2 lo) {
  const float cr = v.x * lo.x - v.y * lo.y;
  const float ci = v.x * lo.y + v.y * lo.x;
  if (__builtin_fabsf(hi.x) >= __builtin_fabsf(hi.y))
    return make_float2(__builtin_fmaf(v.x, hi.x, __builtin_fmaf(-v.y, hi.y, cr)),
                       __builtin_fmaf(v.y, hi.x, __builtin_fmaf(v.x, hi.y, ci)));
  return make_float2(__builtin_fmaf(-v.y, hi.y, __builtin_fmaf(v.x, hi.x, cr)),
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

// e^x for |x| < 2^-10 (fifth-order series, relative error below 1e-21), exp
// otherwise.
__device__ inline double exp_small(double x) {
  if (fabs(x) < 0x1p-10)
    return 1.0 + x * (1.0 + 0.5 * x * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x * (1.0 + 0.2 * x))));
  return exp(x);
}

// sin and cos in FP64 accuracy: Taylor polynomials for |x| < 2^-4 (next terms
// below 1e-21), the library sincos otherwise.  The split-operator phases are
// small angles, so the polynomial is the common path.
__device__ inline void sincos_small(double x, double *s, double *c) {
  if (fabs(x) < 0x1p-4) {
    const double x2 = x * x;
    *s = x * (1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0)))));
    *c = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 + x2 * (-1.0 / 3628800.0)))));
  } else {
    sincos(x, s, c);
  }
}

// psi e^{-i V dt/2} for complex V and complex dt, e^{a + i p}, with the factor
// evaluated in FP64.  In FP32 it is applied as its rounded value plus
// residual, so that its modulus is exact: in a bulk region V is nearly
// uniform, every point would see the same rounding of |e^{-iV dt/2}| twice
// per step, and an FP32-rounded factor drifts the norm by ~1e-8 per step.
template <typename T>
__device__ inline cplx<T> apply_vphase(cplx<T> psi, cplx<T> v, double dtr, double dti) {
  const double a = 0.5 * ((double)v.x * dti + (double)v.y * dtr);
  const double p = -0.5 * ((double)v.x * dtr - (double)v.y * dti);
  double s, c;
  sincos_small(p, &s, &c);
  const double m = exp_small(a), pr = m * c, pi = m * s;
  if constexpr (sizeof(T) == 4) {
    const float hr = (float)pr, hi = (float)pi;
    return cmul_hl_phase(psi, make_float2(hr, hi), make_float2((float)(pr - hr), (float)(pi - hi)));
  } else {
    return cmul(psi, cmk<cplx<T>>(pr, pi));
  }
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
  const void *k[8];    // real half-spectrum arrays [nx][ny][PP]
  const void *sinx, *siny, *sinz;  // sin(2 pi m / n) / h tables
};

}  // namespace fft
}  // namespace dft
