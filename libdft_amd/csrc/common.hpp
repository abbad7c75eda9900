// Shared device/host helpers for the MI355X (gfx950) OT hot path.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#define DFT_HIP_CHECK(expr)                                                                       \
  do {                                                                                            \
    hipError_t e__ = (expr);                                                                      \
    if (e__ != hipSuccess) {                                                                      \
      fprintf(stderr, "libdft: HIP error %s at %s:%d: %s\n", hipGetErrorName(e__), __FILE__,      \
              __LINE__, #expr);                                                                   \
      abort();                                                                                    \
    }                                                                                             \
  } while (0)

// Report a failed launch (bad configuration, too much LDS) where it happens.
#define DFT_LAUNCH_CHECK() DFT_HIP_CHECK(hipGetLastError())

namespace dft {
// Per-launch timing with HIP events on the launch stream (off unless enabled
// through dft_profile_enable).  `bytes` is the launch's algorithmic traffic:
// every input and output field counted once.
extern bool g_prof_on;
void prof_begin(hipStream_t s);
void prof_end(hipStream_t s, const char *name, double bytes);
}  // namespace dft

#define DFT_PROF_BEGIN(s) \
  if (dft::g_prof_on) dft::prof_begin(s)
#define DFT_PROF_END(s, name, bytes) \
  if (dft::g_prof_on) dft::prof_end(s, name, bytes)

namespace dft {

template <typename T> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename T> using cplx = typename cplx_of<T>::type;

template <typename C> __host__ __device__ inline C cmk(decltype(C::x) a, decltype(C::x) b) {
  C r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename C> __device__ inline C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ inline C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ inline C cmul(C a, C b) {
  return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ inline C cconj(C a) { return cmk<C>(a.x, -a.y); }
template <typename C, typename T> __device__ inline C cscale(C a, T s) { return cmk<C>(a.x * s, a.y * s); }

// Accurate exp / sincos in the field precision (no fast intrinsics).
__device__ inline float dexp(float x) { return expf(x); }
__device__ inline double dexp(double x) { return exp(x); }
__device__ inline void dsincos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ inline void dsincos(double x, double *s, double *c) { sincos(x, s, c); }

constexpr int kBlock = 256;

// Row pitch (complex elements) of a half spectrum [nx][ny][nz/2+1]: the row
// holds nz/2+1 values and is padded to a multiple of 16 (128 bytes in FP32),
// so that column tiles of the x and y passes read whole cache lines.
inline int half_pitch(int nz) { return ((nz / 2 + 1) + 15) & ~15; }

// Run BODY with `T` bound to the storage type of a grid precision.
#define DFT_BY_PREC(prec, ...)  \
  do {                          \
    if ((prec) == 1) {          \
      using T = float;          \
      __VA_ARGS__;              \
    } else {                    \
      using T = double;         \
      __VA_ARGS__;              \
    }                           \
  } while (0)

// Launch configuration for elementwise kernels: grid-stride loops over n items.
inline unsigned grid_for(int64_t n, int per_thread = 1) {
  int64_t b = (n + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
  if (b < 1) b = 1;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return (unsigned)b;
}

}  // namespace dft
