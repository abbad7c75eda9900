// Grid handles (libgrid-compatible subset), elementwise device kernels,
// FP64 reductions and the split-operator propagator API.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace dft;

namespace dft {

static hipStream_t g_stream = nullptr;
static int g_prec = GRID_FP64;
static std::mutex g_mu;

hipStream_t lib_stream() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_stream) DFT_HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  return g_stream;
}

void *dmalloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    fprintf(stderr, "libdft: hipMalloc(%zu) failed: %s\n", bytes, hipGetErrorString(e));
    return nullptr;
  }
  return p;
}

void dfree(void *p) {
  if (p) DFT_HIP_CHECK(hipFree(p));
}

// ------------------------------------------------------------- profiling --
bool g_prof_on = false;
struct ProfRec {
  const char *name;
  hipEvent_t a, b;
  double bytes;
};
struct ProfAgg {
  std::string name;
  int64_t launches;
  double ms, bytes;
};
static std::vector<ProfRec> g_prof;
static std::vector<ProfAgg> g_agg;
static std::vector<hipEvent_t> g_evpool;
static size_t g_evnext = 0;
static hipEvent_t g_pending = nullptr;

static hipEvent_t ev_get() {
  if (g_evnext == g_evpool.size()) {
    hipEvent_t e;
    DFT_HIP_CHECK(hipEventCreate(&e));
    g_evpool.push_back(e);
  }
  return g_evpool[g_evnext++];
}

void prof_begin(hipStream_t s) {
  g_pending = ev_get();
  DFT_HIP_CHECK(hipEventRecord(g_pending, s));
}

void prof_end(hipStream_t s, const char *name, double bytes) {
  hipEvent_t b = ev_get();
  DFT_HIP_CHECK(hipEventRecord(b, s));
  g_prof.push_back({name, g_pending, b, bytes});
}

// --------------------------------------------------------------- kernels --
#define GRID_STRIDE(i, n) for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

template <typename T> __global__ void k_fill(T *a, T v, int64_t n) { GRID_STRIDE(i, n) a[i] = v; }
template <typename T> __global__ void k_fillc(cplx<T> *a, T re, T im, int64_t n) {
  GRID_STRIDE(i, n) a[i] = cmk<cplx<T>>(re, im);
}
template <typename T> __global__ void k_addc(cplx<T> *a, T re, T im, int64_t n) {
  GRID_STRIDE(i, n) a[i] = cmk<cplx<T>>(a[i].x + re, a[i].y + im);
}
template <typename T> __global__ void k_mulc(cplx<T> *a, T re, T im, int64_t n) {
  GRID_STRIDE(i, n) a[i] = cmul(a[i], cmk<cplx<T>>(re, im));
}
template <typename T> __global__ void k_add(T *a, T c, int64_t n) { GRID_STRIDE(i, n) a[i] += c; }
template <typename T> __global__ void k_mul(T *a, T c, int64_t n) { GRID_STRIDE(i, n) a[i] *= c; }
template <typename T> __global__ void k_add_scaled_product(T *d, T s, const T *a, const T *b, int64_t n) {
  GRID_STRIDE(i, n) d[i] += s * a[i] * b[i];
}
template <typename T> __global__ void k_r2c(cplx<T> *d, const T *r, int add, int imag, int64_t n) {
  GRID_STRIDE(i, n) {
    cplx<T> v = add ? d[i] : cmk<cplx<T>>(0, 0);
    if (imag) v.y += r[i];
    else v.x += r[i];
    d[i] = v;
  }
}
template <typename T> __global__ void k_density(const cplx<T> *psi, T *rho, int64_t n) {
  GRID_STRIDE(i, n) rho[i] = psi[i].x * psi[i].x + psi[i].y * psi[i].y;
}
template <typename Td, typename Ts> __global__ void k_convert(Td *d, const Ts *s, int64_t n) {
  GRID_STRIDE(i, n) d[i] = (Td)s[i];
}

// Periodic central difference along `dir` of a real field.
template <typename T>
__global__ void k_gradient(const T *f, T *g, int nx, int ny, int nz, int dir, T inv2h) {
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    int64_t p, m;
    if (dir == 0) {
      p = c + (int64_t)(i == nx - 1 ? -(nx - 1) : 1) * ny * nz;
      m = c + (int64_t)(i == 0 ? (nx - 1) : -1) * ny * nz;
    } else if (dir == 1) {
      p = c + (int64_t)(j == ny - 1 ? -(ny - 1) : 1) * nz;
      m = c + (int64_t)(j == 0 ? (ny - 1) : -1) * nz;
    } else {
      p = c + (k == nz - 1 ? -(nz - 1) : 1);
      m = c + (k == 0 ? (nz - 1) : -1);
    }
    g[c] = inv2h * (f[p] - f[m]);
  }
}

// v_d = Im(psi* (psi_{+d} - psi_{-d})/2h) / (M (|psi|^2 + eps))
template <typename T>
__global__ void k_velocity(const cplx<T> *psi, T *vx, T *vy, T *vz, int nx, int ny, int nz, T inv2h, T invM, T eps) {
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const cplx<T> s = psi[c];
    const T den = invM / (s.x * s.x + s.y * s.y + eps);
    T *out[3] = {vx, vy, vz};
    for (int d = 0; d < 3; ++d) {
      if (!out[d]) continue;
      int64_t p, m;
      if (d == 0) {
        p = c + (int64_t)(i == nx - 1 ? -(nx - 1) : 1) * ny * nz;
        m = c + (int64_t)(i == 0 ? (nx - 1) : -1) * ny * nz;
      } else if (d == 1) {
        p = c + (int64_t)(j == ny - 1 ? -(ny - 1) : 1) * nz;
        m = c + (int64_t)(j == 0 ? (ny - 1) : -1) * nz;
      } else {
        p = c + (k == nz - 1 ? -(nz - 1) : 1);
        m = c + (k == 0 ? (nz - 1) : -1);
      }
      const T dx = (psi[p].x - psi[m].x) * inv2h, dy = (psi[p].y - psi[m].y) * inv2h;
      out[d][c] = (s.x * dy - s.y * dx) * den;
    }
  }
}

// ----------------------------------------------------------- reductions --
constexpr int kRedBlocks = 1024;

int64_t red_blocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks;
}

__device__ inline void block_sum_store(double acc, double *partial) {
  __shared__ double s[kBlock];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// sum a (b == null) or sum a*b
template <typename T> __global__ void k_red_sum(const T *a, const T *b, int64_t n, double *partial) {
  double acc = 0.0;
  GRID_STRIDE(i, n) acc += b ? (double)a[i] * (double)b[i] : (double)a[i];
  block_sum_store(acc, partial);
}
template <typename T> __global__ void k_red_abs2(const cplx<T> *a, int64_t n, double *partial) {
  double acc = 0.0;
  GRID_STRIDE(i, n) acc += (double)a[i].x * (double)a[i].x + (double)a[i].y * (double)a[i].y;
  block_sum_store(acc, partial);
}
// sum_k K(k) |spec(k)|^2 over a full complex spectrum, K = cx mx^2 + cy my^2 + cz mz^2.
template <typename T>
__global__ void k_red_kinetic(const cplx<T> *a, int nx, int ny, int nz, double cx, double cy, double cz,
                              double ox, double oy, double oz, double *partial) {
  double acc = 0.0;
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const double mx = (double)(i < nx / 2 ? i : i - nx) - ox, my = (double)(j < ny / 2 ? j : j - ny) - oy,
                 mz = (double)(k < nz / 2 ? k : k - nz) - oz;
    const double K = cx * mx * mx + cy * my * my + cz * mz * mz;
    acc += K * ((double)a[c].x * (double)a[c].x + (double)a[c].y * (double)a[c].y);
  }
  block_sum_store(acc, partial);
}

double sum_partials(const double *dpartials, int64_t n, hipStream_t s) {
  std::vector<double> h(n);
  DFT_HIP_CHECK(hipMemcpyAsync(h.data(), dpartials, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  DFT_HIP_CHECK(hipStreamSynchronize(s));
  // Neumaier compensated summation
  double sum = 0.0, c = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double t = sum + h[i];
    if (std::fabs(sum) >= std::fabs(h[i])) c += (sum - t) + h[i];
    else c += (h[i] - t) + sum;
    sum = t;
  }
  return sum + c;
}

template <typename T> void reduce_sum(const T *a, const T *b, int64_t n, double *partial, hipStream_t s) {
  hipLaunchKernelGGL(k_red_sum<T>, dim3((unsigned)red_blocks(n)), dim3(kBlock), 0, s, a, b, n, partial);
  DFT_LAUNCH_CHECK();
}
template <typename T> void reduce_abs2(const cplx<T> *a, int64_t n, double *partial, hipStream_t s) {
  hipLaunchKernelGGL(k_red_abs2<T>, dim3((unsigned)red_blocks(n)), dim3(kBlock), 0, s, a, n, partial);
  DFT_LAUNCH_CHECK();
}
template <typename T>
void reduce_kinetic(const cplx<T> *spec, int nx, int ny, int nz, double cx, double cy, double cz, double *partial,
                    hipStream_t s, const double *o) {
  hipLaunchKernelGGL(k_red_kinetic<T>, dim3((unsigned)red_blocks((int64_t)nx * ny * nz)), dim3(kBlock), 0, s, spec,
                     nx, ny, nz, cx, cy, cz, o ? o[0] : 0.0, o ? o[1] : 0.0, o ? o[2] : 0.0, partial);
  DFT_LAUNCH_CHECK();
}
template <typename T> __global__ void k_red_density_re(const cplx<T> *a, const cplx<T> *v, int64_t n, double *partial) {
  double acc = 0.0;
  GRID_STRIDE(i, n) acc += ((double)a[i].x * (double)a[i].x + (double)a[i].y * (double)a[i].y) * (double)v[i].x;
  block_sum_store(acc, partial);
}
template <typename T>
void reduce_density_re(const cplx<T> *a, const cplx<T> *v, int64_t n, double *partial, hipStream_t s) {
  hipLaunchKernelGGL(k_red_density_re<T>, dim3((unsigned)red_blocks(n)), dim3(kBlock), 0, s, a, v, n, partial);
  DFT_LAUNCH_CHECK();
}
template void reduce_density_re<float>(const float2 *, const float2 *, int64_t, double *, hipStream_t);
template void reduce_density_re<double>(const double2 *, const double2 *, int64_t, double *, hipStream_t);

template <typename T> void ew_density(const cplx<T> *psi, T *rho, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_density<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, psi, rho, n);
  DFT_LAUNCH_CHECK();
}

template void reduce_sum<float>(const float *, const float *, int64_t, double *, hipStream_t);
template void reduce_sum<double>(const double *, const double *, int64_t, double *, hipStream_t);
template void reduce_abs2<float>(const float2 *, int64_t, double *, hipStream_t);
template void reduce_abs2<double>(const double2 *, int64_t, double *, hipStream_t);
template void reduce_kinetic<float>(const float2 *, int, int, int, double, double, double, double *, hipStream_t,
                                    const double *);
template void reduce_kinetic<double>(const double2 *, int, int, int, double, double, double, double *, hipStream_t,
                                     const double *);
template void ew_density<float>(const float2 *, float *, int64_t, hipStream_t);
template void ew_density<double>(const double2 *, double *, int64_t, hipStream_t);

// dst = sum_d (f_{+d} - 2 f + f_{-d}) / h^2, periodic (complex field)
template <typename T> __global__ void k_fd_laplace(const cplx<T> *f, cplx<T> *dst, int nx, int ny, int nz, T inv_h2) {
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const int64_t sx = (int64_t)ny * nz;
    const int64_t xp = c + (i == nx - 1 ? -(int64_t)(nx - 1) * sx : sx), xm = c + (i == 0 ? (int64_t)(nx - 1) * sx : -sx);
    const int64_t yp = c + (j == ny - 1 ? -(int64_t)(ny - 1) * nz : nz), ym = c + (j == 0 ? (int64_t)(ny - 1) * nz : -(int64_t)nz);
    const int64_t zp = c + (k == nz - 1 ? -(int64_t)(nz - 1) : 1), zm = c + (k == 0 ? (int64_t)(nz - 1) : -1);
    const cplx<T> v = f[c];
    const T re = (f[xp].x + f[xm].x) + (f[yp].x + f[ym].x) + (f[zp].x + f[zm].x) - (T)6 * v.x;
    const T im = (f[xp].y + f[xm].y) + (f[yp].y + f[ym].y) + (f[zp].y + f[zm].y) - (T)6 * v.y;
    dst[c] = cmk<cplx<T>>(re * inv_h2, im * inv_h2);
  }
}

// |curl (fx, fy, fz)| with periodic central differences
template <typename T>
__global__ void k_abs_rot(T *rot, const T *fx, const T *fy, const T *fz, int nx, int ny, int nz, T inv2h) {
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const int64_t sx = (int64_t)ny * nz;
    const int64_t xp = c + (i == nx - 1 ? -(int64_t)(nx - 1) * sx : sx), xm = c + (i == 0 ? (int64_t)(nx - 1) * sx : -sx);
    const int64_t yp = c + (j == ny - 1 ? -(int64_t)(ny - 1) * nz : nz), ym = c + (j == 0 ? (int64_t)(ny - 1) * nz : -(int64_t)nz);
    const int64_t zp = c + (k == nz - 1 ? -(int64_t)(nz - 1) : 1), zm = c + (k == 0 ? (int64_t)(nz - 1) : -1);
    const T cx = (fz[yp] - fz[ym] - (fy[zp] - fy[zm])) * inv2h;
    const T cy = (fx[zp] - fx[zm] - (fz[xp] - fz[xm])) * inv2h;
    const T cz = (fy[xp] - fy[xm] - (fx[yp] - fx[ym])) * inv2h;
    rot[c] = sqrt(cx * cx + cy * cy + cz * cz);
  }
}

// dst = |src|^p (real)
template <typename T> __global__ void k_rpower(T *dst, const T *src, T p, int64_t n) {
  GRID_STRIDE(i, n) dst[i] = pow(src[i], p);
}
template <typename T> __global__ void k_rproduct(T *dst, const T *a, const T *b, int64_t n) {
  GRID_STRIDE(i, n) dst[i] = a[i] * b[i];
}
template <typename T> __global__ void k_rdivision_eps(T *dst, const T *a, const T *b, T eps, int64_t n) {
  GRID_STRIDE(i, n) dst[i] = a[i] / (b[i] + eps);
}
// per-block maximum of sgn * a (sgn = -1: minus the minimum)
template <typename T> __global__ void k_red_max(const T *a, T sgn, int64_t n, double *partial) {
  __shared__ double s[kBlock];
  double m = -INFINITY;
  GRID_STRIDE(i, n) m = fmax(m, (double)(sgn * a[i]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}
template <typename T> __global__ void k_rabs_power(T *dst, const T *src, T p, int64_t n) {
  GRID_STRIDE(c, n) dst[c] = pow(fabs(src[c]), p);
}

// j_d = Im(psi* (psi_{+d} - psi_{-d}) / 2h) / M (probability flux)
template <typename T>
__global__ void k_flux(const cplx<T> *psi, T *jx, T *jy, T *jz, int nx, int ny, int nz, T inv2h, T invM) {
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const int64_t sx = (int64_t)ny * nz;
    const int64_t p[3] = {c + (i == nx - 1 ? -(int64_t)(nx - 1) * sx : sx), c + (j == ny - 1 ? -(int64_t)(ny - 1) * nz : nz),
                          c + (k == nz - 1 ? -(int64_t)(nz - 1) : 1)};
    const int64_t m[3] = {c + (i == 0 ? (int64_t)(nx - 1) * sx : -sx), c + (j == 0 ? (int64_t)(ny - 1) * nz : -(int64_t)nz),
                          c + (k == 0 ? (int64_t)(nz - 1) : -1)};
    T *out[3] = {jx, jy, jz};
    const cplx<T> s = psi[c];
    for (int d = 0; d < 3; ++d) {
      if (!out[d]) continue;
      const T dx = (psi[p[d]].x - psi[m[d]].x) * inv2h, dy = (psi[p[d]].y - psi[m[d]].y) * inv2h;
      out[d][c] = (s.x * dy - s.y * dx) * invM;
    }
  }
}

// dst = conj(a) b
template <typename T> __global__ void k_conj_product(cplx<T> *dst, const cplx<T> *a, const cplx<T> *b, int64_t n) {
  GRID_STRIDE(c, n) {
    const cplx<T> x = a[c], y = b[c];
    dst[c] = cmk<cplx<T>>(x.x * y.x + x.y * y.y, x.x * y.y - x.y * y.x);
  }
}

// dst = |src|^p
template <typename T> __global__ void k_abs_power(cplx<T> *dst, const cplx<T> *src, T p, int64_t n) {
  GRID_STRIDE(c, n) {
    const cplx<T> x = src[c];
    dst[c] = cmk<cplx<T>>(pow(sqrt(x.x * x.x + x.y * x.y), p), (T)0);
  }
}

// Index box of a region sum: i in [b[0], b[1]], j in [b[2], b[3]], k in [b[4], b[5]].
struct Box {
  int b[6];
};

// sum of a real field (comp = 1) or of the real and imaginary parts of a
// complex field (comp = 2, the second sum into partial + gridDim.x) over a box
template <typename T>
__global__ void k_red_region(const T *a, int comp, int ny, int nz, Box bx, int64_t n, double *partial) {
  double acc0 = 0.0, acc1 = 0.0;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    if (i < bx.b[0] || i > bx.b[1] || j < bx.b[2] || j > bx.b[3] || k < bx.b[4] || k > bx.b[5]) continue;
    acc0 += (double)a[comp * c];
    if (comp == 2) acc1 += (double)a[2 * c + 1];
  }
  block_sum_store(acc0, partial);
  if (comp == 2) block_sum_store(acc1, partial + gridDim.x);
}

// sums of Re(conj(a) b) and Im(conj(a) b) (second into partial + gridDim.x)
template <typename T>
__global__ void k_red_conj_product(const cplx<T> *a, const cplx<T> *b, int64_t n, double *partial) {
  double re = 0.0, im = 0.0;
  GRID_STRIDE(c, n) {
    const double ar = a[c].x, ai = a[c].y, br = b[c].x, bi = b[c].y;
    re += ar * br + ai * bi;
    im += ar * bi - ai * br;
  }
  block_sum_store(re, partial);
  block_sum_store(im, partial + gridDim.x);
}

}  // namespace dft

// Scratch for FP64 partial sums shared by the library (stream-ordered use only).
static double *g_partials = nullptr;
static int64_t g_npartials = 0;
double *dft_scratch_partials(int64_t n) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (n > g_npartials) {
    if (g_partials) {
      DFT_HIP_CHECK(hipDeviceSynchronize());
      dfree(g_partials);
    }
    g_npartials = n < 4096 ? 4096 : n;
    g_partials = (double *)dmalloc(sizeof(double) * g_npartials);
  }
  return g_partials;
}

// ================================================================== C ABI ==

namespace {

int64_t npts(INT nx, INT ny, INT nz) { return (int64_t)nx * ny * nz; }
int64_t half_elems(INT nx, INT ny, INT nz) { return (int64_t)nx * ny * (nz / 2 + 1); }
int64_t half_storage(INT nx, INT ny, INT nz) { return (int64_t)nx * ny * dft::half_pitch((int)nz); }

bool dims_ok(INT nx, INT ny, INT nz) {
  return fft_size_ok(nx) && fft_size_ok(ny) && fft_size_ok(nz);
}

size_t esize(int prec) { return prec == GRID_FP32 ? sizeof(float) : sizeof(double); }

INT wrap_index(INT i, INT n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// false (with a libdft: message) when two grids differ in size or precision;
// the calling entry point then returns without doing anything
bool need_same(const char *fn, int64_t a, int64_t b, int pa, int pb) {
  if (a != b || pa != pb) {
    fprintf(stderr, "libdft: %s: grid shape or precision mismatch\n", fn);
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

// Periodic boundary: the value at the index wrapped into the grid.
REAL rgrid_value_outside_periodic(rgrid *g, INT i, INT j, INT k) {
  return rgrid_value_at_index(g, wrap_index(i, g->nx), wrap_index(j, g->ny), wrap_index(k, g->nz));
}
REAL_COMPLEX cgrid_value_outside_periodic(cgrid *g, INT i, INT j, INT k) {
  return cgrid_value_at_index(g, wrap_index(i, g->nx), wrap_index(j, g->ny), wrap_index(k, g->nz));
}

void grid_set_precision(int prec) {
  if (prec != GRID_FP32 && prec != GRID_FP64) {
    fprintf(stderr, "libdft: grid_set_precision: unknown precision %d\n", prec);
    return;
  }
  g_prec = prec;
}
int grid_get_precision(void) { return g_prec; }
void grid_threads_init(INT) {}
void grid_set_fftw_flags(char) {}
void grid_fft_read_wisdom(char *) {}
void grid_fft_write_wisdom(char *) {}
void cuda_enable(char, int, int *) {}
void grid_synchronize(void) { DFT_HIP_CHECK(hipStreamSynchronize(lib_stream())); }
void *grid_get_stream(void) { return (void *)lib_stream(); }

rgrid *rgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL (*value_outside)(rgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id) {
  if (!dims_ok(nx, ny, nz)) {
    fprintf(stderr, "libdft: rgrid_alloc: dimensions must be powers of two in [4, 512] (got %ld %ld %ld)\n",
            (long)nx, (long)ny, (long)nz);
    return nullptr;
  }
  rgrid *g = (rgrid *)calloc(1, sizeof(rgrid));
  if (!g) return nullptr;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->step = step;
  g->prec = g_prec;
  // room for a real field or a real half spectrum
  g->dev_elems = npts(nx, ny, nz) > half_storage(nx, ny, nz) ? npts(nx, ny, nz) : half_storage(nx, ny, nz);
  g->dev = dmalloc(esize(g->prec) * g->dev_elems);
  if (!g->dev) {
    free(g);
    return nullptr;
  }
  DFT_HIP_CHECK(hipMemsetAsync(g->dev, 0, esize(g->prec) * g->dev_elems, lib_stream()));
  if (id) strncpy(g->id, id, sizeof(g->id) - 1);
  g->value_outside = value_outside ? value_outside : rgrid_value_outside_periodic;
  g->outside_params_ptr = outside_params_ptr;
  return g;
}

cgrid *cgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL_COMPLEX (*value_outside)(cgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id) {
  if (!dims_ok(nx, ny, nz)) {
    fprintf(stderr, "libdft: cgrid_alloc: dimensions must be powers of two in [4, 512] (got %ld %ld %ld)\n",
            (long)nx, (long)ny, (long)nz);
    return nullptr;
  }
  cgrid *g = (cgrid *)calloc(1, sizeof(cgrid));
  if (!g) return nullptr;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->step = step;
  g->prec = g_prec;
  g->dev_elems = npts(nx, ny, nz);
  g->dev = dmalloc(2 * esize(g->prec) * g->dev_elems);
  if (!g->dev) {
    free(g);
    return nullptr;
  }
  DFT_HIP_CHECK(hipMemsetAsync(g->dev, 0, 2 * esize(g->prec) * g->dev_elems, lib_stream()));
  if (id) strncpy(g->id, id, sizeof(g->id) - 1);
  g->value_outside = value_outside ? value_outside : cgrid_value_outside_periodic;
  g->outside_params_ptr = outside_params_ptr;
  return g;
}

rgrid *rgrid_clone(rgrid *src, char *id) {
  int save = g_prec;
  g_prec = src->prec;
  rgrid *g = rgrid_alloc(src->nx, src->ny, src->nz, src->step, src->value_outside, src->outside_params_ptr, id);
  g_prec = save;
  if (g) rgrid_set_origin(g, src->x0, src->y0, src->z0);
  return g;
}

cgrid *cgrid_clone(cgrid *src, char *id) {
  int save = g_prec;
  g_prec = src->prec;
  cgrid *g = cgrid_alloc(src->nx, src->ny, src->nz, src->step, src->value_outside, src->outside_params_ptr, id);
  g_prec = save;
  if (g) {
    cgrid_set_origin(g, src->x0, src->y0, src->z0);
    cgrid_set_momentum(g, src->kx0, src->ky0, src->kz0);
  }
  return g;
}

void rgrid_free(rgrid *g) {
  if (!g) return;
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  dfree(g->dev);
  free(g);
}
void cgrid_free(cgrid *g) {
  if (!g) return;
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  dfree(g->dev);
  free(g);
}
void rgrid_set_origin(rgrid *g, REAL x0, REAL y0, REAL z0) {
  g->x0 = x0;
  g->y0 = y0;
  g->z0 = z0;
}
void cgrid_set_origin(cgrid *g, REAL x0, REAL y0, REAL z0) {
  g->x0 = x0;
  g->y0 = y0;
  g->z0 = z0;
}
void rgrid_claim(rgrid *g) {
  if (g->flag) {
    fprintf(stderr, "libdft: grid %s already claimed\n", g->id);
    abort();
  }
  g->flag = 1;
}
void rgrid_release(rgrid *g) {
  if (!g->flag) {
    fprintf(stderr, "libdft: grid %s released without claim\n", g->id);
    abort();
  }
  g->flag = 0;
}

wf *grid_wf_alloc(INT nx, INT ny, INT nz, REAL step, REAL mass, char boundary, char propagator, char *id) {
  if (boundary != WF_PERIODIC_BOUNDARY || propagator != WF_2ND_ORDER_FFT) {
    fprintf(stderr, "libdft: grid_wf_alloc: only periodic boundaries with the 2nd order FFT propagator\n");
    return nullptr;
  }
  wf *w = (wf *)calloc(1, sizeof(wf));
  if (!w) return nullptr;
  w->grid = cgrid_alloc(nx, ny, nz, step, CGRID_PERIODIC_BOUNDARY, nullptr, id);
  if (!w->grid) {
    free(w);
    return nullptr;
  }
  w->mass = mass;
  w->norm = 1.0;
  w->boundary = boundary;
  w->propagator = propagator;
  return w;
}

wf *grid_wf_clone(wf *src, char *id) {
  wf *w = (wf *)calloc(1, sizeof(wf));
  if (!w) return nullptr;
  *w = *src;
  w->cworkspace = nullptr;
  w->grid = cgrid_clone(src->grid, id);
  if (!w->grid) {
    free(w);
    return nullptr;
  }
  return w;
}

void grid_wf_free(wf *w) {
  if (!w) return;
  cgrid_free(w->grid);
  if (w->cworkspace) cgrid_free(w->cworkspace);
  free(w);
}

// ------------------------------------------------------------ transfers --
// For a spectral grid (a functional's kernel field) this copies a real half
// spectrum, nx*ny*(nz/2+1) values, into rows stored with a padded pitch.
void rgrid_copy_from_host(rgrid *g, const REAL *src) {
  hipStream_t s = lib_stream();
  const size_t es = esize(g->prec);
  const int64_t rows = (int64_t)g->nx * g->ny, P = g->nz / 2 + 1, PP = dft::half_pitch((int)g->nz);
  const int64_t n = g->spectral ? rows * P : npts(g->nx, g->ny, g->nz);
  if ((g->spectral ? rows * PP : n) > g->dev_elems) {
    fprintf(stderr, "libdft: rgrid_copy_from_host: %lld values do not fit grid %s\n", (long long)n, g->id);
    return;
  }
  std::vector<float> t;
  const void *host = src;
  if (g->prec != GRID_FP64) {
    t.resize(n);
    for (int64_t i = 0; i < n; ++i) t[i] = (float)src[i];
    host = t.data();
  }
  if (g->spectral)
    DFT_HIP_CHECK(hipMemcpy2DAsync(g->dev, es * PP, host, es * P, es * P, rows, hipMemcpyHostToDevice, s));
  else
    DFT_HIP_CHECK(hipMemcpyAsync(g->dev, host, es * n, hipMemcpyHostToDevice, s));
  DFT_HIP_CHECK(hipStreamSynchronize(s));
}

// For a spectral grid (a functional's kernel field) this copies the real half
// spectrum, nx*ny*(nz/2+1) values (rows are stored with a padded pitch).
void rgrid_copy_to_host(const rgrid *g, REAL *dst) {
  hipStream_t s = lib_stream();
  const size_t es = esize(g->prec);
  int64_t n;
  std::vector<char> t;
  void *host = dst;
  if (g->prec != GRID_FP64) {
    t.resize(es * (g->spectral ? half_elems(g->nx, g->ny, g->nz) : npts(g->nx, g->ny, g->nz)));
    host = t.data();
  }
  if (g->spectral) {
    const int64_t rows = (int64_t)g->nx * g->ny, P = g->nz / 2 + 1, PP = dft::half_pitch((int)g->nz);
    n = rows * P;
    DFT_HIP_CHECK(hipMemcpy2DAsync(host, es * P, g->dev, es * PP, es * P, rows, hipMemcpyDeviceToHost, s));
  } else {
    n = npts(g->nx, g->ny, g->nz);
    DFT_HIP_CHECK(hipMemcpyAsync(host, g->dev, es * n, hipMemcpyDeviceToHost, s));
  }
  DFT_HIP_CHECK(hipStreamSynchronize(s));
  if (g->prec != GRID_FP64) {
    const float *f = (const float *)t.data();
    for (int64_t i = 0; i < n; ++i) dst[i] = f[i];
  }
}

void cgrid_copy_from_host(cgrid *g, const REAL_COMPLEX *src) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  hipStream_t s = lib_stream();
  if (g->prec == GRID_FP64) {
    DFT_HIP_CHECK(hipMemcpyAsync(g->dev, src, 2 * sizeof(double) * n, hipMemcpyHostToDevice, s));
  } else {
    const double *d = (const double *)src;
    std::vector<float> t(2 * n);
    for (int64_t i = 0; i < 2 * n; ++i) t[i] = (float)d[i];
    DFT_HIP_CHECK(hipMemcpyAsync(g->dev, t.data(), 2 * sizeof(float) * n, hipMemcpyHostToDevice, s));
  }
  DFT_HIP_CHECK(hipStreamSynchronize(s));
}

void cgrid_copy_to_host(const cgrid *g, REAL_COMPLEX *dst) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  hipStream_t s = lib_stream();
  if (g->prec == GRID_FP64) {
    DFT_HIP_CHECK(hipMemcpyAsync(dst, g->dev, 2 * sizeof(double) * n, hipMemcpyDeviceToHost, s));
    DFT_HIP_CHECK(hipStreamSynchronize(s));
  } else {
    std::vector<float> t(2 * n);
    DFT_HIP_CHECK(hipMemcpyAsync(t.data(), g->dev, 2 * sizeof(float) * n, hipMemcpyDeviceToHost, s));
    DFT_HIP_CHECK(hipStreamSynchronize(s));
    double *d = (double *)dst;
    for (int64_t i = 0; i < 2 * n; ++i) d[i] = t[i];
  }
}

// An index outside the grid gives the grid's boundary value (periodic: the
// wrapped index).  For a spectral grid (i, j, k) addresses the real half
// spectrum, k < nz/2 + 1.
REAL rgrid_value_at_index(rgrid *g, INT i, INT j, INT k) {
  const INT nk = g->spectral ? g->nz / 2 + 1 : g->nz;
  if (i < 0 || i >= g->nx || j < 0 || j >= g->ny || k < 0 || k >= nk) {
    if (g->spectral || !g->value_outside) return rgrid_value_at_index(g, wrap_index(i, g->nx), wrap_index(j, g->ny), wrap_index(k, nk));
    return g->value_outside(g, i, j, k);
  }
  const int64_t pitch = g->spectral ? dft::half_pitch((int)g->nz) : g->nz;
  const int64_t c = (i * g->ny + j) * pitch + k;
  hipStream_t s = lib_stream();
  double v = 0;
  if (g->prec == GRID_FP64) {
    DFT_HIP_CHECK(hipMemcpyAsync(&v, (double *)g->dev + c, sizeof(double), hipMemcpyDeviceToHost, s));
    DFT_HIP_CHECK(hipStreamSynchronize(s));
  } else {
    float f = 0;
    DFT_HIP_CHECK(hipMemcpyAsync(&f, (float *)g->dev + c, sizeof(float), hipMemcpyDeviceToHost, s));
    DFT_HIP_CHECK(hipStreamSynchronize(s));
    v = f;
  }
  return v;
}

REAL_COMPLEX cgrid_value_at_index(cgrid *g, INT i, INT j, INT k) {
  if (i < 0 || i >= g->nx || j < 0 || j >= g->ny || k < 0 || k >= g->nz) {
    if (!g->value_outside) return cgrid_value_at_index(g, wrap_index(i, g->nx), wrap_index(j, g->ny), wrap_index(k, g->nz));
    return g->value_outside(g, i, j, k);
  }
  const int64_t c = (i * g->ny + j) * g->nz + k;
  hipStream_t s = lib_stream();
  double v[2] = {0, 0};
  if (g->prec == GRID_FP64) {
    DFT_HIP_CHECK(hipMemcpyAsync(v, (double *)g->dev + 2 * c, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    DFT_HIP_CHECK(hipStreamSynchronize(s));
  } else {
    float f[2] = {0, 0};
    DFT_HIP_CHECK(hipMemcpyAsync(f, (float *)g->dev + 2 * c, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    DFT_HIP_CHECK(hipStreamSynchronize(s));
    v[0] = f[0];
    v[1] = f[1];
  }
  REAL_COMPLEX r;
  __real__ r = v[0];
  __imag__ r = v[1];
  return r;
}

void rgrid_map(rgrid *g, REAL (*func)(void *, REAL, REAL, REAL), void *farg) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  std::vector<double> h(n);
  for (INT i = 0; i < g->nx; ++i)
    for (INT j = 0; j < g->ny; ++j)
      for (INT k = 0; k < g->nz; ++k)
        h[(i * g->ny + j) * g->nz + k] = func(farg, (double)(i - g->nx / 2) * g->step - g->x0,
                                              (double)(j - g->ny / 2) * g->step - g->y0,
                                              (double)(k - g->nz / 2) * g->step - g->z0);
  rgrid_copy_from_host(g, h.data());
}

void cgrid_map(cgrid *g, REAL_COMPLEX (*func)(void *, REAL, REAL, REAL), void *farg) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  std::vector<REAL_COMPLEX> h(n);
  for (INT i = 0; i < g->nx; ++i)
    for (INT j = 0; j < g->ny; ++j)
      for (INT k = 0; k < g->nz; ++k)
        h[(i * g->ny + j) * g->nz + k] = func(farg, (double)(i - g->nx / 2) * g->step - g->x0,
                                              (double)(j - g->ny / 2) * g->step - g->y0,
                                              (double)(k - g->nz / 2) * g->step - g->z0);
  cgrid_copy_from_host(g, h.data());
}

void grid_wf_map(wf *w, REAL_COMPLEX (*func)(void *, REAL, REAL, REAL), void *farg) { cgrid_map(w->grid, func, farg); }

// ------------------------------------------------------------ operations --
void rgrid_zero(rgrid *g) { rgrid_constant(g, 0.0); }
void cgrid_zero(cgrid *g) { cgrid_constant(g, 0.0); }

#define LAUNCH(kern, n, ...) \
  do {                                                                                    \
    hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(kBlock), 0, lib_stream(), __VA_ARGS__); \
    DFT_LAUNCH_CHECK();                                                                    \
  } while (0)

void rgrid_constant(rgrid *g, REAL c) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  DFT_BY_PREC(g->prec, LAUNCH(k_fill<T>, n, (T *)g->dev, (T)c, n));
}

void cgrid_constant(cgrid *g, REAL_COMPLEX c) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  DFT_BY_PREC(g->prec, LAUNCH(k_fillc<T>, n, (cplx<T> *)g->dev, (T)__real__ c, (T)__imag__ c, n));
}

void grid_wf_constant(wf *w, REAL_COMPLEX c) { cgrid_constant(w->grid, c); }

void rgrid_copy(rgrid *dst, rgrid *src) {
  if (!need_same("rgrid_copy", npts(dst->nx, dst->ny, dst->nz), npts(src->nx, src->ny, src->nz), dst->prec, src->prec)) return;
  DFT_HIP_CHECK(hipMemcpyAsync(dst->dev, src->dev, esize(src->prec) * npts(src->nx, src->ny, src->nz),
                               hipMemcpyDeviceToDevice, lib_stream()));
}

void cgrid_copy(cgrid *dst, cgrid *src) {
  if (!need_same("cgrid_copy", npts(dst->nx, dst->ny, dst->nz), npts(src->nx, src->ny, src->nz), dst->prec, src->prec)) return;
  DFT_HIP_CHECK(hipMemcpyAsync(dst->dev, src->dev, 2 * esize(src->prec) * npts(src->nx, src->ny, src->nz),
                               hipMemcpyDeviceToDevice, lib_stream()));
}

void cgrid_add(cgrid *g, REAL_COMPLEX c) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  DFT_BY_PREC(g->prec, LAUNCH(k_addc<T>, n, (cplx<T> *)g->dev, (T)__real__ c, (T)__imag__ c, n));
}

void cgrid_multiply(cgrid *g, REAL_COMPLEX c) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  DFT_BY_PREC(g->prec, LAUNCH(k_mulc<T>, n, (cplx<T> *)g->dev, (T)__real__ c, (T)__imag__ c, n));
}

void rgrid_add(rgrid *g, REAL c) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  DFT_BY_PREC(g->prec, LAUNCH(k_add<T>, n, (T *)g->dev, (T)c, n));
}

void rgrid_multiply(rgrid *g, REAL c) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  DFT_BY_PREC(g->prec, LAUNCH(k_mul<T>, n, (T *)g->dev, (T)c, n));
}

void rgrid_add_scaled_product(rgrid *d, REAL s, rgrid *a, rgrid *b) {
  const int64_t n = npts(d->nx, d->ny, d->nz);
  if (!need_same("rgrid_add_scaled_product", n, npts(a->nx, a->ny, a->nz), d->prec, a->prec)) return;
  if (!need_same("rgrid_add_scaled_product", n, npts(b->nx, b->ny, b->nz), d->prec, b->prec)) return;
  DFT_BY_PREC(d->prec, LAUNCH(k_add_scaled_product<T>, n, (T *)d->dev, (T)s, (const T *)a->dev, (const T *)b->dev, n));
}

static void r2c_op(cgrid *dst, rgrid *src, int add, int imag, const char *fn) {
  const int64_t n = npts(dst->nx, dst->ny, dst->nz);
  if (!need_same(fn, n, npts(src->nx, src->ny, src->nz), dst->prec, src->prec)) return;
  DFT_BY_PREC(dst->prec, LAUNCH(k_r2c<T>, n, (cplx<T> *)dst->dev, (const T *)src->dev, add, imag, n));
}
void grid_real_to_complex_re(cgrid *dst, rgrid *src) { r2c_op(dst, src, 0, 0, "grid_real_to_complex_re"); }
void grid_add_real_to_complex_re(cgrid *dst, rgrid *src) { r2c_op(dst, src, 1, 0, "grid_add_real_to_complex_re"); }
void grid_add_real_to_complex_im(cgrid *dst, rgrid *src) { r2c_op(dst, src, 1, 1, "grid_add_real_to_complex_im"); }

void grid_wf_density(wf *w, rgrid *rho) {
  cgrid *g = w->grid;
  const int64_t n = npts(g->nx, g->ny, g->nz);
  if (!need_same("grid_wf_density", n, npts(rho->nx, rho->ny, rho->nz), g->prec, rho->prec)) return;
  DFT_BY_PREC(g->prec, ew_density<T>((const cplx<T> *)g->dev, (T *)rho->dev, n, lib_stream()));
}

void grid_wf_velocity(wf *w, rgrid *vx, rgrid *vy, rgrid *vz, REAL eps) {
  cgrid *g = w->grid;
  const int64_t n = npts(g->nx, g->ny, g->nz);
  for (rgrid *v : {vx, vy, vz})
    if (v && !need_same("grid_wf_velocity", n, npts(v->nx, v->ny, v->nz), g->prec, v->prec)) return;
  DFT_BY_PREC(g->prec, LAUNCH(k_velocity<T>, n, (const cplx<T> *)g->dev, vx ? (T *)vx->dev : nullptr,
                              vy ? (T *)vy->dev : nullptr, vz ? (T *)vz->dev : nullptr, (int)g->nx, (int)g->ny,
                              (int)g->nz, (T)(0.5 / g->step), (T)(1.0 / w->mass), (T)eps));
}

static void gradient_op(rgrid *src, rgrid *dst, int dir) {
  const int64_t n = npts(src->nx, src->ny, src->nz);
  if (!need_same("rgrid_gradient", n, npts(dst->nx, dst->ny, dst->nz), src->prec, dst->prec)) return;
  DFT_BY_PREC(src->prec, LAUNCH(k_gradient<T>, n, (const T *)src->dev, (T *)dst->dev, (int)src->nx, (int)src->ny,
                                (int)src->nz, dir, (T)(0.5 / src->step)));
}
void rgrid_gradient_x(rgrid *src, rgrid *dst) { gradient_op(src, dst, 0); }
void rgrid_gradient_y(rgrid *src, rgrid *dst) { gradient_op(src, dst, 1); }
void rgrid_gradient_z(rgrid *src, rgrid *dst) { gradient_op(src, dst, 2); }

// ----------------------------------------------------------- reductions --
REAL rgrid_integral(rgrid *g) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  double *part = dft_scratch_partials(red_blocks(n));
  DFT_BY_PREC(g->prec, reduce_sum<T>((const T *)g->dev, nullptr, n, part, lib_stream()));
  return sum_partials(part, red_blocks(n), lib_stream()) * g->step * g->step * g->step;
}

// h^3 sum a b (libgrid rgrid_integral_of_product)
REAL rgrid_integral_of_product(rgrid *a, rgrid *b) {
  const int64_t n = npts(a->nx, a->ny, a->nz);
  if (!need_same("rgrid_integral_of_product", n, npts(b->nx, b->ny, b->nz), a->prec, b->prec)) return NAN;
  double *part = dft_scratch_partials(red_blocks(n));
  DFT_BY_PREC(a->prec, reduce_sum<T>((const T *)a->dev, (const T *)b->dev, n, part, lib_stream()));
  return sum_partials(part, red_blocks(n), lib_stream()) * a->step * a->step * a->step;
}

// h^3 sum |psi|^2 Re V (libgrid grid_wf_potential_energy)
REAL grid_wf_potential_energy(wf *w, cgrid *potential) {
  cgrid *g = w->grid;
  const int64_t n = npts(g->nx, g->ny, g->nz);
  if (!need_same("grid_wf_potential_energy", n, npts(potential->nx, potential->ny, potential->nz), g->prec,
            potential->prec)) return NAN;
  double *part = dft_scratch_partials(red_blocks(n));
  DFT_BY_PREC(g->prec, reduce_density_re<T>((const cplx<T> *)g->dev, (const cplx<T> *)potential->dev, n, part,
                                            lib_stream()));
  return sum_partials(part, red_blocks(n), lib_stream()) * g->step * g->step * g->step;
}

REAL grid_wf_norm(wf *w) {
  cgrid *g = w->grid;
  const int64_t n = npts(g->nx, g->ny, g->nz);
  double *part = dft_scratch_partials(red_blocks(n));
  DFT_BY_PREC(g->prec, reduce_abs2<T>((const cplx<T> *)g->dev, n, part, lib_stream()));
  return sum_partials(part, red_blocks(n), lib_stream()) * g->step * g->step * g->step;
}

// ------------------------------------------ further libgrid driver calls --
// (libgrid is not part of the reference tree; call sites:
// examples/vortex-ring/vring.c:63-71,149, examples/gas-bubble/bubble.c:104.)

// Momentum frame of the propagator and the kinetic energy: (k - k0)^2 / 2M.
void cgrid_set_momentum(cgrid *g, REAL kx0, REAL ky0, REAL kz0) {
  g->kx0 = kx0;
  g->ky0 = ky0;
  g->kz0 = kz0;
}

// dst = Laplacian of src by the 2nd-order central difference, periodic.
void cgrid_fd_laplace(cgrid *src, cgrid *dst) {
  const int64_t n = npts(src->nx, src->ny, src->nz);
  if (!need_same("cgrid_fd_laplace", n, npts(dst->nx, dst->ny, dst->nz), src->prec, dst->prec)) return;
  DFT_BY_PREC(src->prec, LAUNCH(k_fd_laplace<T>, n, (const cplx<T> *)src->dev, (cplx<T> *)dst->dev, (int)src->nx,
                                (int)src->ny, (int)src->nz, (T)(1.0 / (src->step * src->step))));
}

// dst = conj(a) * b
void cgrid_conjugate_product(cgrid *dst, cgrid *a, cgrid *b) {
  const int64_t n = npts(dst->nx, dst->ny, dst->nz);
  if (!need_same("cgrid_conjugate_product", n, npts(a->nx, a->ny, a->nz), dst->prec, a->prec) ||
      !need_same("cgrid_conjugate_product", n, npts(b->nx, b->ny, b->nz), dst->prec, b->prec))
    return;
  DFT_BY_PREC(dst->prec, LAUNCH(k_conj_product<T>, n, (cplx<T> *)dst->dev, (const cplx<T> *)a->dev,
                                (const cplx<T> *)b->dev, n));
}

// dst = |src|^p (real values in a complex grid)
void cgrid_abs_power(cgrid *dst, cgrid *src, REAL p) {
  const int64_t n = npts(dst->nx, dst->ny, dst->nz);
  if (!need_same("cgrid_abs_power", n, npts(src->nx, src->ny, src->nz), dst->prec, src->prec)) return;
  DFT_BY_PREC(dst->prec, LAUNCH(k_abs_power<T>, n, (cplx<T> *)dst->dev, (const cplx<T> *)src->dev, (T)p, n));
}

// |curl (fx, fy, fz)| (examples/gas-bubble/analyze.c:99)
void rgrid_abs_rot(rgrid *rot, rgrid *fx, rgrid *fy, rgrid *fz) {
  const int64_t n = npts(rot->nx, rot->ny, rot->nz);
  for (rgrid *f : {fx, fy, fz})
    if (!need_same("rgrid_abs_rot", n, npts(f->nx, f->ny, f->nz), rot->prec, f->prec)) return;
  DFT_BY_PREC(rot->prec, LAUNCH(k_abs_rot<T>, n, (T *)rot->dev, (const T *)fx->dev, (const T *)fy->dev,
                                (const T *)fz->dev, (int)rot->nx, (int)rot->ny, (int)rot->nz, (T)(0.5 / rot->step)));
}

// dst = |src|^p (examples/gas-bubble/analyze.c:100)
void rgrid_abs_power(rgrid *dst, rgrid *src, REAL p) {
  const int64_t n = npts(dst->nx, dst->ny, dst->nz);
  if (!need_same("rgrid_abs_power", n, npts(src->nx, src->ny, src->nz), dst->prec, src->prec)) return;
  DFT_BY_PREC(dst->prec, LAUNCH(k_rabs_power<T>, n, (T *)dst->dev, (const T *)src->dev, (T)p, n));
}

// dst = src^p (libgrid rgrid_power; examples/soliton/soliton.c:96)
void rgrid_power(rgrid *dst, rgrid *src, REAL p) {
  const int64_t n = npts(dst->nx, dst->ny, dst->nz);
  if (!need_same("rgrid_power", n, npts(src->nx, src->ny, src->nz), dst->prec, src->prec)) return;
  DFT_BY_PREC(dst->prec, LAUNCH(k_rpower<T>, n, (T *)dst->dev, (const T *)src->dev, (T)p, n));
}

// dst = a b (libgrid rgrid_product)
void rgrid_product(rgrid *dst, rgrid *a, rgrid *b) {
  const int64_t n = npts(dst->nx, dst->ny, dst->nz);
  for (rgrid *f : {a, b})
    if (!need_same("rgrid_product", n, npts(f->nx, f->ny, f->nz), dst->prec, f->prec)) return;
  DFT_BY_PREC(dst->prec, LAUNCH(k_rproduct<T>, n, (T *)dst->dev, (const T *)a->dev, (const T *)b->dev, n));
}

// dst = a / (b + eps) (libgrid rgrid_division_eps; examples/vortex-trace/trace.c:183)
void rgrid_division_eps(rgrid *dst, rgrid *a, rgrid *b, REAL eps) {
  const int64_t n = npts(dst->nx, dst->ny, dst->nz);
  for (rgrid *f : {a, b})
    if (!need_same("rgrid_division_eps", n, npts(f->nx, f->ny, f->nz), dst->prec, f->prec)) return;
  DFT_BY_PREC(dst->prec, LAUNCH(k_rdivision_eps<T>, n, (T *)dst->dev, (const T *)a->dev, (const T *)b->dev, (T)eps, n));
}

static double rgrid_extreme(rgrid *g, double sgn) {
  const int64_t n = npts(g->nx, g->ny, g->nz), nb = red_blocks(n);
  double *part = dft_scratch_partials(nb);
  DFT_BY_PREC(g->prec, hipLaunchKernelGGL(k_red_max<T>, dim3((unsigned)nb), dim3(kBlock), 0, lib_stream(),
                                          (const T *)g->dev, (T)sgn, n, part));
  DFT_LAUNCH_CHECK();
  std::vector<double> h(nb);
  DFT_HIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, lib_stream()));
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  double m = -INFINITY;
  for (double x : h) m = std::fmax(m, x);
  return sgn * m;
}
// largest and smallest value (libgrid rgrid_max / rgrid_min; examples/vortex-trace/trace.c:185)
REAL rgrid_max(rgrid *g) { return rgrid_extreme(g, 1.0); }
REAL rgrid_min(rgrid *g) { return rgrid_extreme(g, -1.0); }

// Probability flux j = (hbar / M) Im(psi* grad psi), central differences
// (examples/gas-bubble/analyze.c:78).  Any of jx, jy, jz may be NULL.
void grid_wf_probability_flux(wf *w, rgrid *jx, rgrid *jy, rgrid *jz) {
  cgrid *g = w->grid;
  const int64_t n = npts(g->nx, g->ny, g->nz);
  for (rgrid *v : {jx, jy, jz})
    if (v && !need_same("grid_wf_probability_flux", n, npts(v->nx, v->ny, v->nz), g->prec, v->prec)) return;
  DFT_BY_PREC(g->prec, LAUNCH(k_flux<T>, n, (const cplx<T> *)g->dev, jx ? (T *)jx->dev : nullptr,
                              jy ? (T *)jy->dev : nullptr, jz ? (T *)jz->dev : nullptr, (int)g->nx, (int)g->ny,
                              (int)g->nz, (T)(0.5 / g->step), (T)(1.0 / w->mass)));
}

// one component of the flux (libgrid grid_wf_probability_flux_x/y/z)
void grid_wf_probability_flux_x(wf *w, rgrid *j) { grid_wf_probability_flux(w, j, nullptr, nullptr); }
void grid_wf_probability_flux_y(wf *w, rgrid *j) { grid_wf_probability_flux(w, nullptr, j, nullptr); }
void grid_wf_probability_flux_z(wf *w, rgrid *j) { grid_wf_probability_flux(w, nullptr, nullptr, j); }

// h^3 sum f(x, y, z) g over the grid, f evaluated on the host
// (examples/gas-bubble/analyze.c:88).
REAL rgrid_weighted_integral(rgrid *g, REAL (*func)(void *arg, REAL x, REAL y, REAL z), void *farg) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  std::vector<double> h(n);
  rgrid_copy_to_host(g, h.data());
  double sum = 0.0, c = 0.0;
  for (INT i = 0; i < g->nx; ++i)
    for (INT j = 0; j < g->ny; ++j)
      for (INT k = 0; k < g->nz; ++k) {
        const double v = h[(i * g->ny + j) * g->nz + k] *
                         func(farg, (double)(i - g->nx / 2) * g->step - g->x0, (double)(j - g->ny / 2) * g->step - g->y0,
                              (double)(k - g->nz / 2) * g->step - g->z0);
        const double t = sum + v;
        c += std::fabs(sum) >= std::fabs(v) ? (sum - t) + v : (v - t) + sum;
        sum = t;
      }
  return (sum + c) * g->step * g->step * g->step;
}

}  // extern "C"

// Index range of the points i whose coordinate (i - n/2) h - o lies in [lo, hi].
static void axis_range(INT n, REAL h, REAL o, REAL lo, REAL hi, int *a, int *b) {
  const double ia = std::ceil((double)(n / 2) + (lo + o) / h - 1e-9), ib = std::floor((double)(n / 2) + (hi + o) / h + 1e-9);
  *a = (int)std::max(0.0, ia);
  *b = (int)std::min((double)(n - 1), ib);
}

template <typename G> static dft::Box region_box(const G *g, REAL xl, REAL xu, REAL yl, REAL yu, REAL zl, REAL zu) {
  dft::Box bx;
  axis_range(g->nx, g->step, g->x0, xl, xu, &bx.b[0], &bx.b[1]);
  axis_range(g->ny, g->step, g->y0, yl, yu, &bx.b[2], &bx.b[3]);
  axis_range(g->nz, g->step, g->z0, zl, zu, &bx.b[4], &bx.b[5]);
  return bx;
}

extern "C" {

// h^3 sum over the points with x in [xl, xu], y in [yl, yu], z in [zl, zu]
// (coordinates x_i = (i - nx/2) h - x0, bounds inclusive).
REAL rgrid_integral_region(rgrid *g, REAL xl, REAL xu, REAL yl, REAL yu, REAL zl, REAL zu) {
  const int64_t n = npts(g->nx, g->ny, g->nz);
  const dft::Box bx = region_box(g, xl, xu, yl, yu, zl, zu);
  double *part = dft_scratch_partials(red_blocks(n));
  DFT_BY_PREC(g->prec, hipLaunchKernelGGL(k_red_region<T>, dim3((unsigned)red_blocks(n)), dim3(kBlock), 0,
                                          lib_stream(), (const T *)g->dev, 1, (int)g->ny, (int)g->nz, bx, n, part));
  DFT_LAUNCH_CHECK();
  return sum_partials(part, red_blocks(n), lib_stream()) * g->step * g->step * g->step;
}

REAL_COMPLEX cgrid_integral_region(cgrid *g, REAL xl, REAL xu, REAL yl, REAL yu, REAL zl, REAL zu) {
  const int64_t n = npts(g->nx, g->ny, g->nz), nb = red_blocks(n);
  const dft::Box bx = region_box(g, xl, xu, yl, yu, zl, zu);
  double *part = dft_scratch_partials(2 * nb);
  DFT_BY_PREC(g->prec, hipLaunchKernelGGL(k_red_region<T>, dim3((unsigned)nb), dim3(kBlock), 0, lib_stream(),
                                          (const T *)g->dev, 2, (int)g->ny, (int)g->nz, bx, n, part));
  DFT_LAUNCH_CHECK();
  const double h3 = g->step * g->step * g->step;
  REAL_COMPLEX r;
  __real__ r = sum_partials(part, nb, lib_stream()) * h3;
  __imag__ r = sum_partials(part + nb, nb, lib_stream()) * h3;
  return r;
}

// h^3 sum psi (whole grid)
REAL_COMPLEX cgrid_integral(cgrid *g) {
  const double big = 1e300;
  return cgrid_integral_region(g, -big, big, -big, big, -big, big);
}

// h^3 sum conj(a) b
REAL_COMPLEX cgrid_integral_of_conjugate_product(cgrid *a, cgrid *b) {
  const int64_t n = npts(a->nx, a->ny, a->nz), nb = red_blocks(n);
  REAL_COMPLEX r;
  __real__ r = NAN;
  __imag__ r = NAN;
  if (!need_same("cgrid_integral_of_conjugate_product", n, npts(b->nx, b->ny, b->nz), a->prec, b->prec)) return r;
  double *part = dft_scratch_partials(2 * nb);
  DFT_BY_PREC(a->prec, hipLaunchKernelGGL(k_red_conj_product<T>, dim3((unsigned)nb), dim3(kBlock), 0, lib_stream(),
                                          (const cplx<T> *)a->dev, (const cplx<T> *)b->dev, n, part));
  DFT_LAUNCH_CHECK();
  const double h3 = a->step * a->step * a->step;
  __real__ r = sum_partials(part, nb, lib_stream()) * h3;
  __imag__ r = sum_partials(part + nb, nb, lib_stream()) * h3;
  return r;
}

void grid_wf_normalize(wf *w) {
  const double nrm = grid_wf_norm(w);
  if (!(nrm > 0)) {
    fprintf(stderr, "libdft: grid_wf_normalize: zero norm\n");
    return;
  }
  REAL_COMPLEX c = std::sqrt(w->norm / nrm);
  cgrid_multiply(w->grid, c);
}

}  // extern "C"

// Plans shared by the propagator entry points, one per (shape, step, precision).
struct PlanCache {
  int nx = 0, ny = 0, nz = 0, prec = 0;
  double step = 0;
  void *plan = nullptr;
  cgrid *scratch = nullptr;  // spectrum scratch for grid_wf_energy
};
static PlanCache g_pc;

// Spectrum scratch of grid_wf_energy, kept for the last shape and precision
// (allocated outside the lock: allocation queues a clear on the library stream).
static cgrid *energy_scratch(cgrid *g) {
  cgrid *old = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    cgrid *c = g_pc.scratch;
    if (c && c->nx == g->nx && c->ny == g->ny && c->nz == g->nz && c->prec == g->prec) return c;
    old = c;
    g_pc.scratch = nullptr;
  }
  if (old) cgrid_free(old);
  cgrid *c = cgrid_clone(g, (char *)"energy scratch");
  std::lock_guard<std::mutex> lk(g_mu);
  g_pc.scratch = c;
  return c;
}

template <typename T> static Plan3D<T> *plan_for(cgrid *g) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_pc.plan && (g_pc.nx != g->nx || g_pc.ny != g->ny || g_pc.nz != g->nz || g_pc.prec != g->prec ||
                    g_pc.step != g->step)) {
    DFT_HIP_CHECK(hipStreamSynchronize(g_stream));
    if (g_pc.prec == GRID_FP32) {
      ((Plan3D<float> *)g_pc.plan)->destroy();
      delete (Plan3D<float> *)g_pc.plan;
    } else {
      ((Plan3D<double> *)g_pc.plan)->destroy();
      delete (Plan3D<double> *)g_pc.plan;
    }
    g_pc.plan = nullptr;
  }
  if (!g_pc.plan) {
    auto *p = new Plan3D<T>();
    p->init((int)g->nx, (int)g->ny, (int)g->nz, g->step);
    g_pc.plan = p;
    g_pc.nx = (int)g->nx;
    g_pc.ny = (int)g->ny;
    g_pc.nz = (int)g->nz;
    g_pc.prec = g->prec;
    g_pc.step = g->step;
  }
  return (Plan3D<T> *)g_pc.plan;
}

extern "C" {

// Momentum frame of a grid in units of each axis' 2 pi / L.
static void kframe(const cgrid *g, double o[3]) {
  o[0] = g->kx0 * g->nx * g->step / (2.0 * M_PI);
  o[1] = g->ky0 * g->ny * g->step / (2.0 * M_PI);
  o[2] = g->kz0 * g->nz * g->step / (2.0 * M_PI);
}

// Kinetic energy h^3/N sum_k ((k - k0)^2 / 2M) |psi_k|^2 by FFT (libgrid's FFT
// kinetic energy, the propagator's own symbol), plus h^3 sum |psi|^2 Re V
// when a potential is given.
REAL grid_wf_energy(wf *w, cgrid *potential) {
  cgrid *g = w->grid;
  const int64_t n = npts(g->nx, g->ny, g->nz);
  if (potential && !need_same("grid_wf_energy", n, npts(potential->nx, potential->ny, potential->nz), g->prec,
                              potential->prec))
    return NAN;
  cgrid *tmp = energy_scratch(g);
  if (!tmp) {
    fprintf(stderr, "libdft: grid_wf_energy: out of device memory\n");
    return NAN;
  }
  double *part = dft_scratch_partials(red_blocks(n));
  const double cx = std::pow(2.0 * M_PI / (g->nx * g->step), 2) / (2.0 * w->mass);
  const double cy = std::pow(2.0 * M_PI / (g->ny * g->step), 2) / (2.0 * w->mass);
  const double cz = std::pow(2.0 * M_PI / (g->nz * g->step), 2) / (2.0 * w->mass);
  double o[3];
  kframe(g, o);
  DFT_BY_PREC(g->prec, {
    Plan3D<T> *p = plan_for<T>(g);
    c2c_fwd<T>(*p, (const cplx<T> *)g->dev, (cplx<T> *)tmp->dev, lib_stream());
    reduce_kinetic<T>((const cplx<T> *)tmp->dev, (int)g->nx, (int)g->ny, (int)g->nz, cx, cy, cz, part, lib_stream(), o);
  });
  const double s = sum_partials(part, red_blocks(n), lib_stream());
  double e = s * g->step * g->step * g->step / (double)n;
  if (potential) e += grid_wf_potential_energy(w, potential);
  return e;
}

// ----------------------------------------------------------- propagation --
void grid_wf_propagate(wf *w, cgrid *potential, REAL_COMPLEX time) {
  cgrid *g = w->grid;
  if (!need_same("grid_wf_propagate", npts(g->nx, g->ny, g->nz), npts(potential->nx, potential->ny, potential->nz),
            g->prec, potential->prec)) return;
  DFT_BY_PREC(g->prec, {
    Plan3D<T> *p = plan_for<T>(g);
    const double k0[3] = {g->kx0, g->ky0, g->kz0};
    strang<T>(*p, (cplx<T> *)g->dev, (const cplx<T> *)potential->dev, __real__ time, __imag__ time, w->mass, nullptr,
              lib_stream(), k0);
  });
}

void grid_wf_propagate_predict(wf *w, wf *wp, cgrid *potential, REAL_COMPLEX time) {
  cgrid_copy(wp->grid, w->grid);
  grid_wf_propagate(wp, potential, time);
}

void grid_wf_propagate_correct(wf *w, cgrid *potential, REAL_COMPLEX time) { grid_wf_propagate(w, potential, time); }

// ------------------------------------------------------------- profiling --
// Per-launch HIP-event timing of the library's kernels (all on one stream).
void dft_profile_enable(int on) {
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  g_prof_on = on != 0;
}

void dft_profile_reset(void) {
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  g_prof.clear();
  g_agg.clear();
  g_evnext = 0;
}

// Folds the recorded launches into per-kernel totals; returns the number of kernels.
int dft_profile_collect(void) {
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  for (const ProfRec &r : g_prof) {
    float ms = 0.f;
    DFT_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    ProfAgg *a = nullptr;
    for (ProfAgg &x : g_agg)
      if (x.name == r.name) a = &x;
    if (!a) {
      g_agg.push_back({r.name, 0, 0.0, 0.0});
      a = &g_agg.back();
    }
    a->launches += 1;
    a->ms += ms;
    a->bytes += r.bytes;
  }
  g_prof.clear();
  g_evnext = 0;
  return (int)g_agg.size();
}

int dft_profile_entry(int i, char *name, int len, int64_t *launches, double *ms, double *bytes) {
  if (i < 0 || i >= (int)g_agg.size()) return -1;
  snprintf(name, len, "%s", g_agg[i].name.c_str());
  *launches = g_agg[i].launches;
  *ms = g_agg[i].ms;
  *bytes = g_agg[i].bytes;
  return 0;
}

// Plain copy of n doubles (device buffers of the library's grids), for
// calibrating hardware traffic counters against a kernel of known traffic.
static __global__ void k_copy_d(const double *__restrict__ a, double *__restrict__ b, int64_t n) {
  GRID_STRIDE(i, n) b[i] = a[i];
}
void grid_bench_copy(rgrid *dst, rgrid *src) {
  const int64_t n = npts(src->nx, src->ny, src->nz) * (int64_t)esize(src->prec) / (int64_t)sizeof(double);
  if (!need_same("grid_bench_copy", npts(dst->nx, dst->ny, dst->nz), npts(src->nx, src->ny, src->nz), dst->prec,
            src->prec)) return;
  DFT_PROF_BEGIN(lib_stream());
  LAUNCH(k_copy_d, n, (const double *)src->dev, (double *)dst->dev, n);
  DFT_PROF_END(lib_stream(), "copy", 2.0 * n * sizeof(double));
}

// ---------------------------------------------------------------- timers --
void grid_timer_start(grid_timer *t) {
  t->t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
REAL grid_timer_wall_clock_time(grid_timer *t) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t->t0;
}

}  // extern "C"
