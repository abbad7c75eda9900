// grid_wf_normalize_noalloc: renormalisation of a wave function entirely on
// the device, with no allocation and no host synchronisation, for callers
// that cannot allocate (code under HIP graph capture, tight stepping loops).
//
// Its reduction partials live in `scratch`, a file-scope device array of
// kScratch doubles (8.2 KB) that each GPU holds until the process ends.  The
// host reaches it only through the address hipGetSymbolAddress returns,
// looked up once per device at the first call there.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race (the library
// orders its own calls on one stream).  Each call overwrites every partial it
// reads, so nothing needs clearing between calls.
#include <mutex>

#include "internal.hpp"

using namespace dft;

namespace {

// Partials of the |psi|^2 reduction (at most red_blocks(n) <= 1024 for any n,
// the reduction kernels being grid-stride), then the norm and the scale.
constexpr int kPartials = 1024;
constexpr int kScratch = kPartials + 2;
static __device__ __align__(256) double scratch[kScratch];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (nullptr, with a message, on failure).
double *scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fprintf(stderr, "libdft: grid_wf_normalize_noalloc: no usable current device\n");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      fprintf(stderr, "libdft: grid_wf_normalize_noalloc: hipGetSymbolAddress failed: %s\n", hipGetErrorString(e));
      return nullptr;
    }
    g_ptr[dev] = (double *)p;
  }
  return g_ptr[dev];
}

// One block: norm = h3 * sum of the partials, scale = sqrt(target / norm)
// (1 for a zero norm), written after the partials.
__global__ void k_norm_scale(double *s, int nb, double h3, double target) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) acc += s[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nrm = red[0] * h3;
    s[kPartials] = nrm;
    s[kPartials + 1] = nrm > 0.0 ? sqrt(target / nrm) : 1.0;
  }
}

template <typename T> __global__ void k_scale_by(cplx<T> *psi, const double *s, int64_t n) {
  const T f = (T)s[kPartials + 1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    psi[i] = cmk<cplx<T>>(psi[i].x * f, psi[i].y * f);
}

}  // namespace

extern "C" {

// grid_wf_normalize (libgrid; call site examples/imag_time_rot/imag_time.c:208)
// without allocation or host synchronisation: psi *= sqrt(gwf->norm / N),
// N = h^3 sum |psi|^2 in FP64, all on the library stream.
int grid_wf_normalize_noalloc(wf *gwf) {
  if (!gwf || !gwf->grid) {
    fprintf(stderr, "libdft: grid_wf_normalize_noalloc: no wave function\n");
    return -1;
  }
  cgrid *g = gwf->grid;
  const int64_t n = g->nx * g->ny * g->nz;
  const int64_t nb = red_blocks(n);
  if (nb > kPartials) {
    fprintf(stderr, "libdft: grid_wf_normalize_noalloc: %lld partial sums exceed the %d-entry scratch array\n",
            (long long)nb, kPartials);
    return -1;
  }
  double *s = scratch_ptr();
  if (!s) return -1;
  hipStream_t st = lib_stream();
  const double h3 = g->step * g->step * g->step;
  DFT_BY_PREC(g->prec, {
    reduce_abs2<T>((const cplx<T> *)g->dev, n, s, st);
    hipLaunchKernelGGL(k_norm_scale, dim3(1), dim3(kBlock), 0, st, s, (int)nb, h3, gwf->norm);
    DFT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scale_by<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, (cplx<T> *)g->dev, (const double *)s, n);
    DFT_LAUNCH_CHECK();
  });
  return 0;
}

}  // extern "C"
