// Host side of the FFT: plans (FP64-computed twiddles), per-size dispatch and
// the 3-D transform sequences built from the pass kernels.
#include <cmath>
#include <vector>

#include "fft_kernels.hpp"
#include "internal.hpp"

namespace dft {

bool fft_size_ok(int64_t n) {
  return n == 4 || n == 8 || n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512;
}

// Twiddle table of 2n entries: [0, n) holds exp(-2 pi i m / n) rounded to T;
// for FP32, [n, 2n) holds the rounding residual of each entry, so that the
// kernels can multiply by the twiddle to ~1e-14 in modulus (the inverse
// transform uses the conjugates).  For FP64 the second half is zero and unused.
template <typename T> static cplx<T> *make_twiddles(int n) {
  std::vector<cplx<T>> h(2 * n);
  for (int m = 0; m < n; ++m) {
    const double a = 2.0 * M_PI * (double)m / (double)n, c = std::cos(a), s = -std::sin(a);
    const T ch = (T)c, sh = (T)s;
    h[m] = cmk<cplx<T>>(ch, sh);
    h[n + m] = sizeof(T) == 4 ? cmk<cplx<T>>((T)(c - (double)ch), (T)(s - (double)sh)) : cmk<cplx<T>>(0, 0);
  }
  auto *d = (cplx<T> *)dmalloc(sizeof(cplx<T>) * 2 * n);
  DFT_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(cplx<T>) * 2 * n, hipMemcpyHostToDevice));
  return d;
}

// sin(2 pi m / n) / h for m in [0, cnt): the symbol of the periodic central difference.
template <typename T> static T *make_sin(int n, int cnt, double step) {
  std::vector<T> h(cnt);
  for (int m = 0; m < cnt; ++m) h[m] = (T)(std::sin(2.0 * M_PI * (double)m / (double)n) / step);
  auto *d = (T *)dmalloc(sizeof(T) * cnt);
  DFT_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(T) * cnt, hipMemcpyHostToDevice));
  return d;
}

template <typename T> void Plan3D<T>::init(int nx_, int ny_, int nz_, double step_) {
  nx = nx_;
  ny = ny_;
  nz = nz_;
  step = step_;
  P = nz / 2 + 1;
  PP = half_pitch(nz);
  twx = make_twiddles<T>(nx);
  twy = make_twiddles<T>(ny);
  twz = make_twiddles<T>(nz);
  sinx = make_sin<T>(nx, nx, step);
  siny = make_sin<T>(ny, ny, step);
  sinz = make_sin<T>(nz, P, step);
}

template <typename T> void Plan3D<T>::destroy() {
  dfree(twx);
  dfree(twy);
  dfree(twz);
  dfree(sinx);
  dfree(siny);
  dfree(sinz);
  twx = twy = twz = nullptr;
  sinx = siny = sinz = nullptr;
}

namespace {

template <typename T, int N>
void launch_zr2c(const Plan3D<T> &p, const fft::Batch &b, int nb, hipStream_t s) {
  using G = rf::LineCfg<T, N>;
  const int64_t npairs = (int64_t)p.nx * p.ny / 2;
  dim3 grid((unsigned)((npairs + G::LINES - 1) / G::LINES), nb);
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_zr2c<T, N>), grid, dim3(G::BS), 0, s, b, npairs, p.PP, p.twz);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "fft_z_r2c", (double)nb * (p.n() * sizeof(T) + p.nhalf() * sizeof(cplx<T>)));
}

template <typename T, int N>
void launch_zc2r(const Plan3D<T> &p, const fft::Batch &b, int nb, T scale, hipStream_t s) {
  using G = rf::LineCfg<T, N>;
  const int64_t npairs = (int64_t)p.nx * p.ny / 2;
  dim3 grid((unsigned)((npairs + G::LINES - 1) / G::LINES), nb);
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_zc2r<T, N>), grid, dim3(G::BS), 0, s, b, npairs, p.PP, p.twz, scale);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "fft_z_c2r", (double)nb * (p.n() * sizeof(T) + p.nhalf() * sizeof(cplx<T>)));
}

// Strided pass along y (axis 1) or x (axis 0) over columns [0, cols) of rows
// of `pitch` elements.
template <typename T, int N, int DIR>
void launch_strided(const Plan3D<T> &p, const fft::Batch &b, int nb, int axis, int cols, int pitch, T scale,
                    hipStream_t s) {
  using G = fft::StrCfg<T, N>;
  const int ntile = (cols + G::COLS - 1) / G::COLS;
  int64_t es, os;
  int outer;
  if (axis == 1) {
    es = pitch;
    os = (int64_t)p.ny * pitch;
    outer = p.nx;
  } else {
    es = (int64_t)p.ny * pitch;
    os = pitch;
    outer = p.ny;
  }
  dim3 grid((unsigned)(outer * ntile), nb);
  const cplx<T> *tw = axis == 1 ? p.twy : p.twx;
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_strided<T, N, DIR>), grid, dim3(G::BS), 0, s, b, cols, es, os, ntile, tw, scale);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, axis == 1 ? (DIR < 0 ? "fft_y_c2c_fwd" : "fft_y_c2c_inv") : (DIR < 0 ? "fft_x_c2c_fwd" : "fft_x_c2c_inv"),
               2.0 * nb * (double)p.nx * p.ny * cols * sizeof(cplx<T>));
}

template <typename T, int N>
void launch_xcombine(const Plan3D<T> &p, const fft::Batch &b, const fft::CombineSet &cs, const fft::KernTab &kt,
                     int nb, hipStream_t s) {
  using G = rf::ColCfg<T, N>;
  const int ntile = (p.P + G::COLS - 1) / G::COLS;
  dim3 grid((unsigned)(p.ny * ntile), nb);
  // traffic: each term's spectrum and kernel, plus the output
  double per = 0;
  for (int i = 0; i < nb; ++i) {
    per += sizeof(cplx<T>);
    for (int t = 0; t < cs.c[i].nterms; ++t)
      per += sizeof(cplx<T>) + (cs.c[i].t[t].kern >= 0 ? sizeof(T) : 0);
  }
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_xcombine_inv<T, N>), grid, dim3(G::BS), 0, s, b, cs, kt, p.ny, p.P, p.PP, ntile,
                     p.twx);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "fft_x_combine_c2c", per * (double)p.nhalf());
}

template <typename T, int N, int DIR>
void launch_zc2c(const Plan3D<T> &p, const cplx<T> *src, cplx<T> *dst, const cplx<T> *pot, double dtr, double dti,
                 T scale, double *partial, hipStream_t s) {
  using G = rf::LineCfg<T, N>;
  const int64_t nlines = (int64_t)p.nx * p.ny;
  dim3 grid((unsigned)((nlines + G::LINES - 1) / G::LINES));
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_zc2c_pot<T, N, DIR>), grid, dim3(G::BS), 0, s, src, dst, pot, nlines, p.twz, dtr,
                     dti, scale, partial);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, DIR < 0 ? "fft_z_c2c_vphase_fwd" : "fft_z_c2c_vphase_inv",
               (double)p.n() * sizeof(cplx<T>) * (pot ? 3.0 : 2.0));
}

template <typename T, int N>
void launch_xkin(const Plan3D<T> &p, cplx<T> *psi, double cx, double cy, double cz, double dtr, double dti,
                 const double *k0, hipStream_t s) {
  // momentum frame: (k - k0)^2, k0 in units of each axis' 2 pi / L
  const double ox = k0 ? k0[0] * p.nx * p.step / (2.0 * M_PI) : 0.0, oy = k0 ? k0[1] * p.ny * p.step / (2.0 * M_PI) : 0.0,
               oz = k0 ? k0[2] * p.nz * p.step / (2.0 * M_PI) : 0.0;
  using G = fft::KinCfg<T, N>;
  const int ntile = (p.nz + G::COLS - 1) / G::COLS;
  dim3 grid((unsigned)(p.ny * ntile));
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_xkin<T, N>), grid, dim3(G::BS), 0, s, psi, p.ny, p.nz, ntile, p.twx, cx, cy, cz,
                     dtr, dti, ox, oy, oz);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "fft_x_kinetic_c2c", 2.0 * (double)p.n() * sizeof(cplx<T>));
}

// Inverse z pass of a step: transform, scale, and multiply by e^{-iV dt/2}
// with V = veff, the potential the step's first half used.
template <typename T, int N>
void launch_zc2c_phase_inv(const Plan3D<T> &p, cplx<T> *psi, const cplx<T> *veff, double dtr, double dti, T scale,
                           hipStream_t s) {
  using G = rf::LineCfg<T, N>;
  const int64_t nlines = (int64_t)p.nx * p.ny;
  dim3 grid((unsigned)((nlines + G::LINES - 1) / G::LINES));
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_zc2c_pot<T, N, +1>), grid, dim3(G::BS), 0, s, psi, psi, veff, nlines, p.twz, dtr, dti,
                     scale, (double *)nullptr);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "fft_z_c2c_phase_inv", 3.0 * (double)p.n() * sizeof(cplx<T>));
}

// The same pass followed by the renormalisation of rn: the pass writes its
// per-block sums of |psi|^2, one block forms the factor, a last pass scales psi.
// (A single cooperative kernel summing through a grid barrier measured slower
// at 64^3: 8.5k against 12.1k steps/s, the cooperative launch's own cost.)
template <typename T, int N>
void launch_zc2c_phase_inv_norm(const Plan3D<T> &p, cplx<T> *psi, const cplx<T> *veff, double dtr, double dti,
                                T scale, const Renorm &rn, hipStream_t s) {
  using G = rf::LineCfg<T, N>;
  const int64_t nlines = (int64_t)p.nx * p.ny;
  const int nb = (int)((nlines + G::LINES - 1) / G::LINES);
  DFT_PROF_BEGIN(s);
  hipLaunchKernelGGL((fft::k_zc2c_pot<T, N, +1>), dim3((unsigned)nb), dim3(G::BS), 0, s, psi, psi, veff, nlines,
                     p.twz, dtr, dti, scale, rn.partial);
  DFT_LAUNCH_CHECK();
  hipLaunchKernelGGL(fft::k_norm_factor, dim3(1), dim3(256), 0, s, rn.partial, nb, rn.h3, rn.target);
  DFT_LAUNCH_CHECK();
  hipLaunchKernelGGL(fft::k_scale_psi<T>, dim3(grid_for(p.n())), dim3(kBlock), 0, s, psi,
                     (const double *)rn.partial + nb, p.n());
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "fft_z_c2c_phase_inv_norm", 5.0 * (double)p.n() * sizeof(cplx<T>));
}

template <typename T> int zc2c_lines(int nz) {
  int t = 1;
  DFT_DISPATCH_N(nz, (t = rf::LineCfg<T, N_>::LINES));
  return t;
}

}  // namespace

int64_t strang_partials(int nx, int ny, int nz, int prec) {
  const int t = prec == GRID_FP32 ? zc2c_lines<float>(nz) : zc2c_lines<double>(nz);
  return ((int64_t)nx * ny + t - 1) / t;
}

template <typename T>
void r2c3d(const Plan3D<T> &p, int nb, const T *const *in, cplx<T> *const *out, hipStream_t s) {
  for (int b0 = 0; b0 < nb; b0 += fft::kMaxBatch) {
    const int cnt = nb - b0 < fft::kMaxBatch ? nb - b0 : fft::kMaxBatch;
    fft::Batch bz{}, by{};
    for (int i = 0; i < cnt; ++i) {
      bz.in[i] = in[b0 + i];
      bz.out[i] = out[b0 + i];
      by.in[i] = out[b0 + i];
      by.out[i] = out[b0 + i];
    }
    DFT_DISPATCH_N(p.nz, (launch_zr2c<T, N_>(p, bz, cnt, s)));
    DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, -1>(p, by, cnt, 1, p.P, p.PP, (T)1, s)));
    DFT_DISPATCH_N(p.nx, (launch_strided<T, N_, -1>(p, by, cnt, 0, p.P, p.PP, (T)1, s)));
  }
}

template <typename T>
void c2r3d(const Plan3D<T> &p, int nin, const cplx<T> *const *in, int nout, const fft::Combine *cmb,
           const fft::KernTab &kt, cplx<T> *const *tmp, T *const *out, hipStream_t s) {
  if (nin > fft::kMaxBatch) {
    fprintf(stderr, "libdft: c2r3d: too many inputs\n");
    abort();
  }
  for (int b0 = 0; b0 < nout; b0 += fft::kMaxBatch) {
    const int cnt = nout - b0 < fft::kMaxBatch ? nout - b0 : fft::kMaxBatch;
    fft::Batch bx{}, by{}, bz{};
    fft::CombineSet cs{};
    for (int i = 0; i < nin; ++i) bx.in[i] = in[i];
    for (int i = 0; i < cnt; ++i) {
      bx.out[i] = tmp[b0 + i];
      by.in[i] = tmp[b0 + i];
      by.out[i] = tmp[b0 + i];
      bz.in[i] = tmp[b0 + i];
      bz.out[i] = out[b0 + i];
      cs.c[i] = cmb[b0 + i];
    }
    DFT_DISPATCH_N(p.nx, (launch_xcombine<T, N_>(p, bx, cs, kt, cnt, s)));
    DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, +1>(p, by, cnt, 1, p.P, p.PP, (T)1, s)));
    DFT_DISPATCH_N(p.nz, (launch_zc2r<T, N_>(p, bz, cnt, (T)1, s)));
  }
}

template <typename T>
void strang(const Plan3D<T> &p, cplx<T> *psi, const cplx<T> *pot, double dtr, double dti, double mass,
            double *partial, hipStream_t s, const double *k0) {
  fft::Batch by{};
  by.in[0] = psi;
  by.out[0] = psi;
  const double L2x = std::pow(2.0 * M_PI / (p.nx * p.step), 2), L2y = std::pow(2.0 * M_PI / (p.ny * p.step), 2),
               L2z = std::pow(2.0 * M_PI / (p.nz * p.step), 2);
  const double cx = L2x / (2.0 * mass), cy = L2y / (2.0 * mass), cz = L2z / (2.0 * mass);
  DFT_DISPATCH_N(p.nz, (launch_zc2c<T, N_, -1>(p, psi, psi, pot, dtr, dti, (T)1, nullptr, s)));
  DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, -1>(p, by, 1, 1, p.nz, p.nz, (T)1, s)));
  DFT_DISPATCH_N(p.nx, (launch_xkin<T, N_>(p, psi, cx, cy, cz, dtr, dti, k0, s)));
  DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, +1>(p, by, 1, 1, p.nz, p.nz, (T)1, s)));
  const T inv = (T)(1.0 / (double)p.n());
  DFT_DISPATCH_N(p.nz, (launch_zc2c<T, N_, +1>(p, psi, psi, pot, dtr, dti, inv, partial, s)));
}

template <typename T>
void ypass(const Plan3D<T> &p, int nb, cplx<T> *const *bufs, int dir, int cols, int pitch, hipStream_t s) {
  for (int b0 = 0; b0 < nb; b0 += fft::kMaxBatch) {
    const int cnt = nb - b0 < fft::kMaxBatch ? nb - b0 : fft::kMaxBatch;
    fft::Batch b{};
    for (int i = 0; i < cnt; ++i) b.in[i] = b.out[i] = bufs[b0 + i];
    if (dir < 0) {
      DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, -1>(p, b, cnt, 1, cols, pitch, (T)1, s)));
    } else {
      DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, +1>(p, b, cnt, 1, cols, pitch, (T)1, s)));
    }
  }
}

template <typename T>
void strang_tail(const Plan3D<T> &p, cplx<T> *psi, const cplx<T> *phase, double dtr, double dti, double mass,
                 hipStream_t s, const double *k0, const Renorm *rn) {
  fft::Batch by{};
  by.in[0] = psi;
  by.out[0] = psi;
  const double L2x = std::pow(2.0 * M_PI / (p.nx * p.step), 2), L2y = std::pow(2.0 * M_PI / (p.ny * p.step), 2),
               L2z = std::pow(2.0 * M_PI / (p.nz * p.step), 2);
  const double cx = L2x / (2.0 * mass), cy = L2y / (2.0 * mass), cz = L2z / (2.0 * mass);
  DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, -1>(p, by, 1, 1, p.nz, p.nz, (T)1, s)));
  DFT_DISPATCH_N(p.nx, (launch_xkin<T, N_>(p, psi, cx, cy, cz, dtr, dti, k0, s)));
  DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, +1>(p, by, 1, 1, p.nz, p.nz, (T)1, s)));
  const T inv = (T)(1.0 / (double)p.n());
  if (rn) {
    DFT_DISPATCH_N(p.nz, (launch_zc2c_phase_inv_norm<T, N_>(p, psi, phase, dtr, dti, inv, *rn, s)));
  } else {
    DFT_DISPATCH_N(p.nz, (launch_zc2c_phase_inv<T, N_>(p, psi, phase, dtr, dti, inv, s)));
  }
}

template <typename T> void c2c_fwd(const Plan3D<T> &p, const cplx<T> *in, cplx<T> *out, hipStream_t s) {
  fft::Batch b{};
  b.in[0] = out;
  b.out[0] = out;
  DFT_DISPATCH_N(p.nz, (launch_zc2c<T, N_, -1>(p, in, out, nullptr, 0.0, 0.0, (T)1, nullptr, s)));
  DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, -1>(p, b, 1, 1, p.nz, p.nz, (T)1, s)));
  DFT_DISPATCH_N(p.nx, (launch_strided<T, N_, -1>(p, b, 1, 0, p.nz, p.nz, (T)1, s)));
}

// Inverse of c2c_fwd, normalised by 1/N (out of place).
template <typename T> void c2c_inv(const Plan3D<T> &p, const cplx<T> *in, cplx<T> *out, hipStream_t s) {
  fft::Batch b{}, bx{};
  bx.in[0] = in;
  bx.out[0] = out;
  b.in[0] = out;
  b.out[0] = out;
  DFT_DISPATCH_N(p.nx, (launch_strided<T, N_, +1>(p, bx, 1, 0, p.nz, p.nz, (T)1, s)));
  DFT_DISPATCH_N(p.ny, (launch_strided<T, N_, +1>(p, b, 1, 1, p.nz, p.nz, (T)1, s)));
  const T inv = (T)(1.0 / (double)p.n());
  DFT_DISPATCH_N(p.nz, (launch_zc2c<T, N_, +1>(p, out, out, nullptr, 0.0, 0.0, inv, nullptr, s)));
}

#define DFT_INST(T)                                                                                          \
  template struct Plan3D<T>;                                                                                 \
  template void r2c3d<T>(const Plan3D<T> &, int, const T *const *, cplx<T> *const *, hipStream_t);           \
  template void c2r3d<T>(const Plan3D<T> &, int, const cplx<T> *const *, int, const fft::Combine *,          \
                         const fft::KernTab &, cplx<T> *const *, T *const *, hipStream_t);                   \
  template void strang<T>(const Plan3D<T> &, cplx<T> *, const cplx<T> *, double, double, double, double *,   \
                          hipStream_t, const double *);                                                      \
  template void c2c_fwd<T>(const Plan3D<T> &, const cplx<T> *, cplx<T> *, hipStream_t);                    \
  template void c2c_inv<T>(const Plan3D<T> &, const cplx<T> *, cplx<T> *, hipStream_t);                    \
  template void ypass<T>(const Plan3D<T> &, int, cplx<T> *const *, int, int, int, hipStream_t);                  \
  template void strang_tail<T>(const Plan3D<T> &, cplx<T> *, const cplx<T> *, double, double, double,        \
                               hipStream_t, const double *, const Renorm *);
DFT_INST(float)
DFT_INST(double)

}  // namespace dft
