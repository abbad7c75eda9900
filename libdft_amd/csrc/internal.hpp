// Internal C++ interfaces shared by the library's translation units.
#pragma once
#include "common.hpp"
#include "fft.hpp"
#include "../../include/libdft_mi355x.h"

namespace dft {

hipStream_t lib_stream();
void *dmalloc(size_t bytes);
void dfree(void *p);

// ------------------------------------------------------------------ FFT ----
template <typename T> struct Plan3D {
  int nx = 0, ny = 0, nz = 0, P = 0;
  int PP = 0;  // row pitch of the half spectra (half_pitch(nz) >= P)
  double step = 0;
  cplx<T> *twx = nullptr, *twy = nullptr, *twz = nullptr;
  T *sinx = nullptr, *siny = nullptr, *sinz = nullptr;
  void init(int nx_, int ny_, int nz_, double step_);
  void destroy();
  int64_t n() const { return (int64_t)nx * ny * nz; }
  int64_t nhalf() const { return (int64_t)nx * ny * P; }      // values (traffic counts)
  int64_t nhalf_alloc() const { return (int64_t)nx * ny * PP; }  // storage
};

bool fft_size_ok(int64_t n);

// Real fields -> half spectra [nx][ny][nz/2+1] (unnormalised forward transform).
template <typename T>
void r2c3d(const Plan3D<T> &p, int nb, const T *const *in, cplx<T> *const *out, hipStream_t s);
// Half spectra -> real fields: out_m = F^-1[ sum_t s_t K_t D_t in_t ], one combine per output.
// tmp: nout half-spectrum scratch buffers.
template <typename T>
void c2r3d(const Plan3D<T> &p, int nin, const cplx<T> *const *in, int nout, const fft::Combine *cmb,
           const fft::KernTab &kt, cplx<T> *const *tmp, T *const *out, hipStream_t s);
// One Strang step on psi; partial (may be null) receives per-block sums of |psi|^2.
// k0 (may be null): momentum frame (kx0, ky0, kz0), kinetic symbol (k - k0)^2 / 2M.
template <typename T>
void strang(const Plan3D<T> &p, cplx<T> *psi, const cplx<T> *pot, double dtr, double dti, double mass,
            double *partial, hipStream_t s, const double *k0 = nullptr);
int64_t strang_partials(int nx, int ny, int nz, int prec);
// y pass (dir -1 forward, +1 inverse) over nb fields of `cols` columns with
// row pitch `pitch`, in place.
template <typename T>
void ypass(const Plan3D<T> &p, int nb, cplx<T> *const *bufs, int dir, int cols, int pitch, hipStream_t s);
// Renormalisation fused into a step's closing z pass: psi *= sqrt(target /
// (h3 sum |psi|^2)).  partial: strang_partials() + 1 doubles of device memory.
struct Renorm {
  double *partial;
  double h3, target;
};
// Rest of a Strang step after the fused z pass: y, x (with kinetic phase), y,
// and the inverse z pass multiplying by e^{-iV dt/2} with V = `phase` (the
// potential the first half step used); with rn, that pass also renormalises.
template <typename T>
void strang_tail(const Plan3D<T> &p, cplx<T> *psi, const cplx<T> *phase, double dtr, double dti, double mass,
                 hipStream_t s, const double *k0 = nullptr, const Renorm *rn = nullptr);
// Full complex forward transform (out of place).
template <typename T> void c2c_fwd(const Plan3D<T> &p, const cplx<T> *in, cplx<T> *out, hipStream_t s);
// Its inverse, normalised by 1/N (out of place).
template <typename T> void c2c_inv(const Plan3D<T> &p, const cplx<T> *in, cplx<T> *out, hipStream_t s);

// ----------------------------------------------------------- reductions ----
// Neumaier sum of n FP64 partials (device pointer), synchronising.
double sum_partials(const double *dpartials, int64_t n, hipStream_t s);
int64_t red_blocks(int64_t n);
template <typename T> void reduce_sum(const T *a, const T *b, int64_t n, double *partial, hipStream_t s);
template <typename T> void reduce_abs2(const cplx<T> *a, int64_t n, double *partial, hipStream_t s);
template <typename T>
void reduce_density_re(const cplx<T> *a, const cplx<T> *v, int64_t n, double *partial, hipStream_t s);
// K = cx (mx - ox)^2 + cy (my - oy)^2 + cz (mz - oz)^2 (o: momentum frame in units of 2 pi / L)
template <typename T>
void reduce_kinetic(const cplx<T> *spec, int nx, int ny, int nz, double cx, double cy, double cz, double *partial,
                    hipStream_t s, const double *o = nullptr);

// ------------------------------------------------------------- elementwise --
template <typename T> void ew_density(const cplx<T> *psi, T *rho, int64_t n, hipStream_t s);

}  // namespace dft

// Library-wide scratch for FP64 reduction partials (grows on demand).
double *dft_scratch_partials(int64_t n);
