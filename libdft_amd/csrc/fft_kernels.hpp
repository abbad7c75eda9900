// FFT pass kernels (see fft.hpp for the scheme).  Included by fft.hip only.
#pragma once
#include "fft.hpp"

namespace dft {
namespace fft {

// ------------------------------------------------ z pass: real -> half spec --
// Pairs of adjacent real z-lines (2p, 2p+1) -> half spectra rows 2p, 2p+1.
template <typename T, int N>
__global__ void __launch_bounds__((Tile<T, N>::THREADS))
    k_zr2c(Batch b, int64_t npairs, int P, const cplx<T> *__restrict__ g_tw) {
  using C = cplx<T>;
  constexpr int TILE = Tile<T, N>::TILE, PITCH = Shape<N>::PITCH, BS = Tile<T, N>::THREADS;
  __shared__ C s_line[TILE * PITCH];
  __shared__ C s_tw[2 * N];
  const T *__restrict__ in = (const T *)b.in[blockIdx.y];
  C *__restrict__ out = (C *)b.out[blockIdx.y];
  const int64_t p0 = (int64_t)blockIdx.x * TILE;
  const int npair = (int)((npairs - p0) < TILE ? (npairs - p0) : TILE);
  load_twiddles<T, N>(s_tw, g_tw);
  const T *src = in + p0 * 2 * N;
  for (int e = threadIdx.x; e < npair * 2 * N; e += BS) {
    const int pr = e / (2 * N), w = e - pr * 2 * N, which = w / N, el = w - which * N;
    T *dst = (T *)&s_line[pr * PITCH + el];
    dst[which] = src[e];
  }
  __syncthreads();
  const int line = threadIdx.x % TILE, tl = threadIdx.x / TILE;
  line_fft<T, N, -1>(&s_line[line * PITCH], s_tw, tl, line < npair);
  C *dst = out + p0 * 2 * P;
  for (int e = threadIdx.x; e < npair * 2 * P; e += BS) {
    const int pr = e / (2 * P), w = e - pr * 2 * P, which = w / P, k = w - which * P;
    const C z = s_line[pr * PITCH + k], zm = s_line[pr * PITCH + ((N - k) & (N - 1))];
    C r;
    if (which == 0) r = cmk<C>((T)0.5 * (z.x + zm.x), (T)0.5 * (z.y - zm.y));
    else r = cmk<C>((T)0.5 * (z.y + zm.y), (T)0.5 * (zm.x - z.x));
    dst[e] = r;
  }
}

// ------------------------------------------------ z pass: half spec -> real --
template <typename T, int N>
__global__ void __launch_bounds__((Tile<T, N, 2>::THREADS))
    k_zc2r(Batch b, int64_t npairs, int P, const cplx<T> *__restrict__ g_tw, T scale) {
  using C = cplx<T>;
  constexpr int TILE = Tile<T, N, 2>::TILE, PITCH = Shape<N>::PITCH, BS = Tile<T, N, 2>::THREADS;
  constexpr int RP = N / 2 + 1;
  __shared__ C s_line[TILE * PITCH];
  __shared__ C s_raw[TILE * 2 * RP];
  __shared__ C s_tw[2 * N];
  const C *__restrict__ in = (const C *)b.in[blockIdx.y];
  T *__restrict__ out = (T *)b.out[blockIdx.y];
  const int64_t p0 = (int64_t)blockIdx.x * TILE;
  const int npair = (int)((npairs - p0) < TILE ? (npairs - p0) : TILE);
  load_twiddles<T, N>(s_tw, g_tw);
  const C *src = in + p0 * 2 * P;
  for (int e = threadIdx.x; e < npair * 2 * P; e += BS) {
    const int pr = e / (2 * P), w = e - pr * 2 * P, which = w / P, k = w - which * P;
    s_raw[(pr * 2 + which) * RP + k] = src[e];
  }
  __syncthreads();
  for (int m = threadIdx.x; m < npair * N; m += BS) {
    const int pr = m / N, idx = m - pr * N;
    C a, bb;
    if (idx <= N / 2) {
      a = s_raw[(pr * 2) * RP + idx];
      bb = s_raw[(pr * 2 + 1) * RP + idx];
    } else {
      a = cconj(s_raw[(pr * 2) * RP + N - idx]);
      bb = cconj(s_raw[(pr * 2 + 1) * RP + N - idx]);
    }
    s_line[pr * PITCH + idx] = cmk<C>(a.x - bb.y, a.y + bb.x);
  }
  __syncthreads();
  const int line = threadIdx.x % TILE, tl = threadIdx.x / TILE;
  line_fft<T, N, +1>(&s_line[line * PITCH], s_tw, tl, line < npair);
  T *dst = out + p0 * 2 * N;
  for (int e = threadIdx.x; e < npair * 2 * N; e += BS) {
    const int pr = e / (2 * N), w = e - pr * 2 * N, which = w / N, el = w - which * N;
    const C z = s_line[pr * PITCH + el];
    dst[e] = scale * (which ? z.y : z.x);
  }
}

// ------------------------------------------- strided c2c pass (y or x axis) --
// Line (o, c): element e at o*os + e*es + c, c in [0, P).  blockIdx.x = o*ntile + tile.
template <typename T, int N, int DIR>
__global__ void __launch_bounds__((Tile<T, N>::THREADS))
    k_strided(Batch b, int P, int64_t es, int64_t os, int ntile, const cplx<T> *__restrict__ g_tw, T scale) {
  using C = cplx<T>;
  constexpr int TILE = Tile<T, N>::TILE, PITCH = Shape<N>::PITCH, BS = Tile<T, N>::THREADS;
  __shared__ C s_line[TILE * PITCH];
  __shared__ C s_tw[2 * N];
  const C *__restrict__ in = (const C *)b.in[blockIdx.y];
  C *__restrict__ out = (C *)b.out[blockIdx.y];
  const int o = blockIdx.x / ntile, ct = blockIdx.x - o * ntile, c0 = ct * TILE;
  const int ncol = (P - c0) < TILE ? (P - c0) : TILE;
  load_twiddles<T, N>(s_tw, g_tw);
  const int64_t base = (int64_t)o * os + c0;
  for (int e = threadIdx.x; e < TILE * N; e += BS) {
    const int col = e % TILE, el = e / TILE;
    if (col < ncol) s_line[col * PITCH + el] = in[base + el * es + col];
  }
  __syncthreads();
  const int line = threadIdx.x % TILE, tl = threadIdx.x / TILE;
  line_fft<T, N, DIR>(&s_line[line * PITCH], s_tw, tl, line < ncol);
  for (int e = threadIdx.x; e < TILE * N; e += BS) {
    const int col = e % TILE, el = e / TILE;
    if (col < ncol) out[base + el * es + col] = cscale(s_line[col * PITCH + el], scale);
  }
}

// ------------------------------- inverse x pass with fused k-space combine --
// Spectra are [nx][ny][P]; x line (ky, kz) at ky*P + kx*ny*P + kz.
template <typename T, int N>
__global__ void __launch_bounds__((Tile<T, N>::THREADS))
    k_xcombine_inv(Batch b, CombineSet cs, KernTab kt, int ny, int P, int ntile, const cplx<T> *__restrict__ g_tw) {
  using C = cplx<T>;
  constexpr int TILE = Tile<T, N>::TILE, PITCH = Shape<N>::PITCH, BS = Tile<T, N>::THREADS;
  __shared__ C s_line[TILE * PITCH];
  __shared__ C s_tw[2 * N];
  const Combine cb = cs.c[blockIdx.y];
  C *__restrict__ out = (C *)b.out[blockIdx.y];
  const int o = blockIdx.x / ntile, ct = blockIdx.x - o * ntile, c0 = ct * TILE;
  const int ncol = (P - c0) < TILE ? (P - c0) : TILE;
  const int64_t es = (int64_t)ny * P;
  load_twiddles<T, N>(s_tw, g_tw);
  const int64_t base = (int64_t)o * P + c0;
  const T *sx = (const T *)kt.sinx, *sy = (const T *)kt.siny, *sz = (const T *)kt.sinz;
  for (int e = threadIdx.x; e < TILE * N; e += BS) {
    const int col = e % TILE, el = e / TILE;
    if (col >= ncol) continue;
    const int64_t gi = base + el * es + col;
    T accx = 0, accy = 0;
    for (int t = 0; t < cb.nterms; ++t) {
      const Term tm = cb.t[t];
      C v = ((const C *)b.in[tm.in])[gi];
      T s = (T)tm.scale;
      if (tm.kern >= 0) s *= ((const T *)kt.k[tm.kern])[gi];
      if (tm.deriv >= 0) {
        const T d = tm.deriv == 0 ? sx[el] : (tm.deriv == 1 ? sy[o] : sz[c0 + col]);
        s *= d;
        v = cmk<C>(-v.y, v.x);  // times i
      }
      accx += s * v.x;
      accy += s * v.y;
    }
    s_line[col * PITCH + el] = cmk<C>(accx, accy);
  }
  __syncthreads();
  const int line = threadIdx.x % TILE, tl = threadIdx.x / TILE;
  line_fft<T, N, +1>(&s_line[line * PITCH], s_tw, tl, line < ncol);
  for (int e = threadIdx.x; e < TILE * N; e += BS) {
    const int col = e % TILE, el = e / TILE;
    if (col < ncol) out[base + el * es + col] = s_line[col * PITCH + el];
  }
}

// ---------------------------------------- propagator passes (complex psi) --
// z pass over a full complex field, multiplying by e^{-iV dt/2} before (DIR<0)
// or after (DIR>0, together with `scale`) the transform.  With PHASE the
// `pot` array already holds e^{-iV dt/2}.  After the inverse pass each block
// writes sum |psi|^2 to partial[block] (FP64).
template <typename T, int N, int DIR, bool PHASE = false>
__global__ void __launch_bounds__((Tile<T, N>::THREADS))
    k_zc2c_pot(const cplx<T> *src, cplx<T> *dst, const cplx<T> *__restrict__ pot, int64_t nlines,
               const cplx<T> *__restrict__ g_tw, T dtr, T dti, T scale, double *__restrict__ partial) {
  using C = cplx<T>;
  constexpr int TILE = Tile<T, N>::TILE, PITCH = Shape<N>::PITCH, BS = Tile<T, N>::THREADS;
  __shared__ C s_line[TILE * PITCH];
  __shared__ C s_tw[2 * N];
  __shared__ double s_red[BS];
  const int64_t l0 = (int64_t)blockIdx.x * TILE;
  const int nl = (int)((nlines - l0) < TILE ? (nlines - l0) : TILE);
  load_twiddles<T, N>(s_tw, g_tw);
  const int64_t g0 = l0 * N;
  for (int e = threadIdx.x; e < nl * N; e += BS) {
    const int ln = e / N, el = e - ln * N;
    C v = src[g0 + e];
    if (DIR < 0 && pot) v = cmul(v, PHASE ? pot[g0 + e] : vphase<T>(pot[g0 + e], dtr, dti));
    s_line[ln * PITCH + el] = v;
  }
  __syncthreads();
  const int line = threadIdx.x % TILE, tl = threadIdx.x / TILE;
  line_fft<T, N, DIR>(&s_line[line * PITCH], s_tw, tl, line < nl);
  double acc = 0.0;
  for (int e = threadIdx.x; e < nl * N; e += BS) {
    const int ln = e / N, el = e - ln * N;
    C v = s_line[ln * PITCH + el];
    if (DIR > 0) {
      v = cscale(v, scale);
      if (pot) v = cmul(v, PHASE ? pot[g0 + e] : vphase<T>(pot[g0 + e], dtr, dti));
      acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    dst[g0 + e] = v;
  }
  if (DIR > 0 && partial) {
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s_red[0];
  }
}

// Fused x pass of the propagator: forward x FFT, kinetic phase
// exp(-i k^2 dt / 2M), inverse x FFT.  Field is [nx][ny][nz] complex.
template <typename T, int N>
__global__ void __launch_bounds__((Tile<T, N>::THREADS))
    k_xkin(cplx<T> *__restrict__ psi, int ny, int nz, int ntile, const cplx<T> *__restrict__ g_tw,
           T cx, T cy, T cz, T dtr, T dti) {
  using C = cplx<T>;
  constexpr int TILE = Tile<T, N>::TILE, PITCH = Shape<N>::PITCH, BS = Tile<T, N>::THREADS;
  __shared__ C s_line[TILE * PITCH];
  __shared__ C s_tw[2 * N];
  const int o = blockIdx.x / ntile, ct = blockIdx.x - o * ntile, c0 = ct * TILE;
  const int ncol = (nz - c0) < TILE ? (nz - c0) : TILE;
  const int64_t es = (int64_t)ny * nz, base = (int64_t)o * nz + c0;
  load_twiddles<T, N>(s_tw, g_tw);
  for (int e = threadIdx.x; e < TILE * N; e += BS) {
    const int col = e % TILE, el = e / TILE;
    if (col < ncol) s_line[col * PITCH + el] = psi[base + el * es + col];
  }
  __syncthreads();
  const int line = threadIdx.x % TILE, tl = threadIdx.x / TILE;
  line_fft<T, N, -1>(&s_line[line * PITCH], s_tw, tl, line < ncol);
  // kinetic phase: K = cx*mx^2 + cy*my^2 + cz*mz^2 (signed frequency indices)
  const int my = o < ny / 2 ? o : o - ny;
  for (int e = threadIdx.x; e < TILE * N; e += BS) {
    const int col = e % TILE, el = e / TILE;
    if (col >= ncol) continue;
    const int mx = el < N / 2 ? el : el - N;
    const int kz = c0 + col, mz = kz < nz / 2 ? kz : kz - nz;
    const T K = cx * (T)(mx * mx) + cy * (T)(my * my) + cz * (T)(mz * mz);
    // exp(-i K dt) = exp(K dti) (cos(K dtr) - i sin(K dtr))
    T s, c;
    dsincos(K * dtr, &s, &c);
    const T m = dexp(K * dti);
    s_line[col * PITCH + el] = cmul(s_line[col * PITCH + el], cmk<C>(m * c, -m * s));
  }
  __syncthreads();
  line_fft<T, N, +1>(&s_line[line * PITCH], s_tw, tl, line < ncol);
  for (int e = threadIdx.x; e < TILE * N; e += BS) {
    const int col = e % TILE, el = e / TILE;
    if (col < ncol) psi[base + el * es + col] = s_line[col * PITCH + el];
  }
}

}  // namespace fft
}  // namespace dft
