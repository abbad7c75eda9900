// Orsay-Trento functional on gfx950: kernel spectra, fused pointwise stages
// and the FFT schedule of dft_ot_potential / dft_ot_energy_density.
//
// Schedule of one full-OT potential (PLAIN|KC|BACKFLOW), src/ot.c:314-991:
//   P1   psi -> rho, v_d, q_d = rho v_d, q4 = rho v^2           (one stencil pass)
//   F1   r2c {rho, q_x, q_y, q_z, q4}
//   I1   c2r {Pi*rho, F*rho, dF_d*rho (3), U*rho, U*q_d (3), U*q4}
//   P2   local terms, Q = (c2 rhobar + c3 rhobar^2) rho, G_d = d_d rho (1 - F*rho/rho_0s),
//        BF real part, W_d = v_d A - B_d, (d_d rho/(rho+eps)) W_d
//   F2   r2c {Q, G_x, G_y, G_z}
//   I2   c2r {J_d = F*G_d (3), T1 = F*sum_d d_d G_d}
//   P3   H = sum_d d_d rho J_d, KC terms 1 and 3, BF sum_d d_d W_d
//   F3   r2c {H}
//   I3   c2r {V_LJ*rho + Pi*Q + k23 c F*H}  (all linear terms summed in k space)
//   P4   V += ...
// i.e. 10 forward + 15 inverse real 3-D transforms, against 34 in the
// reference.  The derivative kernels dF_d are never stored: their spectra
// are F^ * i sin(k_d h)/h, the symbol of the periodic central difference
// the reference applies to F before transforming (src/ot.c:196).
#include <cmath>
#include <cstring>
#include <vector>

#include "ideal_gas.hpp"
#include "internal.hpp"
#include "ot_common.hpp"
#include "ot_fused.hpp"

using namespace dft;

namespace dft {

#define GRID_STRIDE(i, n) \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// ------------------------------------------------------- kernel functions --
struct KParams {
  double lj_h, lj_sigma, lj_eps, lj_cval;
  double sa_radius;
  double g_inv_width;
  double g11, g12, g21, g22, a1, a2;
};

// src/common.c:31-66 (LJ, zero core), 131-138 (spherical average),
// 193-200 (gaussian F), src/ot-private.h:11-22 (backflow U_J).
__device__ inline double kfun(const KParams &p, int which, double x, double y, double z) {
#pragma clang fp contract(off)
  const double r2 = x * x + y * y + z * z;
  if (which == 0) {
    if (r2 <= p.lj_h * p.lj_h) return p.lj_cval;
    double s = p.lj_sigma * p.lj_sigma / r2;
    s = s * s * s;
    return 4.0 * p.lj_eps * s * (s - 1.0);
  } else if (which == 1) {
    const double h = p.sa_radius;
    return (r2 <= h * h) ? 3.0 / (4.0 * M_PI * h * h * h) : 0.0;
  } else if (which == 2) {
    const double iw = p.g_inv_width;
    double nrm = 0.5 * (2.0 / sqrt(M_PI)) * iw;
    nrm = nrm * nrm * nrm;
    return nrm * exp(-r2 * iw * iw);
  }
  return (p.g11 + p.g12 * r2) * exp(-p.a1 * r2) + (p.g21 + p.g22 * r2) * exp(-p.a2 * r2);
}

// rgrid_adaptive_map (libgrid): cell average over an ns^3 midpoint sub-grid,
// ns = min_sub, 2 min_sub, ... max_sub, stopping when two successive averages
// differ by less than tol.  Sampled about the index origin (wrapped
// coordinates), which equals the reference's r = 0 at n/2 with the
// (-1)^(i+j+k) factor of rgrid_fft_convolute.
__global__ void k_kernel_map(double *out, int nx, int ny, int nz, double h, int which, KParams kp, int min_sub,
                             int max_sub, double tol) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const double x = (double)(i < nx / 2 ? i : i - nx) * h, y = (double)(j < ny / 2 ? j : j - ny) * h,
                 z = (double)(k < nz / 2 ? k : k - nz) * h;
    double prev = 0.0;
    for (int ns = min_sub; ns <= max_sub; ns *= 2) {
      double s = 0.0;
      for (int a = 0; a < ns; ++a) {
        const double xs = x + (((double)a + 0.5) / (double)ns - 0.5) * h;
        for (int b = 0; b < ns; ++b) {
          const double ys = y + (((double)b + 0.5) / (double)ns - 0.5) * h;
          for (int cc = 0; cc < ns; ++cc) {
            const double zs = z + (((double)cc + 0.5) / (double)ns - 0.5) * h;
            s += kfun(kp, which, xs, ys, zs);
          }
        }
      }
      const double cur = s / ((double)ns * (double)ns * (double)ns);
      const bool done = (ns > min_sub) && fabs(cur - prev) < tol;
      prev = cur;
      if (done) break;
    }
    out[c] = prev;
  }
}

template <typename T> __global__ void k_spec_real(const double2 *spec, T *out, int64_t nh, double scale) {
  GRID_STRIDE(i, nh) out[i] = (T)(spec[i].x * scale);
}

// P1: psi -> rho and, for backflow, v_d (grid_wf_velocity), q_d = v_d rho,
// q4 = |v|^2 rho (src/ot.c:385-396, 852-891); with the HD / HD2 backflow
// switch rho~ = G rho (stored in rt) takes rho's place in q_d and q4.
template <typename T, bool BF>
__global__ void k_prep(const cplx<T> *__restrict__ psi, T *__restrict__ rho, T *__restrict__ vx, T *__restrict__ vy,
                       T *__restrict__ vz, T *__restrict__ qx, T *__restrict__ qy, T *__restrict__ qz,
                       T *__restrict__ q4, T *__restrict__ rt, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    const cplx<T> s = psi[c];
    const T r = s.x * s.x + s.y * s.y;
    rho[c] = r;
    if constexpr (BF) {
      const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
      const T den = pw.invM / (r + pw.eps_v);
      T v[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const cplx<T> a = psi[nb.p[d]], b = psi[nb.m[d]];
        const T dx = (a.x - b.x) * pw.inv2h, dy = (a.y - b.y) * pw.inv2h;
        v[d] = (s.x * dy - s.y * dx) * den;
        if (pw.clamp_v) v[d] = clampv(v[d], pw.vmax);
      }
      vx[c] = v[0];
      vy[c] = v[1];
      vz[c] = v[2];
      const T w = pw.bfhd ? bf_rt(r, pw) : r;
      if (pw.bfhd) rt[c] = w;
      qx[c] = v[0] * w;
      qy[c] = v[1] * w;
      qz[c] = v[2] * w;
      q4[c] = (v[2] * v[2] + v[1] * v[1] + v[0] * v[0]) * w;
    }
  }
}

// P1 for given density and velocities (dft_ot_potential2, dft_ot_backflow_potential).
template <typename T>
__global__ void k_prep_v(const T *__restrict__ rho, T *vx, T *vy, T *vz, T *__restrict__ qx, T *__restrict__ qy,
                         T *__restrict__ qz, T *__restrict__ q4, T *__restrict__ rt, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    const T r0 = rho[c];
    const T r = pw.bfhd ? bf_rt(r0, pw) : r0;
    if (pw.bfhd) rt[c] = r;
    T v[3] = {vx[c], vy[c], vz[c]};
    if (pw.clamp_v) {
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = clampv(v[d], pw.vmax);
      vx[c] = v[0];
      vy[c] = v[1];
      vz[c] = v[2];
    }
    qx[c] = v[0] * r;
    qy[c] = v[1] * r;
    qz[c] = v[2] * r;
    q4[c] = (v[2] * v[2] + v[1] * v[1] + v[0] * v[0]) * r;
  }
}

struct StageBIn {
  const void *rho, *rhobar, *rhot, *vx, *vy, *vz, *A, *Bx, *By, *Bz, *C;
};
struct StageBOut {
  void *Q, *Gx, *Gy, *Gz, *Vre, *Vim, *Wx, *Wy, *Wz;
};

// P2: src/ot.c:547-586 (local terms), 604-656 (G_d), 893-961 (BF real part,
// W_d and the (d rho/(rho+eps)) W_d half of the imaginary part; with the
// HD / HD2 backflow switch the real part carries d(G rho)/d rho and the
// gradient is that of G rho).
template <typename T, bool LOC, bool KC, bool BF>
__global__ void k_stageB(StageBIn in, StageBOut out, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  const T *rho = (const T *)in.rho;
  GRID_STRIDE(c, n) {
    const T r = rho[c];
    T vre = 0;
    if constexpr (LOC) {
      const T rb = ((const T *)in.rhobar)[c];
      vre = rb * rb * (pw.c2 * (T)0.5) + rb * rb * rb * (pw.c3 / (T)3);
      if (pw.hdC != (T)0) vre += hd_pot(r, pw);
      ((T *)out.Q)[c] = (rb * pw.c2 + rb * rb * pw.c3) * r;
    }
    T g[3] = {0, 0, 0};
    if constexpr (KC || BF) {
      const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
#pragma unroll
      for (int d = 0; d < 3; ++d) g[d] = (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h;
    }
    if constexpr (KC) {
      const T rst = ((const T *)in.rhot)[c] * (-pw.inv_rho0s) + (T)1;
      ((T *)out.Gx)[c] = g[0] * rst;
      ((T *)out.Gy)[c] = g[1] * rst;
      ((T *)out.Gz)[c] = g[2] * rst;
    }
    if constexpr (BF) {
      const T vx = ((const T *)in.vx)[c], vy = ((const T *)in.vy)[c], vz = ((const T *)in.vz)[c];
      const T A = ((const T *)in.A)[c], Bx = ((const T *)in.Bx)[c], By = ((const T *)in.By)[c],
              Bz = ((const T *)in.Bz)[c], Cc = ((const T *)in.C)[c];
      T t = (vz * vz + vy * vy + vx * vx) * A;
      t += (T)-2 * vx * Bx;
      t += (T)-2 * vy * By;
      t += (T)-2 * vz * Bz;
      t = t + Cc;
      if (pw.bfhd) t = t * bf_dg(r, pw);
      T re = t * ((T)-0.5 * pw.mass);
      if (pw.clamp_bf) re = clampv(re, pw.bfmax);
      vre += re * pw.cbf;
      const T Wx = vx * A - Bx, Wy = vy * A - By, Wz = vz * A - Bz;
      ((T *)out.Wx)[c] = Wx;
      ((T *)out.Wy)[c] = Wy;
      ((T *)out.Wz)[c] = Wz;
      if (pw.bfhd) {
        const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
#pragma unroll
        for (int d = 0; d < 3; ++d) g[d] = (bf_rt(rho[nb.p[d]], pw) - bf_rt(rho[nb.m[d]], pw)) * pw.inv2h;
      }
      const T inv = (T)1 / (r + pw.div_eps);
      T im = (T)0.5 * (g[0] * inv) * Wx;
      im += (T)0.5 * (g[1] * inv) * Wy;
      im += (T)0.5 * (g[2] * inv) * Wz;
      ((T *)out.Vim)[c] = im;
    }
    ((T *)out.Vre)[c] = vre;
  }
}

struct StageCIn {
  const void *rho, *rhot, *T1, *Jx, *Jy, *Jz, *Px, *Py, *Pz, *Wx, *Wy, *Wz;
};
struct StageCOut {
  void *H, *Vre, *Vim;
};

// P3: src/ot.c:640-675 (KC terms 1 and 3, H_d = d_d rho J_d summed over d
// because F*(sum_d H_d) = sum_d F*H_d), 963-990 (BF d_d W_d half of Im V,
// times G(rho) with the HD / HD2 backflow switch).
template <typename T, bool KC, bool BF>
__global__ void k_stageC(StageCIn in, StageCOut out, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    if constexpr (KC) {
      const T *rho = (const T *)in.rho;
      const T rst = ((const T *)in.rhot)[c] * (-pw.inv_rho0s) + (T)1;
      const T J[3] = {((const T *)in.Jx)[c], ((const T *)in.Jy)[c], ((const T *)in.Jz)[c]};
      const T Pd[3] = {((const T *)in.Px)[c], ((const T *)in.Py)[c], ((const T *)in.Pz)[c]};
      T H = 0, t3 = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T g = (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h;
        H += g * J[d];
        t3 += Pd[d] * J[d];
      }
      ((T *)out.H)[c] = H;
      ((T *)out.Vre)[c] += ((const T *)in.T1)[c] * rst * pw.ckc - t3 * (pw.ckc * pw.k23);
    }
    if constexpr (BF) {
      const T *W[3] = {(const T *)in.Wx, (const T *)in.Wy, (const T *)in.Wz};
      T im = ((const T *)out.Vim)[c];
      if (pw.bfhd) {
        const T gw = bf_g(((const T *)in.rho)[c], pw);
#pragma unroll
        for (int d = 0; d < 3; ++d) im += (T)0.5 * (((W[d][nb.p[d]] - W[d][nb.m[d]]) * pw.inv2h) * gw);
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) im += (T)0.5 * ((W[d][nb.p[d]] - W[d][nb.m[d]]) * pw.inv2h);
      }
      if (pw.clamp_bf) im = clampv(im, pw.bfmax);
      ((T *)out.Vim)[c] = im * pw.cbf;
    }
  }
}

// P4: V += (Vre + lin) + i Vim
template <typename T>
__global__ void k_final(cplx<T> *V, const T *__restrict__ vre, const T *__restrict__ lin, const T *__restrict__ vim,
                        int64_t n) {
  GRID_STRIDE(c, n) {
    cplx<T> v = V[c];
    T add = vre ? vre[c] : (T)0;
    if (lin) add += lin[c];
    v.x += add;
    if (vim) v.y += vim[c];
    V[c] = v;
  }
}

// Ideal Bose gas of the thermal models (src/ot.c:402-407, src/ot-energy.c:101-104),
// in FP64: out += the potential (MODE 0), the energy per volume (1), or its
// derivative for the energy gradient (2); with MODE 3, out += vbar dV/drho
// (the transposed Jacobian of the potential).
template <typename T, int MODE>
__global__ void k_ideal_gas(T *out, const T *__restrict__ rho, const T *__restrict__ vbar, ig::Params p,
                            int64_t n) {
  GRID_STRIDE(c, n) {
    const double r = (double)rho[c];
    double v;
    if constexpr (MODE == 0) v = ig::potential(r, p);
    else if constexpr (MODE == 1) v = ig::energy(r, p);
    else if constexpr (MODE == 2) v = ig::energy_derivative(r, p);
    else v = (double)vbar[c] * ig::potential_derivative(r, p);
    out[c] += (T)v;
  }
}

// Energy density, src/ot-energy.c:64-93 (plain), 202-288 (BF), and the
// G_d of 144-161 (KC).  acc: add to eps instead of overwriting it.
struct EnergyIn {
  const void *rho, *ljr, *rhobar, *rhot, *vx, *vy, *vz, *A, *Bx, *By, *Bz, *C;
};
template <typename T, bool LOC, bool KC, bool BF>
__global__ void k_energyB(EnergyIn in, T *eps, T *Gx, T *Gy, T *Gz, PW<T> pw, int acc) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  const T *rho = (const T *)in.rho;
  GRID_STRIDE(c, n) {
    const T r = rho[c];
    T e = acc ? eps[c] : (T)0;
    if constexpr (LOC) {
      const T rb = ((const T *)in.rhobar)[c];
      e += (T)0.5 * r * ((const T *)in.ljr)[c];
      e += (pw.c2 * (T)0.5) * (rb * rb * r);
      e += (pw.c3 / (T)3) * (rb * rb * rb * r);
      if (pw.hdC != (T)0) e += hd_eps(r, pw);
    }
    if constexpr (KC) {
      const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
      const T rst = ((const T *)in.rhot)[c] * (-pw.inv_rho0s) + (T)1;
      Gx[c] = (rho[nb.p[0]] - rho[nb.m[0]]) * pw.inv2h * rst;
      Gy[c] = (rho[nb.p[1]] - rho[nb.m[1]]) * pw.inv2h * rst;
      Gz[c] = (rho[nb.p[2]] - rho[nb.m[2]]) * pw.inv2h * rst;
    }
    if constexpr (BF) {
      // G rho in place of rho with the HD / HD2 backflow switch (src/ot-energy.c:221-227)
      const T rw = pw.bfhd ? bf_rt(r, pw) : r;
      const T vx = ((const T *)in.vx)[c], vy = ((const T *)in.vy)[c], vz = ((const T *)in.vz)[c];
      const T v2 = (vx * vx + vy * vy) + vz * vz;
      const T mq = pw.cbf * pw.mass * (T)0.25, mh = pw.cbf * pw.mass * (T)0.5;
      e += -mq * (((const T *)in.A)[c] * v2 * rw);
      e += mh * (((const T *)in.Bx)[c] * rw * vx);
      e += mh * (((const T *)in.By)[c] * rw * vy);
      e += mh * (((const T *)in.Bz)[c] * rw * vz);
      e += -mq * (((const T *)in.C)[c] * rw);
    }
    eps[c] = e;
  }
}

// KC energy: eps += -(alpha_s / 4M) sum_d J_d G_d   (src/ot-energy.c:163-190)
template <typename T>
__global__ void k_energyC(T *eps, const T *__restrict__ Jx, const T *__restrict__ Jy, const T *__restrict__ Jz,
                          const T *__restrict__ Gx, const T *__restrict__ Gy, const T *__restrict__ Gz, T s,
                          int64_t n) {
  GRID_STRIDE(c, n) {
    T e = eps[c];
    e += s * (Jx[c] * Gx[c]);
    e += s * (Jy[c] * Gy[c]);
    e += s * (Jz[c] * Gz[c]);
    eps[c] = e;
  }
}

// Backward of E[rho], KC part: H = sum_d (d_d rho) J_d and M_d = rho_st J_d.
template <typename T>
__global__ void k_grad_kc(const T *__restrict__ rho, const T *__restrict__ rhot, const T *__restrict__ Jx,
                          const T *__restrict__ Jy, const T *__restrict__ Jz, T *__restrict__ H, T *__restrict__ Mx,
                          T *__restrict__ My, T *__restrict__ Mz, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    const T rst = rhot[c] * (-pw.inv_rho0s) + (T)1;
    const T J[3] = {Jx[c], Jy[c], Jz[c]};
    T h = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) h += (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h * J[d];
    H[c] = h;
    Mx[c] = rst * J[0];
    My[c] = rst * J[1];
    Mz[c] = rst * J[2];
  }
}

// grad = local + linear terms + c sum_d D_d M_d (M = null: no KC).
template <typename T>
__global__ void k_grad_final(T *__restrict__ grad, const T *__restrict__ vre, const T *__restrict__ lin,
                             const T *__restrict__ Mx, const T *__restrict__ My, const T *__restrict__ Mz, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    T g = vre[c] + lin[c];
    if (Mx) {
      const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
      const T *M[3] = {Mx, My, Mz};
      T div = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) div += (M[d][nb.p[d]] - M[d][nb.m[d]]) * pw.inv2h;
      g += pw.ckc * div;
    }
    grad[c] = g;
  }
}


// ------------------------------------------------- adjoint (reverse mode) --
// Reverse pass through V_OT(rho) for the PLAIN and KC models: given Vbar, the
// gradient of a scalar with respect to the real potential, accumulate
// R = (dV/drho)^T Vbar.  Each forward stage is transposed in reverse order:
// convolutions with symmetric kernels are self-transposed, the periodic
// central difference D_d transposes to -D_d, and products pass the adjoint
// to each factor.  Forward (what the unfused potential path computes):
//   rb = Pi*rho, rt = F*rho, rst = 1 - rt/rho0s, g_d = D_d rho, G_d = g_d rst,
//   J_d = F*G_d, P_d = D_d rt, T1 = sum_d D_d J_d, H = sum_d g_d J_d,
//   V = L*rho + c2/2 rb^2 + c3/3 rb^3 + Pi*[(c2 rb + c3 rb^2) rho]
//       + ckc [rst T1 - k23 sum_d P_d J_d + k23 F*H].
struct VjpA {
  const void *vbar, *rho, *rb, *qb, *lv, *hp, *rt, *t1, *J[3], *Pd[3];
  void *rp, *rbb, *t1b, *pb[3], *jb[3], *gb[3], *rsb;
};

// Pointwise adjoints of the local terms and of the KC products.
// qb = Pi*Vbar, lv = L*Vbar, hp = F*Vbar.
template <typename T, bool KC> __global__ void k_vjp_a(VjpA a, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    const T vb = ((const T *)a.vbar)[c], r = ((const T *)a.rho)[c], rb = ((const T *)a.rb)[c];
    const T qb = ((const T *)a.qb)[c];
    const T dl = pw.c2 * rb + pw.c3 * rb * rb;  // d(local V)/d rb, and Q/rho
    ((T *)a.rbb)[c] = vb * dl + qb * r * (pw.c2 + (T)2 * pw.c3 * rb);
    ((T *)a.rp)[c] = ((const T *)a.lv)[c] + qb * dl + (pw.hdC != (T)0 ? vb * hd_dpot(r, pw) : (T)0);
    if constexpr (KC) {
      const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
      const T *rho = (const T *)a.rho;
      const T rst = ((const T *)a.rt)[c] * (-pw.inv_rho0s) + (T)1;
      const T ck = pw.ckc, ckk = pw.ckc * pw.k23;
      const T Hb = ckk * ((const T *)a.hp)[c];
      ((T *)a.t1b)[c] = ck * vb * rst;
      ((T *)a.rsb)[c] = ck * vb * ((const T *)a.t1)[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T g = (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h;
        const T J = ((const T *)a.J[d])[c], P = ((const T *)a.Pd[d])[c];
        ((T *)a.pb[d])[c] = -ckk * vb * J;
        ((T *)a.jb[d])[c] = -ckk * vb * P + Hb * g;
        ((T *)a.gb[d])[c] = Hb * J;
      }
    }
  }
}

// Jbar_d -= D_d T1bar (adjoint of T1 = sum_d D_d J_d).
template <typename T> __global__ void k_vjp_b(VjpA a, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  const T *t1b = (const T *)a.t1b;
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
#pragma unroll
    for (int d = 0; d < 3; ++d) ((T *)a.jb[d])[c] -= (t1b[nb.p[d]] - t1b[nb.m[d]]) * pw.inv2h;
  }
}

// With Gbar_d = F*Jbar_d (in a.J): gbar_d += Gbar_d rst, rstbar += sum_d Gbar_d g_d,
// rtbar = -rstbar / rho0s (written to a.rsb).
template <typename T> __global__ void k_vjp_c(VjpA a, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  const T *rho = (const T *)a.rho;
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    const T rst = ((const T *)a.rt)[c] * (-pw.inv_rho0s) + (T)1;
    T rsb = ((const T *)a.rsb)[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T g = (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h;
      const T Gb = ((const T *)a.J[d])[c];
      ((T *)a.gb[d])[c] += Gb * rst;
      rsb += Gb * g;
    }
    ((T *)a.rsb)[c] = -rsb * pw.inv_rho0s;
  }
}

// rtbar -= sum_d D_d Pbar_d (adjoint of P_d = D_d rt); Rpart -= sum_d D_d gbar_d (g_d = D_d rho).
template <typename T> __global__ void k_vjp_d(VjpA a, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    T rt = ((const T *)a.rsb)[c], rp = ((const T *)a.rp)[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T *pb = (const T *)a.pb[d], *gb = (const T *)a.gb[d];
      rt -= (pb[nb.p[d]] - pb[nb.m[d]]) * pw.inv2h;
      rp -= (gb[nb.p[d]] - gb[nb.m[d]]) * pw.inv2h;
    }
    ((T *)a.rsb)[c] = rt;
    ((T *)a.rp)[c] = rp;
  }
}

// R = rp + conv; out = R (rho adjoint), or psibar += 2 R psi when psi is given.
template <typename T>
__global__ void k_vjp_out(T *out, cplx<T> *psib, const cplx<T> *psi, const T *rp, const T *conv, int64_t n) {
  GRID_STRIDE(c, n) {
    const T R = rp[c] + (conv ? conv[c] : (T)0);
    if (psib) {
      cplx<T> b = psib[c];
      b.x += (T)2 * R * psi[c].x;
      b.y += (T)2 * R * psi[c].y;
      psib[c] = b;
    } else {
      out[c] = R;
    }
  }
}

// ----------------------------------------- backflow adjoint (reverse mode) --
// Reverse pass through the backflow potential of src/ot.c:840-991 at the state
// whose intermediates potential_impl left in the slots:
//   v_d (clamped to max_veloc), A = U*rho, B_d = U*(rho v_d), C = U*(rho v^2),
//   t = v^2 A - 2 v.B + C,  Re V += cbf clamp(-(m/2) t),
//   W_d = v_d A - B_d,  g_d = D_d rho,
//   Im V = cbf clamp(1/2 sum_d [(g_d / (rho + eps)) W_d + D_d W_d]).
// Given Vr_bar and Vi_bar, the adjoints of Re V and Im V, the passes below
// accumulate rho_bar and v_bar; clamped points pass no adjoint.
struct BfAdj {
  const void *vrb, *vib;          // adjoints of Re V and Im V
  const void *rho, *v[3], *A, *W[3];
  void *tb, *imb;                 // -(m/2) cbf Vr_bar (masked), (cbf/2) Vi_bar (masked)
  void *Ab, *Bb[3], *gb[3], *vb[3], *rb;
  const void *uA, *uB[3], *uC;    // U*Abar, U*Bbar_d, U*Cbar (Cbar = tb)
};

// d^2(G rho)/d rho^2 = 2 G' + rho G'' of the backflow density switch, with
// G' = -2 xi e/(1+e)^2 and G'' = -4 xi^2 e (1-e)/(1+e)^3 for x <= 0, its
// negative for x > 0 (e = e^{-2|x|}).
template <typename T> __device__ inline T bf_d2(T r, const PW<T> &pw) {
  const T e = bf_e(r, pw), q = (T)1 + e;
  const T g1 = (T)-2 * pw.bfXi * e / (q * q);
  T g2 = (T)-4 * pw.bfXi * pw.bfXi * e * ((T)1 - e) / (q * q * q);
  if (r > pw.bfRho) g2 = -g2;
  return (T)2 * g1 + r * g2;
}
// G' alone
template <typename T> __device__ inline T bf_g1(T r, const PW<T> &pw) {
  const T e = bf_e(r, pw), q = (T)1 + e;
  return (T)-2 * pw.bfXi * e / (q * q);
}

// Pass 1: tb and imb, each zero where the forward clamp (max_bfpot) was
// active; the values before clamping are rebuilt from v, A, B_d, C and W_d.
// With the HD / HD2 backflow switch (rho~ = G rho; A, B_d, C from rho~;
// Re V_BF ~ t d(rho~)/d rho; Im V_BF's gradient half from rho~, its divergence
// half times G) tb also carries d(rho~)/d rho, and rb starts from the terms
// through the density dependence of those factors: Re_bar t d^2(rho~)/d rho^2
// and Im_bar G' sum_d D_d W_d (rb = 0 without the switch).
template <typename T>
__global__ void k_bfadj_a(BfAdj a, const T *__restrict__ B0, const T *__restrict__ B1, const T *__restrict__ B2,
                          const T *__restrict__ Cf, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  const T *rho = (const T *)a.rho;
  GRID_STRIDE(c, n) {
    const T v[3] = {((const T *)a.v[0])[c], ((const T *)a.v[1])[c], ((const T *)a.v[2])[c]};
    const T A = ((const T *)a.A)[c];
    T tb = ((const T *)a.vrb)[c] * pw.cbf * ((T)-0.5 * pw.mass);
    T ib = ((const T *)a.vib)[c] * pw.cbf * (T)0.5;
    T rb = 0;
    if (pw.clamp_bf || pw.bfhd) {
      T t = (v[2] * v[2] + v[1] * v[1] + v[0] * v[0]) * A;
      t += (T)-2 * v[0] * B0[c];
      t += (T)-2 * v[1] * B1[c];
      t += (T)-2 * v[2] * B2[c];
      t = t + Cf[c];
      const T r = rho[c];
      const T dg = pw.bfhd ? bf_dg(r, pw) : (T)1;
      if (pw.clamp_bf && fabs(t * dg * ((T)-0.5 * pw.mass)) > pw.bfmax) tb = 0;
      const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
      const T inv = (T)1 / (r + pw.div_eps);
      T im = 0, div = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T *W = (const T *)a.W[d];
        const T gr = pw.bfhd ? bf_rt(rho[nb.p[d]], pw) - bf_rt(rho[nb.m[d]], pw) : rho[nb.p[d]] - rho[nb.m[d]];
        im += (T)0.5 * (gr * pw.inv2h * inv) * W[c];
        div += (T)0.5 * ((W[nb.p[d]] - W[nb.m[d]]) * pw.inv2h);
      }
      im += pw.bfhd ? bf_g(r, pw) * div : div;
      if (pw.clamp_bf && fabs(im) > pw.bfmax) ib = 0;
      if (pw.bfhd) {
        rb = tb * t * bf_d2(r, pw) + ib * (T)2 * bf_g1(r, pw) * div;
        tb = tb * dg;
      }
    }
    ((T *)a.tb)[c] = tb;
    ((T *)a.imb)[c] = ib;
    ((T *)a.rb)[c] = rb;
  }
}

// Pass 2 (stencils of imb and rho): Wbar_d = imb g_d/(rho+eps) - D_d imb, and
// the adjoints of t and of the W_d products (g_d the central difference of
// rho, or of rho~ with the switch, which also weights D_d imb by G).
template <typename T> __global__ void k_bfadj_b(BfAdj a, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  const T *rho = (const T *)a.rho, *imb = (const T *)a.imb;
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    const T r = rho[c], tb = ((const T *)a.tb)[c], ib = imb[c], A = ((const T *)a.A)[c];
    const T inv = (T)1 / (r + pw.div_eps);
    T Ab = 0, rb = ((const T *)a.rb)[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T v = ((const T *)a.v[d])[c], W = ((const T *)a.W[d])[c];
      T g, Wb;
      if (pw.bfhd) {
        const T rp = rho[nb.p[d]], rm = rho[nb.m[d]];
        g = (bf_rt(rp, pw) - bf_rt(rm, pw)) * pw.inv2h;
        Wb = ib * g * inv - (bf_g(rp, pw) * imb[nb.p[d]] - bf_g(rm, pw) * imb[nb.m[d]]) * pw.inv2h;
      } else {
        g = (rho[nb.p[d]] - rho[nb.m[d]]) * pw.inv2h;
        Wb = ib * g * inv - (imb[nb.p[d]] - imb[nb.m[d]]) * pw.inv2h;
      }
      ((T *)a.vb[d])[c] = (T)2 * tb * W + Wb * A;
      ((T *)a.Bb[d])[c] = (T)-2 * tb * v - Wb;
      ((T *)a.gb[d])[c] = ib * W * inv;
      Ab += tb * v * v + Wb * v;
      rb -= ib * g * W * inv * inv;
    }
    ((T *)a.Ab)[c] = Ab;
    ((T *)a.rb)[c] = rb;
  }
}

// Pass 3 (stencil of gbar), after the convolutions uA = U*Abar, uB_d = U*Bbar_d,
// uC = U*tb: rho_bar and v_bar of the backflow part, v_bar masked by the velocity
// clamp.  With the switch the terms through A, B_d, C and the gradient of rho~
// are the adjoint of rho~ and reach rho through d(rho~)/d rho.
template <typename T> __global__ void k_bfadj_c(BfAdj a, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    const T r0 = ((const T *)a.rho)[c], uC = ((const T *)a.uC)[c];
    if (pw.bfhd) {
      const T r = bf_rt(r0, pw);
      T rt = ((const T *)a.uA)[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T v = ((const T *)a.v[d])[c], uB = ((const T *)a.uB[d])[c];
        const T *gb = (const T *)a.gb[d];
        rt += uC * v * v + uB * v - (gb[nb.p[d]] - gb[nb.m[d]]) * pw.inv2h;
        T vb = ((const T *)a.vb[d])[c] + uB * r + (T)2 * uC * r * v;
        if (pw.clamp_v && fabs(v) >= pw.vmax) vb = 0;
        ((T *)a.vb[d])[c] = vb;
      }
      ((T *)a.rb)[c] = ((const T *)a.rb)[c] + bf_dg(r0, pw) * rt;
      continue;
    }
    const T r = r0;
    T rb = ((const T *)a.rb)[c] + ((const T *)a.uA)[c];
    T v2 = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T v = ((const T *)a.v[d])[c];
      v2 += v * v;
    }
    rb += uC * v2;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T *gb = (const T *)a.gb[d];
      const T v = ((const T *)a.v[d])[c], uB = ((const T *)a.uB[d])[c];
      rb -= (gb[nb.p[d]] - gb[nb.m[d]]) * pw.inv2h;
      rb += uB * v;
      T vb = ((const T *)a.vb[d])[c] + uB * r + (T)2 * uC * r * v;
      if (pw.clamp_v && fabs(v) >= pw.vmax) vb = 0;
      ((T *)a.vb[d])[c] = vb;
    }
    ((T *)a.rb)[c] = rb;
  }
}

// psib += 2 rb psi + the transpose of the velocity v_d = Im(psi* D_d psi) / (M (|psi|^2 + eps))
// applied to vb_d: at the point, v_bar d v/d psi; through D_d psi, -D_d e_d with
// e_d = vb_d i psi / (M (|psi|^2 + eps)) evaluated at the neighbours.
template <typename T>
__global__ void k_bfadj_psi(cplx<T> *psib, const cplx<T> *__restrict__ psi, const T *__restrict__ rb, const T *vb0,
                            const T *vb1, const T *vb2, PW<T> pw) {
  const int64_t n = (int64_t)pw.nx * pw.ny * pw.nz;
  const T *vb[3] = {vb0, vb1, vb2};
  GRID_STRIDE(c, n) {
    const Nbr nb = neighbours(c, pw.nx, pw.ny, pw.nz);
    const cplx<T> s = psi[c];
    const T r = s.x * s.x + s.y * s.y, inv = (T)1 / (r + pw.eps_v);
    T br = (T)2 * rb[c] * s.x, bi = (T)2 * rb[c] * s.y;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const cplx<T> pp = psi[nb.p[d]], pm = psi[nb.m[d]];
      const T dr = (pp.x - pm.x) * pw.inv2h, di = (pp.y - pm.y) * pw.inv2h;
      const T vbd = vb[d][c];
      const T v = (s.x * di - s.y * dr) * pw.invM * inv;
      // local: d v / d psi_r = (di/M - 2 psi_r v) / (rho + eps), d v / d psi_i = (-dr/M - 2 psi_i v) / (rho + eps)
      br += vbd * (di * pw.invM - (T)2 * s.x * v) * inv;
      bi += vbd * (-dr * pw.invM - (T)2 * s.y * v) * inv;
      // through D_d psi: -D_d e_d, e_d = vb_d i psi invM / (rho + eps)
      const T fp = vb[d][nb.p[d]] * pw.invM / (pp.x * pp.x + pp.y * pp.y + pw.eps_v);
      const T fm = vb[d][nb.m[d]] * pw.invM / (pm.x * pm.x + pm.y * pm.y + pw.eps_v);
      br -= (-(fp * pp.y) + fm * pm.y) * pw.inv2h;
      bi -= (fp * pp.x - fm * pm.x) * pw.inv2h;
    }
    cplx<T> b = psib[c];
    b.x += br;
    b.y += bi;
    psib[c] = b;
  }
}

// External potential of a gaussian well, amp exp(-|r - r0|^2 / 2 w^2), with
// r_i = (i - n/2) h (the grid convention of the reference drivers), and its
// derivatives with respect to amp and w.
struct Well {
  double amp, width, x0, y0, z0, h;
  int nx, ny, nz;
};
__device__ inline void well_at(const Well &w, int64_t c, double &v, double &da, double &dw) {
  const int k = (int)(c % w.nz), j = (int)((c / w.nz) % w.ny), i = (int)(c / ((int64_t)w.nz * w.ny));
  const double x = (i - w.nx / 2) * w.h - w.x0, y = (j - w.ny / 2) * w.h - w.y0, z = (k - w.nz / 2) * w.h - w.z0;
  const double r2 = x * x + y * y + z * z;
  const double e = exp(-r2 / (2.0 * w.width * w.width));
  v = w.amp * e;
  da = e;
  dw = w.amp * e * r2 / (w.width * w.width * w.width);
}

// Real field of the well (for the forward steps).
template <typename T> __global__ void k_well(T *out, Well w, int64_t n) {
  GRID_STRIDE(c, n) {
    double v, da, dw;
    well_at(w, c, v, da, dw);
    out[c] = (T)v;
  }
}

// Sum of two FP64 values over the block into pa[block], pb[block].
__device__ inline void block_sum2(double a, double b, double *pa, double *pb) {
  __shared__ double s[2][kBlock];
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pa[blockIdx.x] = s[0][0];
    pb[blockIdx.x] = s[1][0];
  }
}

// veff = Re V_OT + V_ext - mu and veffi = Im V_OT (backflow; may be null).
template <typename T>
__global__ void k_adj_veff(T *veff, T *veffi, const cplx<T> *vot, Well w, double mu, int64_t n) {
  GRID_STRIDE(c, n) {
    double v, da, dw;
    well_at(w, c, v, da, dw);
    veff[c] = (T)((double)vot[c].x + v - mu);
    if (veffi) veffi[c] = vot[c].y;
  }
}

// e^{-i V dt/2} in FP64 for complex V = vr + i vi and complex dt = dtr + i dti.
__device__ inline double2 phase_of(double vr, double vi, double dtr, double dti) {
  double s, c;
  sincos(-0.5 * (vr * dtr - vi * dti), &s, &c);
  const double m = exp(0.5 * (vr * dti + vi * dtr));
  return make_double2(m * c, m * s);
}
template <typename T> __device__ inline double2 phase_at(const T *veff, const T *veffi, int64_t c, double dtr, double dti) {
  return phase_of((double)veff[c], veffi ? (double)veffi[c] : 0.0, dtr, dti);
}

// out = phase(veff) * in
template <typename T>
__global__ void k_adj_phase(cplx<T> *out, const cplx<T> *in, const T *veff, const T *veffi, double dtr, double dti,
                            int64_t n) {
  GRID_STRIDE(c, n) {
    const double2 f = phase_at(veff, veffi, c, dtr, dti);
    const double ar = in[c].x, ai = in[c].y;
    out[c] = cmk<cplx<T>>((T)(f.x * ar - f.y * ai), (T)(f.x * ai + f.y * ar));
  }
}

// Adjoint of psi' = phi u2: u2bar = conj(phi) psibar', phibar = conj(u2) psibar'.
template <typename T>
__global__ void k_adj_out(cplx<T> *u2b, cplx<T> *phib, const cplx<T> *psibn, const cplx<T> *u2, const T *veff,
                          const T *veffi, double dtr, double dti, int64_t n) {
  GRID_STRIDE(c, n) {
    const double2 f = phase_at(veff, veffi, c, dtr, dti);
    const double br = psibn[c].x, bi = psibn[c].y, ur = u2[c].x, ui = u2[c].y;
    u2b[c] = cmk<cplx<T>>((T)(f.x * br + f.y * bi), (T)(f.x * bi - f.y * br));
    phib[c] = cmk<cplx<T>>((T)(ur * br + ui * bi), (T)(ur * bi - ui * br));
  }
}

// Adjoint of u1 = phi psi and of phi = e^{-i V dt/2}: psibar = conj(phi) u1bar,
// phibar += conj(psi) u1bar, Vbar = Re(conj(dphi/dV) phibar) for V's real and
// imaginary parts, dphi/dVr = -i dt/2 phi and dphi/dVi = dt/2 phi; per block:
// sum Vbar_r dV/damp and sum Vbar_r dV/dwidth (V_ext is real).
template <typename T>
__global__ void k_adj_in(cplx<T> *psib, T *vbar, T *vbari, const cplx<T> *u1b, const cplx<T> *psi, const cplx<T> *phib,
                         const T *veff, const T *veffi, double dtr, double dti, Well w, double *pa, double *pw_,
                         int64_t n) {
  double sa = 0.0, sw = 0.0;
  GRID_STRIDE(c, n) {
    const double2 f = phase_at(veff, veffi, c, dtr, dti);
    const double br = u1b[c].x, bi = u1b[c].y, pr = psi[c].x, pi = psi[c].y;
    psib[c] = cmk<cplx<T>>((T)(f.x * br + f.y * bi), (T)(f.x * bi - f.y * br));
    const double qr = phib[c].x + (pr * br + pi * bi), qi = phib[c].y + (pr * bi - pi * br);
    // dphi/dVr = (-i dt / 2) phi = ((dti f.x + dtr f.y) + i (dti f.y - dtr f.x)) / 2
    const double dr = 0.5 * (dti * f.x + dtr * f.y), di = 0.5 * (dti * f.y - dtr * f.x);
    const double vb = dr * qr + di * qi;
    vbar[c] = (T)vb;
    if (vbari) {
      // dphi/dVi = (dt / 2) phi = ((dtr f.x - dti f.y) + i (dti f.x + dtr f.y)) / 2
      const double er = 0.5 * (dtr * f.x - dti * f.y), ei = 0.5 * (dti * f.x + dtr * f.y);
      vbari[c] = (T)(er * qr + ei * qi);
    }
    double v, da, dw;
    well_at(w, c, v, da, dw);
    sa += vb * da;
    sw += vb * dw;
  }
  block_sum2(sa, sw, pa, pw_);
}

// psibar = 2 h^3 [kin + (grad + V_ext) psi], the gradient of the total energy
// (kin = F^-1 K F psi); per block: sum rho dV/damp, sum rho dV/dwidth.
template <typename T>
__global__ void k_adj_seed(cplx<T> *psib, const cplx<T> *kin, const cplx<T> *psi, const T *grad, Well w, double h3,
                           double *pa, double *pw_, int64_t n) {
  double sa = 0.0, sw = 0.0;
  GRID_STRIDE(c, n) {
    double v, da, dw;
    well_at(w, c, v, da, dw);
    const double pr = psi[c].x, pi = psi[c].y, g = (double)grad[c] + v;
    psib[c] = cmk<cplx<T>>((T)(2.0 * h3 * ((double)kin[c].x + g * pr)), (T)(2.0 * h3 * ((double)kin[c].y + g * pi)));
    const double r = pr * pr + pi * pi;
    sa += r * da;
    sw += r * dw;
  }
  block_sum2(sa, sw, pa, pw_);
}

// spec *= K (kinetic symbol cx (mx - ox)^2 + cy (my - oy)^2 + cz (mz - oz)^2;
// o: the momentum frame in units of 2 pi / L)
template <typename T>
__global__ void k_kin_mul(cplx<T> *spec, int nx, int ny, int nz, double cx, double cy, double cz, double ox, double oy,
                          double oz) {
  const int64_t n = (int64_t)nx * ny * nz;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const double mx = (double)(i < nx / 2 ? i : i - nx) - ox, my = (double)(j < ny / 2 ? j : j - ny) - oy,
                 mz = (double)(k < nz / 2 ? k : k - nz) - oz;
    const double K = cx * mx * mx + cy * my * my + cz * mz * mz;
    spec[c] = cmk<cplx<T>>((T)(K * spec[c].x), (T)(K * spec[c].y));
  }
}

}  // namespace dft

// ============================================================ device state ==

enum Kern { K_LJ = 0, K_SA = 1, K_F = 2, K_U = 3 };

// Slot roles (each slot holds a real field or a real-field half spectrum).
enum Slot {
  SP0, SP1, SP2, SP3, SP4, SP5,             // spectra (SP5: rho~ of the HD / HD2 backflow switch)
  VX, VY, VZ, QX, QY, QZ, Q4,               // velocities, products
  RB, RT, PX, PY, PZ, SA_, SBX, SBY, SBZ, SC,  // first inverse batch
  TMP0, TMP1, TMP2, TMP3, TMP4, TMP5, TMP6, TMP7, TMP8, TMP9,
  VRE, VIM, WX, WY, WZ, SH,
  NSLOT
};

struct dft_ot_device {
  int prec;
  int nx, ny, nz, P;
  int PP;                    // row pitch of the half spectra (half_pitch(nz))
  int64_t n, slot_elems;
  double step;
  void *plan;                // Plan3D<float> or Plan3D<double>
  void *slot[NSLOT];
  void *kern[4];             // real half spectra (LJ, SA, F, U)
  void *phase;               // complex: V_eff of the current step, whose e^{-iV dt/2} both z passes apply
  void *vsave;               // complex: predictor potential V_n of a predict-correct step
  double *norm_part;         // per-block sums |psi|^2 and the factor of a renormalising step
  double *igtab;             // pow(k, s) table of the ideal Bose gas (thermal models), or null
  double ig_g1;              // and its g_{3/2}(1)
  // x-middle programs of the fused schedule in device memory, and the host
  // copies they were uploaded from (re-uploaded only when one changes)
  fused::MidProg *mid_dev;
  fused::MidProg mid_host[3];
  bool mid_valid[3];
};

namespace {

// The thermal models above 0 K add the ideal Bose gas term (src/ot.c:402,
// src/ot-energy.c:102: model >= DFT_OT_T400MK, GP / DR / ZERO excluded).
bool has_ig(const dft_ot_functional *otf) { return otf->model >= 32 && otf->model < 524288; }
// ... and its table is on the device (allocated with the functional)
bool ig_on(const dft_ot_functional *otf) { return has_ig(otf) && otf->device->igtab; }
ig::Params ig_params(const dft_ot_functional *otf) {
  ig::Params p;
  p.tab = otf->device->igtab;
  p.ckt = otf->c4 * (otf->temp / GRID_AUTOK);
  p.l3 = ig::lambda3(otf->mass, otf->temp, GRID_AUTOK);
  p.g1 = otf->device->ig_g1;
  return p;
}

template <typename T> Plan3D<T> &plan_of(dft_ot_device *d) { return *(Plan3D<T> *)d->plan; }
template <typename T> T *S(dft_ot_device *d, int i) { return (T *)d->slot[i]; }
template <typename T> cplx<T> *CS(dft_ot_device *d, int i) { return (cplx<T> *)d->slot[i]; }

fft::Combine cmb(int nterms, const fft::Term *t) {
  fft::Combine c{};
  c.nterms = nterms;
  for (int i = 0; i < nterms; ++i) c.t[i] = t[i];
  return c;
}
fft::Combine cmb1(int in, int kern, int deriv, double scale) {
  fft::Term t{in, kern, deriv, scale};
  return cmb(1, &t);
}

template <typename T> fft::KernTab kerntab(dft_ot_device *d) {
  fft::KernTab kt{};
  for (int i = 0; i < 4; ++i) kt.k[i] = d->kern[i];
  Plan3D<T> &p = plan_of<T>(d);
  kt.sinx = p.sinx;
  kt.siny = p.siny;
  kt.sinz = p.sinz;
  return kt;
}

template <typename T> PW<T> make_pw(dft_ot_functional *otf, double wf_mass, double eps_v, bool clamp_v) {
  dft_ot_device *d = otf->device;
  PW<T> pw{};
  pw.nx = d->nx;
  pw.ny = d->ny;
  pw.nz = d->nz;
  pw.inv2h = (T)(0.5 / d->step);
  pw.invM = (T)(1.0 / wf_mass);
  pw.eps_v = (T)eps_v;
  pw.c2 = (T)(otf->c2 * otf->c2_exp / 2.0);
  pw.c3 = (T)(otf->c3 * otf->c3_exp / 3.0);
  pw.inv_rho0s = (T)(1.0 / otf->rho_0s);
  pw.ckc = (T)(otf->alpha_s / (2.0 * otf->mass));
  pw.k23 = (T)otf->kc_coef23;
  pw.mass = (T)otf->mass;
  pw.div_eps = (T)otf->div_epsilon;
  pw.cbf = (T)otf->c_bfpot;
  pw.vmax = (T)otf->max_veloc;
  pw.bfmax = (T)otf->max_bfpot;
  pw.clamp_v = clamp_v && otf->max_veloc < (999.0 / GRID_AUTOMPS);
  pw.clamp_bf = otf->max_bfpot < (10000.0 / GRID_AUTOK);
  pw.hdC = (T)otf->C;
  pw.hdB = (T)otf->beta;
  pw.hdRm = (T)otf->rhom;
  pw.bfXi = (T)otf->xi;
  pw.bfRho = (T)otf->rhobf;
  pw.bfhd = (otf->model & (DFT_OT_HD | DFT_OT_HD2)) && (otf->model & DFT_OT_BACKFLOW);
  return pw;
}

// Pointwise launch; nf = number of N-point real fields read plus written (algorithmic traffic).
#define LAUNCH_PW(name, nf, kern, ...)                                                        \
  do {                                                                                        \
    DFT_PROF_BEGIN(lib_stream());                                                             \
    hipLaunchKernelGGL(kern, dim3(grid_for(d->n)), dim3(kBlock), 0, lib_stream(), __VA_ARGS__); \
    DFT_LAUNCH_CHECK();                                                                       \
    DFT_PROF_END(lib_stream(), name, (double)(nf) * (double)d->n * sizeof(T));               \
  } while (0)

// Which parts of the potential to evaluate.
// loc: local correlation (Pi terms), kc, bf; lj: the Lennard-Jones term
// (-1: with loc, 0 / 1: without / with).
struct PotParts {
  bool loc, kc, bf;
  int lj = -1;
};

// Shared body of dft_ot_potential / potential2 / backflow_potential.
// rho (device) holds the density on entry if psi == nullptr.
template <typename T>
void potential_impl(dft_ot_functional *otf, cgrid *V, const cplx<T> *psi, double wf_mass, T *rho, T *vx, T *vy,
                    T *vz, PotParts parts, bool write_back_w) {
  dft_ot_device *d = otf->device;
  Plan3D<T> &p = plan_of<T>(d);
  hipStream_t s = lib_stream();
  const double h3N = d->step * d->step * d->step / (double)d->n;
  const bool KC = parts.kc, BF = parts.bf, LOC = parts.loc, LJ = parts.lj < 0 ? LOC : parts.lj != 0;
  fft::KernTab kt = kerntab<T>(d);
  PW<T> pw = make_pw<T>(otf, wf_mass, otf->div_epsilon, true);

  // P1
  if (psi) {
    if (BF) {
      vx = S<T>(d, VX);
      vy = S<T>(d, VY);
      vz = S<T>(d, VZ);
      LAUNCH_PW("ot_prep_density_velocity", 10, (k_prep<T, true>), psi, rho, vx, vy, vz, S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, Q4), S<T>(d, SA_), pw);
    } else {
      LAUNCH_PW("ot_prep_density", 3, (k_prep<T, false>), psi, rho, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, pw);
    }
  } else if (BF) {
    LAUNCH_PW("ot_prep_products", 8, (k_prep_v<T>), rho, vx, vy, vz, S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, Q4), S<T>(d, SA_), pw);
  }
  // F1 (rho~ of the HD / HD2 backflow switch in SA_ -> SP5)
  const bool HB = BF && pw.bfhd;
  {
    std::vector<const T *> in{rho};
    std::vector<cplx<T> *> out{CS<T>(d, SP0)};
    if (BF) {
      for (int i : {QX, QY, QZ, Q4}) in.push_back(S<T>(d, i));
      for (int i : {SP1, SP2, SP3, SP4}) out.push_back(CS<T>(d, i));
    }
    if (HB) {
      in.push_back(S<T>(d, SA_));
      out.push_back(CS<T>(d, SP5));
    }
    r2c3d<T>(p, (int)in.size(), in.data(), out.data(), s);
  }
  // I1
  {
    const cplx<T> *in[6] = {CS<T>(d, SP0), CS<T>(d, SP1), CS<T>(d, SP2), CS<T>(d, SP3), CS<T>(d, SP4), CS<T>(d, SP5)};
    std::vector<fft::Combine> c;
    std::vector<T *> out;
    if (LOC) {
      c.push_back(cmb1(0, K_SA, -1, h3N));
      out.push_back(S<T>(d, RB));
    }
    if (KC) {
      c.push_back(cmb1(0, K_F, -1, h3N));
      out.push_back(S<T>(d, RT));
      for (int dd = 0; dd < 3; ++dd) {
        c.push_back(cmb1(0, K_F, dd, h3N));
        out.push_back(S<T>(d, PX + dd));
      }
    }
    if (BF) {
      c.push_back(cmb1(HB ? 5 : 0, K_U, -1, h3N));
      out.push_back(S<T>(d, SA_));
      for (int dd = 0; dd < 3; ++dd) {
        c.push_back(cmb1(1 + dd, K_U, -1, h3N));
        out.push_back(S<T>(d, SBX + dd));
      }
      c.push_back(cmb1(4, K_U, -1, h3N));
      out.push_back(S<T>(d, SC));
    }
    std::vector<cplx<T> *> tmp;
    for (size_t i = 0; i < out.size(); ++i) tmp.push_back(CS<T>(d, TMP0 + (int)i));
    if (!out.empty()) c2r3d<T>(p, HB ? 6 : 5, in, (int)out.size(), c.data(), kt, tmp.data(), out.data(), s);
  }
  // P2
  {
    StageBIn bi{rho, S<T>(d, RB), S<T>(d, RT), vx, vy, vz, S<T>(d, SA_), S<T>(d, SBX), S<T>(d, SBY), S<T>(d, SBZ),
                S<T>(d, SC)};
    StageBOut bo{S<T>(d, Q4), S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, VRE), S<T>(d, VIM),
                 S<T>(d, WX), S<T>(d, WY), S<T>(d, WZ)};
#define SB(L, K, B) LAUNCH_PW("ot_stage_local_kc_bf", 2 + 2 * L + 4 * K + 12 * B, (k_stageB<T, L, K, B>), bi, bo, pw)
    if (LOC && KC && BF) SB(true, true, true);
    else if (LOC && KC) SB(true, true, false);
    else if (LOC && BF) SB(true, false, true);
    else if (LOC) SB(true, false, false);
    else if (KC && BF) SB(false, true, true);
    else if (KC) SB(false, true, false);
    else if (BF) SB(false, false, true);
#undef SB
  }
  // F2
  {
    std::vector<const T *> in;
    std::vector<cplx<T> *> out;
    if (LOC) {
      in.push_back(S<T>(d, Q4));
      out.push_back(CS<T>(d, SP4));
    }
    if (KC)
      for (int dd = 0; dd < 3; ++dd) {
        in.push_back(S<T>(d, QX + dd));
        out.push_back(CS<T>(d, SP1 + dd));
      }
    if (!in.empty()) r2c3d<T>(p, (int)in.size(), in.data(), out.data(), s);
  }
  // I2 (J_d and T1 into the slots F2 has consumed; A = U*rho in SA_ stays for the adjoint)
  if (KC) {
    const cplx<T> *in[3] = {CS<T>(d, SP1), CS<T>(d, SP2), CS<T>(d, SP3)};
    fft::Combine c[4];
    T *out[4] = {S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, Q4)};
    for (int dd = 0; dd < 3; ++dd) c[dd] = cmb1(dd, K_F, -1, h3N);
    fft::Term t[3] = {{0, K_F, 0, h3N}, {1, K_F, 1, h3N}, {2, K_F, 2, h3N}};
    c[3] = cmb(3, t);
    cplx<T> *tmp[4] = {CS<T>(d, TMP0), CS<T>(d, TMP1), CS<T>(d, TMP2), CS<T>(d, TMP3)};
    c2r3d<T>(p, 3, in, 4, c, kt, tmp, out, s);
  }
  // P3
  if (KC || BF) {
    StageCIn ci{rho, S<T>(d, RT), S<T>(d, Q4), S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, PX),
                S<T>(d, PY), S<T>(d, PZ), S<T>(d, WX), S<T>(d, WY), S<T>(d, WZ)};
    StageCOut co{S<T>(d, TMP5), S<T>(d, VRE), S<T>(d, VIM)};
    if (KC && BF) LAUNCH_PW("ot_stage_kc_terms_bf_div", 17, (k_stageC<T, true, true>), ci, co, pw);
    else if (KC) LAUNCH_PW("ot_stage_kc_terms_bf_div", 12, (k_stageC<T, true, false>), ci, co, pw);
    else LAUNCH_PW("ot_stage_kc_terms_bf_div", 5, (k_stageC<T, false, true>), ci, co, pw);
  }
  // F3 + I3
  if (LOC || KC || LJ) {
    if (KC) {
      const T *in[1] = {S<T>(d, TMP5)};
      cplx<T> *out[1] = {CS<T>(d, SH)};
      r2c3d<T>(p, 1, in, out, s);
    }
    const cplx<T> *in[3] = {CS<T>(d, SP0), CS<T>(d, SP4), CS<T>(d, SH)};
    fft::Term t[3];
    int nt = 0;
    if (LJ) t[nt++] = {0, K_LJ, -1, h3N};
    if (LOC) t[nt++] = {1, K_SA, -1, h3N};
    if (KC) t[nt++] = {2, K_F, -1, h3N * otf->kc_coef23 * otf->alpha_s / (2.0 * otf->mass)};
    fft::Combine c = cmb(nt, t);
    // (TMP1 as a real field: RB keeps rhobar for the adjoint pass)
    T *out[1] = {S<T>(d, TMP1)};
    cplx<T> *tmp[1] = {CS<T>(d, TMP0)};
    c2r3d<T>(p, 3, in, 1, &c, kt, tmp, out, s);
  }
  // ideal Bose gas of the thermal models, with the whole potential (dft_ot_potential,
  // dft_ot_potential2: src/ot.c:402-407, 505-510)
  if (LOC && parts.lj < 0 && ig_on(otf)) {
    const ig::Params igp = ig_params(otf);
    LAUNCH_PW("ot_ideal_gas", 3, (k_ideal_gas<T, 0>), S<T>(d, VRE), (const T *)rho, (const T *)nullptr, igp, d->n);
  }
  // P4
  const bool PW_ = LOC || KC || BF;  // pointwise stages ran (V_re in VRE)
  LAUNCH_PW("ot_accumulate_potential", 4 + (PW_ ? 1 : 0) + ((LOC || KC || LJ) ? 1 : 0) + (BF ? 1 : 0), (k_final<T>),
            (cplx<T> *)V->dev, PW_ ? S<T>(d, VRE) : (T *)nullptr, (LOC || KC || LJ) ? S<T>(d, TMP1) : (T *)nullptr,
            BF ? S<T>(d, VIM) : (T *)nullptr, d->n);
  if (write_back_w && BF) {
    // the reference leaves v_d A - B_d in the velocity inputs (src/ot.c:923-931)
    DFT_HIP_CHECK(hipMemcpyAsync(vx, S<T>(d, WX), sizeof(T) * d->n, hipMemcpyDeviceToDevice, s));
    DFT_HIP_CHECK(hipMemcpyAsync(vy, S<T>(d, WY), sizeof(T) * d->n, hipMemcpyDeviceToDevice, s));
    DFT_HIP_CHECK(hipMemcpyAsync(vz, S<T>(d, WZ), sizeof(T) * d->n, hipMemcpyDeviceToDevice, s));
  }
}

// ------------------------------------------------------------ fused path --
// dft_ot_potential on a wave function and the split-operator steps built on
// it.  Schedule (stage kernels in ot_fused.hpp):
//   zA, y, xmid{Pi, F, U products}, y^-1, zAB, y, [xmid{J_d, T1}, y^-1, zBC, y],
//   xmid{V_LJ + Pi Q + k23 c F H}, y^-1, zC
// 7 + 5 forward and 7 + 4 + 1 inverse real 3-D transforms for full OT, with
// all pointwise stages inside the z passes.
struct StepIn {
  const void *vext;   // real field or null
  double mu;
  const void *vavg;   // predictor potential (corrector step) or null
  void *vsave;        // where to keep V (predictor step) or null
  void *phase;
  const void *psi_src;
  void *psi_dst;
  double dtr, dti;
};

template <typename T, int N> void launch_zA(bool BF, const cplx<T> *psi, T *rho, const fused::Ptrs8 &o,
                                            int64_t npairs, int P, const PW<T> &pw, const cplx<T> *tw,
                                            hipStream_t s, double bytes) {
  constexpr int LN = rf::LineCfg<T, N, fused::pair_em<N>()>::LINES;
  const dim3 grid((unsigned)((npairs + LN - 1) / LN));
  DFT_PROF_BEGIN(s);
  if (BF && pw.bfhd) hipLaunchKernelGGL((fused::k_zA_bfhd<T, N>), grid, dim3(fused::PairBS<T, N>), 0, s, psi, rho, o, npairs, P, pw, tw);
  else if (BF) hipLaunchKernelGGL((fused::k_zA<T, N, true>), grid, dim3(fused::PairBS<T, N>), 0, s, psi, rho, o, npairs, P, pw, tw);
  else hipLaunchKernelGGL((fused::k_zA<T, N, false>), grid, dim3(fused::PairBS<T, N>), 0, s, psi, rho, o, npairs, P, pw, tw);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "ot_zA_density_r2c", bytes);
}

// x-middle implementation: 0 = all outputs of a column tile in one block
// (k_xmid), 1 = one output per block (k_xconv) in FP32 for the programs
// without derivative terms, 2 = k_xconv for every FP32 program, 3 = as 1 in
// FP64 too (N <= 256)
#ifndef DFT_XMID_KIND
#define DFT_XMID_KIND 3
#endif

template <typename T, int N, int NK, bool DV>
void launch_xmid1(const fused::MidProg *pg, int ny, int P, int PP, const cplx<T> *tw, hipStream_t s) {
  using Cfg = fused::MidCfg<T, N>;
  hipLaunchKernelGGL((fused::k_xmid<T, N, NK, DV>), dim3((unsigned)fused::mid_blocks<Cfg::COLS>(ny, P)),
                     dim3(Cfg::BS), 0, s, pg, ny, P, PP, tw);
}

template <typename T, int N>
const char *launch_xconv(const fused::MidProg *pg, const fused::MidProg &hp, bool deriv, int ny, int P, int PP,
                         const cplx<T> *tw, hipStream_t s) {
  // leading outputs with a single kernel term take the lean form, the rest
  // the accumulating one
  int n1 = 0;
  while (n1 < hp.nout && hp.o[n1].nterms == 1 && hp.o[n1].t[0].kern >= 0 && hp.o[n1].t[0].deriv < 0) ++n1;
  using Cfg1 = fused::XConvCfg<T, N, true>;
  using Cfg = fused::XConvCfg<T, N, false>;
  if (n1 > 0)
    hipLaunchKernelGGL((fused::k_xconv<T, N, false, true>), dim3((unsigned)(fused::mid_blocks<Cfg1::COLS>(ny, P) * n1)),
                       dim3(Cfg1::BS), 0, s, pg, 0, n1, ny, P, PP, tw);
  if (n1 < hp.nout) {
    const dim3 grid((unsigned)(fused::mid_blocks<Cfg::COLS>(ny, P) * (hp.nout - n1)));
    if (deriv)
      hipLaunchKernelGGL((fused::k_xconv<T, N, true, false>), grid, dim3(Cfg::BS), 0, s, pg, n1, hp.nout - n1, ny, P,
                         PP, tw);
    else
      hipLaunchKernelGGL((fused::k_xconv<T, N, false, false>), grid, dim3(Cfg::BS), 0, s, pg, n1, hp.nout - n1, ny, P,
                         PP, tw);
  }
  return n1 == hp.nout ? "ot_xconv_single" : (n1 == 0 ? "ot_xconv_multi" : "ot_xconv_mixed");
}

// x-middle pass of program pg (device copy; host copy hp).  Programs without
// derivative terms run one output per block (k_xconv; FP64 only up to
// N = 256); the others, and FP64 at N = 512, all outputs of a column tile per
// block (k_xmid, instantiated per number of distinct kernels): measured at
// 256^3 FP32, k_xconv takes 337 and 139 us for the first and last programs
// against 427 and 158, k_xmid 266 us for the derivative program against 380.
template <typename T, int N>
void launch_xmid(const fused::MidProg *pg, const fused::MidProg &hp, int ny, int P, int PP, const cplx<T> *tw,
                 hipStream_t s, double bytes) {
  bool deriv = false;
  for (int m = 0; m < hp.nout; ++m)
    for (int q = 0; q < hp.o[m].nterms; ++q) deriv = deriv || hp.o[m].t[q].deriv >= 0;
  DFT_PROF_BEGIN(s);
  bool one = true;
  for (int m = 0; m < hp.nout; ++m) one = one && hp.o[m].nterms == 1;
  if constexpr (((DFT_XMID_KIND == 1 || DFT_XMID_KIND == 2) && sizeof(T) == 4) ||
                (DFT_XMID_KIND == 3 && (sizeof(T) == 4 || N <= 256))) {
    // per-output blocks unless the program shares inputs between
    // multi-term outputs with derivative symbols (J_d and T1 = F sum_d D_d G_d)
    if (one || !deriv || DFT_XMID_KIND == 2) {
      const char *label = launch_xconv<T, N>(pg, hp, deriv, ny, P, PP, tw, s);
      DFT_LAUNCH_CHECK();
      DFT_PROF_END(s, label, bytes);
      return;
    }
  }
  {
    if (deriv && hp.nkern == 1) launch_xmid1<T, N, 1, true>(pg, ny, P, PP, tw, s);
    else if (!deriv && hp.nkern == 1) launch_xmid1<T, N, 1, false>(pg, ny, P, PP, tw, s);
    else if (!deriv && hp.nkern == 2) launch_xmid1<T, N, 2, false>(pg, ny, P, PP, tw, s);
    else if (!deriv && hp.nkern == 3) launch_xmid1<T, N, 3, false>(pg, ny, P, PP, tw, s);
    else {
      fprintf(stderr, "libdft: x-middle program with %d kernels%s is not instantiated\n", hp.nkern,
              deriv ? " and derivative terms" : "");
      abort();
    }
  }
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "ot_xmid_fused", bytes);
}

// HB: the backflow density switch of HD / HD2 (k_zAB_bfhd)
template <typename T, int N, bool KC, bool BF, bool HB = false>
void launch_zAB1(const fused::ZABArgs &a, int64_t npairs, int P, const PW<T> &pw, const cplx<T> *tw, hipStream_t s) {
  constexpr int LN = rf::LineCfg<T, N, fused::pair_em<N>()>::LINES;
  const dim3 grid((unsigned)((npairs + LN - 1) / LN));
  if constexpr (BF && (sizeof(T) == 8 ? DFT_ZAB_SPLIT64 : (HB ? DFT_ZAB_SPLIT32_HB : DFT_ZAB_SPLIT32))) {
    // the backflow terms, then the local and KC terms (V_re through memory)
    if constexpr (HB) {
      hipLaunchKernelGGL((fused::k_zAB_bfhd<T, N, KC, 1>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
      DFT_LAUNCH_CHECK();
      hipLaunchKernelGGL((fused::k_zAB_bfhd<T, N, KC, 2>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
    } else {
      hipLaunchKernelGGL((fused::k_zAB<T, N, KC, BF, 1>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
      DFT_LAUNCH_CHECK();
      hipLaunchKernelGGL((fused::k_zAB<T, N, KC, BF, 2>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
    }
  } else if constexpr (HB) {
    hipLaunchKernelGGL((fused::k_zAB_bfhd<T, N, KC>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
  } else {
    hipLaunchKernelGGL((fused::k_zAB<T, N, KC, BF>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
  }
}

template <typename T, int N>
void launch_zAB(bool KC, bool BF, const fused::ZABArgs &a, int64_t npairs, int P, const PW<T> &pw,
                const cplx<T> *tw, hipStream_t s, double bytes) {
  DFT_PROF_BEGIN(s);
  if (BF && pw.bfhd) {
    if (KC) launch_zAB1<T, N, true, true, true>(a, npairs, P, pw, tw, s);
    else launch_zAB1<T, N, false, true, true>(a, npairs, P, pw, tw, s);
  } else if (KC && BF) launch_zAB1<T, N, true, true>(a, npairs, P, pw, tw, s);
  else if (KC) launch_zAB1<T, N, true, false>(a, npairs, P, pw, tw, s);
  else if (BF) launch_zAB1<T, N, false, true>(a, npairs, P, pw, tw, s);
  else launch_zAB1<T, N, false, false>(a, npairs, P, pw, tw, s);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "ot_zAB_c2r_local_kc_bf_r2c", bytes);
}

template <typename T, int N>
void launch_zBC(bool BF, const fused::ZBCArgs &a, int64_t npairs, int P, const PW<T> &pw, const cplx<T> *tw,
                hipStream_t s, double bytes) {
  constexpr int LN = rf::LineCfg<T, N, fused::pair_em<N>()>::LINES;
  const dim3 grid((unsigned)((npairs + LN - 1) / LN));
  DFT_PROF_BEGIN(s);
  constexpr bool SPLIT = sizeof(T) == 8 && DFT_ZBC_SPLIT64;
  if constexpr (SPLIT) {
    // FP64: the J_d stages, then T1, the tail and the r2c of H (H and the term-3 sum through memory)
    if (BF && pw.bfhd) {
      hipLaunchKernelGGL((fused::k_zBC_bfhd<T, N, 1>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
      hipLaunchKernelGGL((fused::k_zBC_bfhd<T, N, 2>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
    } else if (BF) {
      hipLaunchKernelGGL((fused::k_zBC<T, N, true, 1>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
      hipLaunchKernelGGL((fused::k_zBC<T, N, true, 2>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
    } else {
      hipLaunchKernelGGL((fused::k_zBC<T, N, false, 1>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
      hipLaunchKernelGGL((fused::k_zBC<T, N, false, 2>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
    }
  } else if (BF && pw.bfhd) {
    hipLaunchKernelGGL((fused::k_zBC_bfhd<T, N>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
  } else if (BF) {
    hipLaunchKernelGGL((fused::k_zBC<T, N, true>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
  } else {
    hipLaunchKernelGGL((fused::k_zBC<T, N, false>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, pw, tw);
  }
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, "ot_zBC_c2r_kc_terms_r2c", bytes);
}

template <typename T, int N>
void launch_zC(bool step, const fused::ZCArgs &a, int64_t npairs, int P, const cplx<T> *tw, hipStream_t s,
               double bytes) {
  constexpr int LN = rf::LineCfg<T, N, fused::pair_em<N>()>::LINES;
  const dim3 grid((unsigned)((npairs + LN - 1) / LN));
  DFT_PROF_BEGIN(s);
  if (step) hipLaunchKernelGGL((fused::k_zC<T, N, true>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, tw);
  else hipLaunchKernelGGL((fused::k_zC<T, N, false>), grid, dim3(fused::PairBS<T, N>), 0, s, a, npairs, P, tw);
  DFT_LAUNCH_CHECK();
  DFT_PROF_END(s, step ? "ot_zC_c2r_phase_z_fwd" : "ot_zC_c2r_accumulate", bytes);
}

// Device copy of x-middle program k (uploaded when it differs from the last one).
const fused::MidProg *mid_prog(dft_ot_device *d, int k, const fused::MidProg &pg, hipStream_t s) {
  if (!d->mid_valid[k] || memcmp(&d->mid_host[k], &pg, sizeof pg) != 0) {
    DFT_HIP_CHECK(hipStreamSynchronize(s));
    memcpy(&d->mid_host[k], &pg, sizeof pg);
    DFT_HIP_CHECK(hipMemcpy(d->mid_dev + k, &d->mid_host[k], sizeof pg, hipMemcpyHostToDevice));
    d->mid_valid[k] = true;
  }
  return d->mid_dev + k;
}

// A zero-filled program (padding included, so that programs compare bytewise).
fused::MidProg new_prog() {
  fused::MidProg pg;
  memset(&pg, 0, sizeof pg);
  return pg;
}

fused::MidTerm mterm(int in, int kern, int deriv, double scale) {
  fused::MidTerm t{};
  t.in = in;
  t.kern = kern;
  t.deriv = deriv;
  t.scale = scale;
  return t;
}

// Algorithmic traffic of an x-middle program: inputs and outputs, and each
// distinct kernel spectrum once.
template <typename T> double mid_bytes(const fused::MidProg &pg, int64_t nh) {
  return (pg.nin + pg.nout) * (double)nh * sizeof(cplx<T>) + pg.nkern * (double)nh * sizeof(T);
}

template <typename T>
void fused_impl(dft_ot_functional *otf, const cplx<T> *psi, double wf_mass, T *rho, cplx<T> *V, const StepIn *st) {
  dft_ot_device *d = otf->device;
  Plan3D<T> &p = plan_of<T>(d);
  hipStream_t s = lib_stream();
  const double h3N = d->step * d->step * d->step / (double)d->n;
  const bool KC = (otf->model & DFT_OT_KC) != 0, BF = (otf->model & DFT_OT_BACKFLOW) != 0;
  const PW<T> pw = make_pw<T>(otf, wf_mass, otf->div_epsilon, true);
  const int64_t npairs = (int64_t)d->nx * d->ny / 2, nh = (int64_t)d->nx * d->ny * d->P, n = d->n;
  const int P = d->P, PP = d->PP;
  const double R = (double)n * sizeof(T), H = (double)nh * sizeof(cplx<T>);
  fft::KernTab kt = kerntab<T>(d);
  // terms name kernels by K_*; the program lists the distinct ones and the terms their slots
  auto setk = [&](fused::MidProg &pg) {
    int slot[4] = {-1, -1, -1, -1};
    pg.nkern = 0;
    for (int m = 0; m < pg.nout; ++m)
      for (int q = 0; q < pg.o[m].nterms; ++q) {
        int &k = pg.o[m].t[q].kern;
        if (k < 0) continue;
        if (slot[k] < 0) {
          if (pg.nkern == fused::kMidMaxKern) {
            fprintf(stderr, "libdft: x-middle program uses more than %d kernels\n", fused::kMidMaxKern);
            abort();
          }
          slot[k] = pg.nkern;
          pg.kern[pg.nkern++] = kt.k[k];
        }
        k = slot[k];
      }
    pg.sinx = kt.sinx;
    pg.siny = kt.siny;
    pg.sinz = kt.sinz;
  };

  // ---- stage A: psi -> {rho, q_d, q4, [rho~]}, forward z and y
  const bool HB = BF && pw.bfhd;
  const int nA = BF ? (HB ? 6 : 5) : 1;
  fused::Ptrs8 oA{};
  for (int i = 0; i < nA; ++i) oA.p[i] = d->slot[SP0 + i];
  DFT_DISPATCH_N(d->nz, (launch_zA<T, N_>(BF, psi, rho, oA, npairs, PP, pw, p.twz, s, 3 * R + nA * H)));
  {
    cplx<T> *b[6];
    for (int i = 0; i < nA; ++i) b[i] = CS<T>(d, SP0 + i);
    ypass<T>(p, nA, b, -1, P, PP, s);
  }
  // x middle: Pi*rho, [F*rho], [U*rho (U*rho~ with the switch), U*q_x, U*q_y, U*q_z, U*q4]
  int nB = 0;
  {
    fused::MidProg pg = new_prog();
    pg.nin = nA;
    for (int i = 0; i < nA; ++i) pg.in[i] = d->slot[SP0 + i];
    pg.multi = -1;
    auto add = [&](int in, int kern) {
      pg.o[pg.nout].nterms = 1;
      pg.o[pg.nout].t[0] = mterm(in, kern, -1, h3N);
      pg.o[pg.nout].out = d->slot[TMP0 + pg.nout];
      pg.nout++;
    };
    add(0, K_SA);
    if (KC) add(0, K_F);
    if (BF) {
      add(HB ? 5 : 0, K_U);
      for (int i = 1; i <= 4; ++i) add(i, K_U);
    }
    setk(pg);
    nB = pg.nout;
    DFT_DISPATCH_N(d->nx, (launch_xmid<T, N_>(mid_prog(d, 0, pg, s), pg, d->ny, P, PP, p.twx, s, mid_bytes<T>(pg, nh))));
    cplx<T> *b[8];
    for (int i = 0; i < nB; ++i) b[i] = CS<T>(d, TMP0 + i);
    ypass<T>(p, nB, b, +1, P, PP, s);
  }
  // ---- stage B: z^-1, local / KC / BF pointwise, z of {Q, G_d}
  const int nQ = KC ? 4 : 1;
  {
    fused::ZABArgs a{};
    for (int i = 0; i < nB; ++i) a.in.p[i] = d->slot[TMP0 + i];
    a.out.p[0] = d->slot[Q4];
    a.out.p[1] = d->slot[QX];
    a.out.p[2] = d->slot[QY];
    a.out.p[3] = d->slot[QZ];
    a.rho = rho;
    a.psi = psi;
    a.rhot = d->slot[RT];
    a.vre = d->slot[VRE];
    a.vim = d->slot[VIM];
    a.wx = d->slot[WX];
    a.wy = d->slot[WY];
    a.wz = d->slot[WZ];
    const double bytes = nB * H + R + (BF ? 2 * R : 0) + nQ * H + R + (KC ? R : 0) + (BF ? 4 * R : 0);
    DFT_DISPATCH_N(d->nz, (launch_zAB<T, N_>(KC, BF, a, npairs, PP, pw, p.twz, s, bytes)));
    cplx<T> *b[4] = {CS<T>(d, Q4), CS<T>(d, QX), CS<T>(d, QY), CS<T>(d, QZ)};
    ypass<T>(p, nQ, b, -1, P, PP, s);
  }
  // ---- KC: J_d = F*G_d, T1 = F*sum_d D_d G_d; then H, terms 1 and 3, BF divergence
  if (KC) {
    fused::MidProg pg = new_prog();
    pg.nin = 3;
    pg.in[0] = d->slot[QX];
    pg.in[1] = d->slot[QY];
    pg.in[2] = d->slot[QZ];
    pg.nout = 4;
    for (int dd = 0; dd < 3; ++dd) {
      pg.o[dd].nterms = 1;
      pg.o[dd].t[0] = mterm(dd, K_F, -1, h3N);
      pg.o[dd].out = d->slot[TMP0 + dd];
    }
    pg.o[3].nterms = 3;
    for (int dd = 0; dd < 3; ++dd) pg.o[3].t[dd] = mterm(dd, K_F, dd, h3N);
    pg.o[3].out = d->slot[TMP3];
    pg.multi = 3;
    setk(pg);
    DFT_DISPATCH_N(d->nx, (launch_xmid<T, N_>(mid_prog(d, 1, pg, s), pg, d->ny, P, PP, p.twx, s, mid_bytes<T>(pg, nh))));
    cplx<T> *b[4] = {CS<T>(d, TMP0), CS<T>(d, TMP1), CS<T>(d, TMP2), CS<T>(d, TMP3)};
    ypass<T>(p, 4, b, +1, P, PP, s);
    fused::ZBCArgs a{};
    for (int i = 0; i < 4; ++i) a.in.p[i] = d->slot[TMP0 + i];
    a.out.p[0] = d->slot[SH];
    a.rho = rho;
    a.rhot = d->slot[RT];
    a.wx = d->slot[WX];
    a.wy = d->slot[WY];
    a.wz = d->slot[WZ];
    a.vre = d->slot[VRE];
    a.vim = d->slot[VIM];
    // scratch of the split form: the G_d spectra in QX, QY are consumed by the x-middle above
    a.hbuf = d->slot[QX];
    a.tbuf = d->slot[QY];
    const double bytes = 4 * H + 3 * R + (BF ? 4 * R : 0) + H + R + (BF ? R : 0);
    DFT_DISPATCH_N(d->nz, (launch_zBC<T, N_>(BF, a, npairs, PP, pw, p.twz, s, bytes)));
    cplx<T> *hb[1] = {CS<T>(d, SH)};
    ypass<T>(p, 1, hb, -1, P, PP, s);
  } else if (BF) {
    StageCIn ci{rho, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                S<T>(d, WX), S<T>(d, WY), S<T>(d, WZ)};
    StageCOut co{nullptr, S<T>(d, VRE), S<T>(d, VIM)};
    LAUNCH_PW("ot_stage_kc_terms_bf_div", 5, (k_stageC<T, false, true>), ci, co, pw);
  }
  // ---- ideal Bose gas of the thermal models: pointwise, into V_re
  if (ig_on(otf)) {
    const ig::Params igp = ig_params(otf);
    LAUNCH_PW("ot_ideal_gas", 3, (k_ideal_gas<T, 0>), (T *)d->slot[VRE], (const T *)rho, (const T *)nullptr, igp, n);
  }
  // ---- stage C: linear terms summed in k space, z^-1, accumulate or start the step
  {
    fused::MidProg pg = new_prog();
    pg.nin = KC ? 3 : 2;
    pg.in[0] = d->slot[SP0];
    pg.in[1] = d->slot[Q4];
    pg.in[2] = d->slot[SH];
    pg.nout = 1;
    pg.multi = 0;
    pg.o[0].nterms = KC ? 3 : 2;
    pg.o[0].t[0] = mterm(0, K_LJ, -1, h3N);
    pg.o[0].t[1] = mterm(1, K_SA, -1, h3N);
    pg.o[0].t[2] = mterm(2, K_F, -1, h3N * otf->kc_coef23 * otf->alpha_s / (2.0 * otf->mass));
    pg.o[0].out = d->slot[TMP4];
    setk(pg);
    DFT_DISPATCH_N(d->nx, (launch_xmid<T, N_>(mid_prog(d, 2, pg, s), pg, d->ny, P, PP, p.twx, s, mid_bytes<T>(pg, nh))));
    cplx<T> *b[1] = {CS<T>(d, TMP4)};
    ypass<T>(p, 1, b, +1, P, PP, s);
    fused::ZCArgs a{};
    a.in.p[0] = d->slot[TMP4];
    a.vre = d->slot[VRE];
    a.vim = BF ? d->slot[VIM] : nullptr;
    double bytes = H + R + (BF ? R : 0);
    if (!st) {
      a.V = V;
      bytes += 4 * R;
    } else {
      a.vext = st->vext;
      a.mu = st->mu;
      a.vavg = st->vavg;
      a.vsave = st->vsave;
      a.phase = st->phase;
      a.psi_src = st->psi_src;
      a.psi_dst = st->psi_dst;
      a.dtr = st->dtr;
      a.dti = st->dti;
      bytes += (st->vext ? R : 0) + (st->vavg ? 2 * R : 0) + (st->vsave ? 2 * R : 0) + 6 * R;
    }
    DFT_DISPATCH_N(d->nz, (launch_zC<T, N_>(st != nullptr, a, npairs, PP, p.twz, s, bytes)));
  }
}

// Energy density: eps = (acc ? eps : 0) + selected parts.  rho holds the
// density on entry when psi == nullptr (the _kc / _bf entry points).
template <typename T>
void energy_impl(dft_ot_functional *otf, T *eps, const cplx<T> *psi, double wf_mass, T *rho, bool rho_given,
                 PotParts parts, bool acc) {
  dft_ot_device *d = otf->device;
  Plan3D<T> &p = plan_of<T>(d);
  hipStream_t s = lib_stream();
  const double h3N = d->step * d->step * d->step / (double)d->n;
  const bool KC = parts.kc, BF = parts.bf, LOC = parts.loc;
  fft::KernTab kt = kerntab<T>(d);
  PW<T> pw = make_pw<T>(otf, wf_mass, DFT_EPS, false);
  T *vx = S<T>(d, VX), *vy = S<T>(d, VY), *vz = S<T>(d, VZ);
  if (BF && !psi) {
    fprintf(stderr, "libdft: backflow energy needs the wave function\n");
    return;
  }
  if (!rho_given) {
    // density (and velocities) from psi
    if (BF)
      LAUNCH_PW("ot_prep_density_velocity", 10, (k_prep<T, true>), psi, rho, vx, vy, vz, S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, Q4), S<T>(d, SA_), pw);
    else
      LAUNCH_PW("ot_prep_density", 3, (k_prep<T, false>), psi, rho, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, pw);
  } else if (BF) {
    // velocities from psi (grid_wf_velocity, src/ot-energy.c:231), products with the given density
    LAUNCH_PW("ot_prep_density_velocity", 10, (k_prep<T, true>), psi, S<T>(d, VRE), vx, vy, vz, S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, Q4),
              S<T>(d, SA_), pw);
    LAUNCH_PW("ot_prep_products", 8, (k_prep_v<T>), rho, vx, vy, vz, S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, Q4), S<T>(d, SA_), pw);
  }
  const bool HB = BF && pw.bfhd;  // rho~ of the HD / HD2 backflow switch in SA_ -> SP5
  {
    std::vector<const T *> in{rho};
    std::vector<cplx<T> *> out{CS<T>(d, SP0)};
    if (BF) {
      for (int i : {QX, QY, QZ, Q4}) in.push_back(S<T>(d, i));
      for (int i : {SP1, SP2, SP3, SP4}) out.push_back(CS<T>(d, i));
    }
    if (HB) {
      in.push_back(S<T>(d, SA_));
      out.push_back(CS<T>(d, SP5));
    }
    r2c3d<T>(p, (int)in.size(), in.data(), out.data(), s);
  }
  {
    const cplx<T> *in[6] = {CS<T>(d, SP0), CS<T>(d, SP1), CS<T>(d, SP2), CS<T>(d, SP3), CS<T>(d, SP4), CS<T>(d, SP5)};
    std::vector<fft::Combine> c;
    std::vector<T *> out;
    if (LOC) {
      c.push_back(cmb1(0, K_LJ, -1, h3N));
      out.push_back(S<T>(d, PX));
      c.push_back(cmb1(0, K_SA, -1, h3N));
      out.push_back(S<T>(d, RB));
    }
    if (KC) {
      c.push_back(cmb1(0, K_F, -1, h3N));
      out.push_back(S<T>(d, RT));
    }
    if (BF) {
      c.push_back(cmb1(HB ? 5 : 0, K_U, -1, h3N));
      out.push_back(S<T>(d, SA_));
      for (int dd = 0; dd < 3; ++dd) {
        c.push_back(cmb1(1 + dd, K_U, -1, h3N));
        out.push_back(S<T>(d, SBX + dd));
      }
      c.push_back(cmb1(4, K_U, -1, h3N));
      out.push_back(S<T>(d, SC));
    }
    std::vector<cplx<T> *> tmp;
    for (size_t i = 0; i < out.size(); ++i) tmp.push_back(CS<T>(d, TMP0 + (int)i));
    c2r3d<T>(p, HB ? 6 : 5, in, (int)out.size(), c.data(), kt, tmp.data(), out.data(), s);
  }
  {
    EnergyIn ei{rho, S<T>(d, PX), S<T>(d, RB), S<T>(d, RT), vx, vy, vz, S<T>(d, SA_), S<T>(d, SBX), S<T>(d, SBY),
                S<T>(d, SBZ), S<T>(d, SC)};
    T *G[3] = {S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ)};
    const int a = acc ? 1 : 0;
#define EB(L, K, B) LAUNCH_PW("ot_energy_local_bf", 2 + 2 * L + 4 * K + 8 * B + a, (k_energyB<T, L, K, B>), ei, eps, G[0], G[1], G[2], pw, a)
    if (LOC && KC && BF) EB(true, true, true);
    else if (LOC && KC) EB(true, true, false);
    else if (LOC && BF) EB(true, false, true);
    else if (LOC) EB(true, false, false);
    else if (KC && BF) EB(false, true, true);
    else if (KC) EB(false, true, false);
    else if (BF) EB(false, false, true);
#undef EB
    if (LOC && ig_on(otf)) {  // src/ot-energy.c:101-104
      const ig::Params igp = ig_params(otf);
      LAUNCH_PW("ot_ideal_gas_energy", 3, (k_ideal_gas<T, 1>), eps, (const T *)rho, (const T *)nullptr, igp, d->n);
    }
    if (KC) {
      const T *in3[3] = {S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ)};
      cplx<T> *sp[3] = {CS<T>(d, SP1), CS<T>(d, SP2), CS<T>(d, SP3)};
      r2c3d<T>(p, 3, in3, sp, s);
      const cplx<T> *cin[3] = {sp[0], sp[1], sp[2]};
      fft::Combine c[3];
      for (int dd = 0; dd < 3; ++dd) c[dd] = cmb1(dd, K_F, -1, h3N);
      T *J[3] = {S<T>(d, SBX), S<T>(d, SBY), S<T>(d, SBZ)};
      cplx<T> *tmp[3] = {CS<T>(d, TMP0), CS<T>(d, TMP1), CS<T>(d, TMP2)};
      c2r3d<T>(p, 3, cin, 3, c, kt, tmp, J, s);
      LAUNCH_PW("ot_energy_kc", 8, (k_energyC<T>), eps, J[0], J[1], J[2], G[0], G[1], G[2], (T)(-otf->alpha_s / (4.0 * otf->mass)),
                d->n);
    }
  }
}

// Backward of E[rho] (LJ, c2/c3, KC): exact discrete dE/drho.  Reverse-mode
// differentiation of the energy of src/ot-energy.c:64-93, 144-190 with the
// central difference antisymmetric and the kernels symmetric:
//   dE/drho = V_LJ*rho + c2/2 rb^2 + c3/3 rb^3 + Pi*[(c2 rb + c3 rb^2) rho]
//           + c sum_d D_d(rho_st J_d) + (c/rho_0s) F*(sum_d (D_d rho) J_d).
template <typename T> void gradient_impl(dft_ot_functional *otf, T *grad, const T *rho) {
  dft_ot_device *d = otf->device;
  Plan3D<T> &p = plan_of<T>(d);
  hipStream_t s = lib_stream();
  const double h3N = d->step * d->step * d->step / (double)d->n;
  const bool KC = (otf->model & DFT_OT_KC) != 0;
  fft::KernTab kt = kerntab<T>(d);
  PW<T> pw = make_pw<T>(otf, 1.0, otf->div_epsilon, false);
  {
    const T *in[1] = {rho};
    cplx<T> *out[1] = {CS<T>(d, SP0)};
    r2c3d<T>(p, 1, in, out, s);
  }
  {
    const cplx<T> *in[1] = {CS<T>(d, SP0)};
    fft::Combine c[2] = {cmb1(0, K_SA, -1, h3N), cmb1(0, K_F, -1, h3N)};
    T *out[2] = {S<T>(d, RB), S<T>(d, RT)};
    cplx<T> *tmp[2] = {CS<T>(d, TMP0), CS<T>(d, TMP1)};
    c2r3d<T>(p, 1, in, KC ? 2 : 1, c, kt, tmp, out, s);
  }
  {
    StageBIn bi{rho, S<T>(d, RB), S<T>(d, RT), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr};
    StageBOut bo{S<T>(d, Q4), S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ), S<T>(d, VRE), nullptr, nullptr, nullptr,
                 nullptr};
    if (KC) LAUNCH_PW("ot_stage_local_kc_bf", 10, (k_stageB<T, true, true, false>), bi, bo, pw);
    else LAUNCH_PW("ot_stage_local_kc_bf", 4, (k_stageB<T, true, false, false>), bi, bo, pw);
  }
  {
    const T *in[4] = {S<T>(d, Q4), S<T>(d, QX), S<T>(d, QY), S<T>(d, QZ)};
    cplx<T> *out[4] = {CS<T>(d, SP4), CS<T>(d, SP1), CS<T>(d, SP2), CS<T>(d, SP3)};
    r2c3d<T>(p, KC ? 4 : 1, in, out, s);
  }
  if (KC) {
    const cplx<T> *in[3] = {CS<T>(d, SP1), CS<T>(d, SP2), CS<T>(d, SP3)};
    fft::Combine c[3] = {cmb1(0, K_F, -1, h3N), cmb1(1, K_F, -1, h3N), cmb1(2, K_F, -1, h3N)};
    T *J[3] = {S<T>(d, SBX), S<T>(d, SBY), S<T>(d, SBZ)};
    cplx<T> *tmp[3] = {CS<T>(d, TMP0), CS<T>(d, TMP1), CS<T>(d, TMP2)};
    c2r3d<T>(p, 3, in, 3, c, kt, tmp, J, s);
    LAUNCH_PW("ot_grad_kc", 9, (k_grad_kc<T>), rho, S<T>(d, RT), J[0], J[1], J[2], S<T>(d, SC), S<T>(d, WX),
              S<T>(d, WY), S<T>(d, WZ), pw);
    const T *hin[1] = {S<T>(d, SC)};
    cplx<T> *hout[1] = {CS<T>(d, SH)};
    r2c3d<T>(p, 1, hin, hout, s);
  }
  {
    const cplx<T> *in[3] = {CS<T>(d, SP0), CS<T>(d, SP4), CS<T>(d, SH)};
    fft::Term t[3] = {{0, K_LJ, -1, h3N}, {1, K_SA, -1, h3N},
                      {2, K_F, -1, h3N * otf->alpha_s / (2.0 * otf->mass) / otf->rho_0s}};
    fft::Combine c = cmb(KC ? 3 : 2, t);
    T *out[1] = {S<T>(d, RB)};
    cplx<T> *tmp[1] = {CS<T>(d, TMP0)};
    c2r3d<T>(p, 3, in, 1, &c, kt, tmp, out, s);
  }
  LAUNCH_PW("ot_grad_final", KC ? 6 : 3, (k_grad_final<T>), grad, S<T>(d, VRE), S<T>(d, RB),
            KC ? S<T>(d, WX) : (T *)nullptr, S<T>(d, WY), S<T>(d, WZ), pw);
  if (ig_on(otf)) {
    const ig::Params igp = ig_params(otf);
    LAUNCH_PW("ot_ideal_gas_grad", 3, (k_ideal_gas<T, 2>), grad, rho, (const T *)nullptr, igp, d->n);
  }
}

// Build the kernel spectra (FP64 on the device), src/ot.c:138-217.
template <typename T> bool build_kernels(dft_ot_functional *otf, int min_sub, int max_sub) {
  dft_ot_device *d = otf->device;
  hipStream_t s = lib_stream();
  const int64_t nh = (int64_t)d->nx * d->ny * d->PP;
  Plan3D<double> p64;
  p64.init(d->nx, d->ny, d->nz, d->step);
  double *field = (double *)dmalloc(sizeof(double) * d->n);
  double2 *spec = (double2 *)dmalloc(sizeof(double2) * nh);
  if (!field || !spec) return false;
  KParams kp{};
  kp.lj_h = otf->lj_params.h;
  kp.lj_sigma = otf->lj_params.sigma;
  kp.lj_eps = otf->lj_params.epsilon;
  kp.lj_cval = otf->lj_params.cval;
  // spherical-average radius: h, or 1.065 h for HD2 (src/ot.c:164-170)
  kp.sa_radius = (otf->model & DFT_OT_HD2) ? otf->lj_params.h * 1.065 : otf->lj_params.h;
  kp.g_inv_width = 1.0 / otf->l_g;
  kp.g11 = otf->bf_params.g11;
  kp.g12 = otf->bf_params.g12;
  kp.g21 = otf->bf_params.g21;
  kp.g22 = otf->bf_params.g22;
  kp.a1 = otf->bf_params.a1;
  kp.a2 = otf->bf_params.a2;
  const double tol = 0.01 / GRID_AUTOK;
  const double h3 = d->step * d->step * d->step;
  const bool need[4] = {true, true, (otf->model & DFT_OT_KC) != 0, (otf->model & DFT_OT_BACKFLOW) != 0};
  for (int w = 0; w < 4; ++w) {
    if (!need[w]) continue;
    hipLaunchKernelGGL(k_kernel_map, dim3(grid_for(d->n)), dim3(kBlock), 0, s, field, d->nx, d->ny, d->nz, d->step,
                       w, kp, min_sub, max_sub, tol);
    DFT_LAUNCH_CHECK();
    const double *in[1] = {field};
    double2 *out[1] = {spec};
    r2c3d<double>(p64, 1, in, out, s);
    double dc[2];
    DFT_HIP_CHECK(hipMemcpyAsync(dc, spec, sizeof(double2), hipMemcpyDeviceToHost, s));
    DFT_HIP_CHECK(hipStreamSynchronize(s));
    double scale = 1.0;
    if (w == K_LJ) scale = otf->b / (h3 * dc[0]);   // integral of V_LJ := b (src/ot.c:160)
    if (w == K_SA) scale = 1.0 / (h3 * dc[0]);      // integral of Pi := 1 (src/ot.c:181)
    hipLaunchKernelGGL(k_spec_real<T>, dim3(grid_for(nh)), dim3(kBlock), 0, s, spec, (T *)d->kern[w], nh, scale);
    DFT_LAUNCH_CHECK();
  }
  DFT_HIP_CHECK(hipStreamSynchronize(s));
  dfree(field);
  dfree(spec);
  p64.destroy();
  return true;
}

void free_device(dft_ot_device *d) {
  if (!d) return;
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  for (int i = 0; i < NSLOT; ++i) dfree(d->slot[i]);
  for (int i = 0; i < 4; ++i) dfree(d->kern[i]);
  dfree(d->phase);
  dfree(d->vsave);
  dfree(d->norm_part);
  dfree(d->igtab);
  dfree(d->mid_dev);
  if (d->plan) {
    if (d->prec == GRID_FP32) {
      ((Plan3D<float> *)d->plan)->destroy();
      delete (Plan3D<float> *)d->plan;
    } else {
      ((Plan3D<double> *)d->plan)->destroy();
      delete (Plan3D<double> *)d->plan;
    }
  }
  delete d;
}

rgrid *spectral_handle(dft_ot_functional *otf, int w, const char *id) {
  dft_ot_device *d = otf->device;
  rgrid *g = (rgrid *)calloc(1, sizeof(rgrid));
  g->nx = d->nx;
  g->ny = d->ny;
  g->nz = d->nz;
  g->step = d->step;
  g->dev = d->kern[w];
  g->dev_elems = (int64_t)d->nx * d->ny * d->PP;
  g->prec = d->prec;
  g->spectral = 1;
  strncpy(g->id, id, sizeof(g->id) - 1);
  return g;
}

// PLAIN, KC and BACKFLOW, and the high-density penalty (HD, HD2) without
// backflow: with HD the reference also switches the backflow to another
// density dependence (src/ot.c:843-982), which is not built here.
bool supported_model(INT model) {
  // PLAIN, KC, HD, HD2, BACKFLOW and the parameter rows T0MK ... T3000MK (bits 16 ... 262144)
  return !(model & ~(INT)(DFT_OT_KC | DFT_OT_BACKFLOW | DFT_OT_HD | DFT_OT_HD2 | ((INT)524288 - 16)));
}

// Argument checks of the entry points: false (with a libdft: message) when
// a handle is missing or its shape or precision differs from the functional's.
bool check_otf(dft_ot_functional *otf, const char *fn) {
  if (!otf || !otf->device) {
    fprintf(stderr, "libdft: %s: no functional\n", fn);
    return false;
  }
  return true;
}

bool check_wf(dft_ot_functional *otf, wf *gwf, const char *fn) {
  if (!check_otf(otf, fn)) return false;
  dft_ot_device *d = otf->device;
  cgrid *g = gwf ? gwf->grid : nullptr;
  if (!g || g->nx != d->nx || g->ny != d->ny || g->nz != d->nz || g->prec != d->prec) {
    fprintf(stderr, "libdft: %s: wave function does not match the functional's grid\n", fn);
    return false;
  }
  return true;
}

bool check_c(dft_ot_functional *otf, cgrid *g, const char *fn) {
  if (!check_otf(otf, fn)) return false;
  dft_ot_device *d = otf->device;
  if (!g || g->nx != d->nx || g->ny != d->ny || g->nz != d->nz || g->prec != d->prec) {
    fprintf(stderr, "libdft: %s: grid does not match the functional's grid\n", fn);
    return false;
  }
  return true;
}

bool check_r(dft_ot_functional *otf, rgrid *g, const char *fn) {
  if (!check_otf(otf, fn)) return false;
  dft_ot_device *d = otf->device;
  if (!g || g->nx != d->nx || g->ny != d->ny || g->nz != d->nz || g->prec != d->prec || g->spectral) {
    fprintf(stderr, "libdft: %s: grid does not match the functional's grid\n", fn);
    return false;
  }
  return true;
}


// -------------------------------------------------------- adjoint sweep --
// Work space of the reverse pass: real fields, complex fields and half spectra,
// and for backflow models the fields of the backflow transpose.
template <typename T> struct AdjWork {
  int64_t n = 0, nh = 0;
  bool bf = false;
  T *r[24] = {};
  T *b[20] = {};
  cplx<T> *c[4] = {};
  cplx<T> *spec[10] = {};
  double *pa = nullptr, *pw = nullptr;
  bool alloc(int64_t n_, int64_t nh_, bool with_bf) {
    n = n_;
    nh = nh_;
    bf = with_bf;
    for (auto &x : r)
      if (!(x = (T *)dmalloc(sizeof(T) * n))) return false;
    if (bf)
      for (auto &x : b)
        if (!(x = (T *)dmalloc(sizeof(T) * n))) return false;
    for (auto &x : c)
      if (!(x = (cplx<T> *)dmalloc(sizeof(cplx<T>) * n))) return false;
    for (auto &x : spec)
      if (!(x = (cplx<T> *)dmalloc(sizeof(cplx<T>) * nh))) return false;
    pa = (double *)dmalloc(sizeof(double) * red_blocks(n));
    pw = (double *)dmalloc(sizeof(double) * red_blocks(n));
    return pa && pw;
  }
  void release() {
    for (auto &x : r) dfree(x);
    for (auto &x : b) dfree(x);
    for (auto &x : c) dfree(x);
    for (auto &x : spec) dfree(x);
    dfree(pa);
    dfree(pw);
  }
};

// Real fields of the work space.
enum { AQB, ALV, AHP, ARP, ARBB, AT1B, APB, AJB = APB + 3, AGB = AJB + 3, ARSB = AGB + 3, AGBAR, ACONV = AGBAR + 3,
       AVBAR, AVEFF, ARHO, AGRAD };
// Backflow fields: Vi_bar, tb, imb, Abar, Bbar_d, gbar_d, vbar_d, rho_bar, U*Abar,
// U*Bbar_d, U*Cbar, Im V_OT of the step.
enum { BVBI, BTB, BIMB, BAB, BBB, BGB = BBB + 3, BVB = BGB + 3, BRB = BVB + 3, BUA, BUB, BUC = BUB + 3, BVIM };

// out = sum_m K_m * in_m over real fields (r2c of the inputs, one c2r output
// per combine; at most 5 of each).
template <typename T>
void adj_conv(dft_ot_device *d, AdjWork<T> &w, int nin, const T *const *in, int nout, const fft::Combine *c,
              T *const *out) {
  Plan3D<T> &p = plan_of<T>(d);
  hipStream_t s = lib_stream();
  cplx<T> *si[5] = {w.spec[0], w.spec[1], w.spec[2], w.spec[3], w.spec[4]};
  r2c3d<T>(p, nin, in, si, s);
  const cplx<T> *ci[5] = {w.spec[0], w.spec[1], w.spec[2], w.spec[3], w.spec[4]};
  cplx<T> *tmp[5] = {w.spec[5], w.spec[6], w.spec[7], w.spec[8], w.spec[9]};
  c2r3d<T>(p, nin, ci, nout, c, kerntab<T>(d), tmp, out, s);
}

// R = (dV_OT/drho)^T vbar (real parts: LJ, local, KC, HD) at the state whose
// intermediates potential_impl left in the slots (rho: the density it used).
// psib += 2 R psi, or rout = R when psib is null.
template <typename T>
void vjp_ot(dft_ot_functional *otf, const T *vbar, const T *rho, cplx<T> *psib, const cplx<T> *psi, T *rout,
            AdjWork<T> &w) {
  dft_ot_device *d = otf->device;
  const double h3N = d->step * d->step * d->step / (double)d->n;
  const bool KC = (otf->model & DFT_OT_KC) != 0;
  PW<T> pw = make_pw<T>(otf, 1.0, otf->div_epsilon, false);
  {
    const T *in[1] = {vbar};
    fft::Combine c[3] = {cmb1(0, K_SA, -1, h3N), cmb1(0, K_LJ, -1, h3N), cmb1(0, K_F, -1, h3N)};
    T *out[3] = {w.r[AQB], w.r[ALV], w.r[AHP]};
    adj_conv<T>(d, w, 1, in, KC ? 3 : 2, c, out);
  }
  VjpA a{};
  a.vbar = vbar;
  a.rho = rho;
  a.rb = S<T>(d, RB);
  a.qb = w.r[AQB];
  a.lv = w.r[ALV];
  a.hp = w.r[AHP];
  a.rt = S<T>(d, RT);
  a.t1 = S<T>(d, Q4);
  for (int k = 0; k < 3; ++k) {
    a.J[k] = S<T>(d, QX + k);
    a.Pd[k] = S<T>(d, PX + k);
    a.pb[k] = w.r[APB + k];
    a.jb[k] = w.r[AJB + k];
    a.gb[k] = w.r[AGB + k];
  }
  a.rp = w.r[ARP];
  a.rbb = w.r[ARBB];
  a.t1b = w.r[AT1B];
  a.rsb = w.r[ARSB];
  if (KC) {
    LAUNCH_PW("ot_adjoint_a", 22, (k_vjp_a<T, true>), a, pw);
    LAUNCH_PW("ot_adjoint_b", 7, (k_vjp_b<T>), a, pw);
    {
      const T *in[3] = {w.r[AJB], w.r[AJB + 1], w.r[AJB + 2]};
      fft::Combine c[3] = {cmb1(0, K_F, -1, h3N), cmb1(1, K_F, -1, h3N), cmb1(2, K_F, -1, h3N)};
      T *out[3] = {w.r[AGBAR], w.r[AGBAR + 1], w.r[AGBAR + 2]};
      adj_conv<T>(d, w, 3, in, 3, c, out);
    }
    VjpA ac = a;
    for (int k = 0; k < 3; ++k) ac.J[k] = w.r[AGBAR + k];
    LAUNCH_PW("ot_adjoint_c", 12, (k_vjp_c<T>), ac, pw);
    LAUNCH_PW("ot_adjoint_d", 10, (k_vjp_d<T>), a, pw);
    const T *in[2] = {w.r[ARBB], w.r[ARSB]};
    fft::Term t[2] = {{0, K_SA, -1, h3N}, {1, K_F, -1, h3N}};
    fft::Combine c = cmb(2, t);
    T *out[1] = {w.r[ACONV]};
    adj_conv<T>(d, w, 2, in, 1, &c, out);
  } else {
    LAUNCH_PW("ot_adjoint_a", 6, (k_vjp_a<T, false>), a, pw);
    const T *in[1] = {w.r[ARBB]};
    fft::Combine c = cmb1(0, K_SA, -1, h3N);
    T *out[1] = {w.r[ACONV]};
    adj_conv<T>(d, w, 1, in, 1, &c, out);
  }
  if (ig_on(otf)) {
    const ig::Params igp = ig_params(otf);
    LAUNCH_PW("ot_adjoint_ideal_gas", 3, (k_ideal_gas<T, 3>), w.r[ARP], rho, vbar, igp, d->n);
  }
  LAUNCH_PW("ot_adjoint_out", 4, (k_vjp_out<T>), rout, psib, psi, (const T *)w.r[ARP], (const T *)w.r[ACONV], d->n);
}

// Backflow part: psib += (dV_BF/dpsi)^T (vrb, vib), the adjoints of Re V and
// Im V, at the state whose intermediates potential_impl left in the slots
// (v_d, A, B_d, C, W_d; rho: the density it used, mass: the wave function's).
template <typename T>
void vjp_bf(dft_ot_functional *otf, const T *vrb, const T *vib, const T *rho, cplx<T> *psib, const cplx<T> *psi,
            double mass, AdjWork<T> &w) {
  dft_ot_device *d = otf->device;
  const double h3N = d->step * d->step * d->step / (double)d->n;
  PW<T> pw = make_pw<T>(otf, mass, otf->div_epsilon, true);
  BfAdj a{};
  a.vrb = vrb;
  a.vib = vib;
  a.rho = rho;
  a.A = S<T>(d, SA_);
  for (int k = 0; k < 3; ++k) {
    a.v[k] = S<T>(d, VX + k);
    a.W[k] = S<T>(d, WX + k);
    a.Bb[k] = w.b[BBB + k];
    a.gb[k] = w.b[BGB + k];
    a.vb[k] = w.b[BVB + k];
    a.uB[k] = w.b[BUB + k];
  }
  a.tb = w.b[BTB];
  a.imb = w.b[BIMB];
  a.Ab = w.b[BAB];
  a.rb = w.b[BRB];
  a.uA = w.b[BUA];
  a.uC = w.b[BUC];
  LAUNCH_PW("ot_adjoint_bf_a", 13, (k_bfadj_a<T>), a, (const T *)S<T>(d, SBX), (const T *)S<T>(d, SBY),
            (const T *)S<T>(d, SBZ), (const T *)S<T>(d, SC), pw);
  LAUNCH_PW("ot_adjoint_bf_b", 19, (k_bfadj_b<T>), a, pw);
  {
    const T *in[5] = {w.b[BAB], w.b[BBB], w.b[BBB + 1], w.b[BBB + 2], w.b[BTB]};
    fft::Combine c[5];
    for (int i = 0; i < 5; ++i) c[i] = cmb1(i, K_U, -1, h3N);
    T *out[5] = {w.b[BUA], w.b[BUB], w.b[BUB + 1], w.b[BUB + 2], w.b[BUC]};
    adj_conv<T>(d, w, 5, in, 5, c, out);
  }
  LAUNCH_PW("ot_adjoint_bf_c", 19, (k_bfadj_c<T>), a, pw);
  LAUNCH_PW("ot_adjoint_bf_psi", 8, (k_bfadj_psi<T>), psib, psi, (const T *)w.b[BRB], (const T *)w.b[BVB],
            (const T *)w.b[BVB + 1], (const T *)w.b[BVB + 2], pw);
}

// V_OT of psi into the complex field w.c[3] (the unfused path, leaving its
// intermediates in the slots), veff = Re V_OT + V_ext - mu and, for backflow,
// Im V_OT in w.b[BVIM].
template <typename T>
void adj_potential(dft_ot_functional *otf, const cplx<T> *psi, double mass, const Well &wl, double mu, AdjWork<T> &w) {
  dft_ot_device *d = otf->device;
  hipStream_t s = lib_stream();
  DFT_HIP_CHECK(hipMemsetAsync(w.c[3], 0, sizeof(cplx<T>) * d->n, s));
  cgrid V{};
  V.nx = d->nx;
  V.ny = d->ny;
  V.nz = d->nz;
  V.step = d->step;
  V.dev = w.c[3];
  V.dev_elems = d->n;
  V.prec = d->prec;
  PotParts parts{true, (otf->model & DFT_OT_KC) != 0, w.bf};
  potential_impl<T>(otf, &V, psi, mass, w.r[ARHO], nullptr, nullptr, nullptr, parts, false);
  hipLaunchKernelGGL(k_adj_veff<T>, dim3(grid_for(d->n)), dim3(kBlock), 0, s, w.r[AVEFF], w.bf ? w.b[BVIM] : nullptr,
                     (const cplx<T> *)w.c[3], wl, mu, d->n);
  DFT_LAUNCH_CHECK();
}

// rb = h3 Re V, vb_d = -mcb h3 rho W_d: the gradient of the backflow energy
// with respect to rho (at fixed v) and to v_d.
template <typename T>
__global__ void k_bfseed(T *rb, T *vb0, T *vb1, T *vb2, const cplx<T> *V, const T *rho, const T *W0, const T *W1,
                         const T *W2, double h3, double mcb, int64_t n, PW<T> pw) {
  GRID_STRIDE(c, n) {
    rb[c] = (T)(h3 * (double)V[c].x);
    const double f = -mcb * h3 * (double)(pw.bfhd ? bf_rt(rho[c], pw) : rho[c]);
    vb0[c] = (T)(f * (double)W0[c]);
    vb1[c] = (T)(f * (double)W1[c]);
    vb2[c] = (T)(f * (double)W2[c]);
  }
}

// psib += dE_BF/dpsi: the backflow energy of src/ot-energy.c:202-288 (velocity
// with DFT_EPS, no clamps) depends on psi through rho and v:
// psib += 2 h^3 Re V_BF psi + (dv/dpsi)^T vb, vb_d = -m c_bf h^3 rho W_d
// (rho~ = G rho in place of rho with the HD / HD2 backflow switch).
template <typename T>
void seed_bf(dft_ot_functional *otf, const cplx<T> *psi, double mass, cplx<T> *psib, AdjWork<T> &w) {
  dft_ot_device *d = otf->device;
  hipStream_t s = lib_stream();
  const double h3 = d->step * d->step * d->step;
  const double keep[3] = {otf->div_epsilon, otf->max_veloc, otf->max_bfpot};
  otf->div_epsilon = DFT_EPS;
  otf->max_veloc = 999.0 / GRID_AUTOMPS;
  otf->max_bfpot = 10000.0 / GRID_AUTOK;
  const PW<T> pw = make_pw<T>(otf, mass, DFT_EPS, false);
  DFT_HIP_CHECK(hipMemsetAsync(w.c[3], 0, sizeof(cplx<T>) * d->n, s));
  cgrid V{};
  V.nx = d->nx;
  V.ny = d->ny;
  V.nz = d->nz;
  V.step = d->step;
  V.dev = w.c[3];
  V.dev_elems = d->n;
  V.prec = d->prec;
  PotParts parts{false, false, true};
  potential_impl<T>(otf, &V, psi, mass, w.r[ARHO], nullptr, nullptr, nullptr, parts, false);
  otf->div_epsilon = keep[0];
  otf->max_veloc = keep[1];
  otf->max_bfpot = keep[2];
  LAUNCH_PW("ot_adjoint_bf_seed", 8, (k_bfseed<T>), w.b[BRB], w.b[BVB], w.b[BVB + 1], w.b[BVB + 2],
            (const cplx<T> *)w.c[3], (const T *)w.r[ARHO], (const T *)S<T>(d, WX), (const T *)S<T>(d, WY),
            (const T *)S<T>(d, WZ), h3, otf->mass * otf->c_bfpot, d->n, pw);
  LAUNCH_PW("ot_adjoint_bf_psi", 8, (k_bfadj_psi<T>), psib, psi, (const T *)w.b[BRB], (const T *)w.b[BVB],
            (const T *)w.b[BVB + 1], (const T *)w.b[BVB + 2], pw);
}

// One step backwards: psib holds dJ/dpsi_{n+1} on entry and dJ/dpsi_n on exit;
// th[0..1] accumulate dJ/d(amp, width) through V_ext of this step.  k0: the
// momentum frame of the kinetic propagator.
template <typename T>
void adj_step(dft_ot_functional *otf, const cplx<T> *psi, cplx<T> *psib, double mass, const Well &wl, double mu,
              double dtr, double dti, const double *k0, AdjWork<T> &w, double th[2]) {
  dft_ot_device *d = otf->device;
  Plan3D<T> &p = plan_of<T>(d);
  hipStream_t s = lib_stream();
  const dim3 g(grid_for(d->n)), b(kBlock);
  adj_potential<T>(otf, psi, mass, wl, mu, w);
  const T *vi = w.bf ? w.b[BVIM] : nullptr;
  // forward: u1 = phi psi, u2 = T u1
  hipLaunchKernelGGL(k_adj_phase<T>, g, b, 0, s, w.c[0], psi, (const T *)w.r[AVEFF], vi, dtr, dti, d->n);
  DFT_LAUNCH_CHECK();
  strang<T>(p, w.c[0], nullptr, dtr, dti, mass, nullptr, s, k0);
  // reverse: u2bar, phibar; u1bar = T^dagger u2bar (T at -conj(dt)); psibar, Vbar
  hipLaunchKernelGGL(k_adj_out<T>, g, b, 0, s, w.c[1], w.c[2], (const cplx<T> *)psib, (const cplx<T> *)w.c[0],
                     (const T *)w.r[AVEFF], vi, dtr, dti, d->n);
  DFT_LAUNCH_CHECK();
  strang<T>(p, w.c[1], nullptr, -dtr, dti, mass, nullptr, s, k0);
  hipLaunchKernelGGL(k_adj_in<T>, dim3((unsigned)red_blocks(d->n)), b, 0, s, psib, w.r[AVBAR],
                     w.bf ? w.b[BVBI] : nullptr, (const cplx<T> *)w.c[1], psi, (const cplx<T> *)w.c[2],
                     (const T *)w.r[AVEFF], vi, dtr, dti, wl, w.pa, w.pw, d->n);
  DFT_LAUNCH_CHECK();
  th[0] += sum_partials(w.pa, red_blocks(d->n), s);
  th[1] += sum_partials(w.pw, red_blocks(d->n), s);
  vjp_ot<T>(otf, w.r[AVBAR], w.r[ARHO], psib, psi, nullptr, w);
  if (w.bf) vjp_bf<T>(otf, w.r[AVBAR], w.b[BVBI], w.r[ARHO], psib, psi, mass, w);
}

// Progress lines of the long-running sweeps on stderr (dft_set_verbose).
int g_verbose = 0;

// Checkpointed adjoint of K single steps of dft_ot_propagate from the state in
// gwf (left unchanged), with the well wl as V_ext.  The forward pass keeps every
// `every`-th state; the reverse pass recomputes each segment from its
// checkpoint.  seed(psi_K, psib, th) sets psib = dJ/dpsi_K (the convention
// dJ = Re sum conj(psib) dpsi) and may start th; on return psib holds dJ/dpsi_0
// and th[0..1] has accumulated dJ/d(amp, width) through V_ext.
template <typename T, class Seed>
int adjoint_sweep(dft_ot_functional *otf, wf *gwf, const Well &wl, double mu, double dtr, double dti, int64_t K,
                  int64_t every, AdjWork<T> &w, cplx<T> *psib, double th[2], Seed seed) {
  dft_ot_device *d = otf->device;
  hipStream_t s = lib_stream();
  const int64_t n = d->n;
  const size_t bytes = sizeof(cplx<T>) * n;
  const int64_t ncp = (K + every - 1) / every;
  std::vector<cplx<T> *> cp(ncp, nullptr), seg(every, nullptr);
  bool ok = true;
  for (auto &x : cp) ok = ok && (x = (cplx<T> *)dmalloc(bytes)) != nullptr;
  for (auto &x : seg) ok = ok && (x = (cplx<T> *)dmalloc(bytes)) != nullptr;
  wf *cur = grid_wf_clone(gwf, (char *)"adjoint state");
  rgrid *vext = rgrid_clone(otf->density, (char *)"adjoint well");
  auto cleanup = [&]() {
    DFT_HIP_CHECK(hipStreamSynchronize(s));
    for (auto x : cp) dfree(x);
    for (auto x : seg) dfree(x);
    if (cur) grid_wf_free(cur);
    if (vext) rgrid_free(vext);
  };
  if (!ok || !cur || !vext) {
    cleanup();
    return -1;
  }
  hipLaunchKernelGGL(k_well<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, (T *)vext->dev, wl, n);
  DFT_LAUNCH_CHECK();
  REAL_COMPLEX dt;
  __real__ dt = dtr;
  __imag__ dt = dti;
  const double mass = gwf->mass;
  const double k0[3] = {gwf->grid->kx0, gwf->grid->ky0, gwf->grid->kz0};
  // forward sweep
  DFT_HIP_CHECK(hipMemcpyAsync(cur->grid->dev, gwf->grid->dev, bytes, hipMemcpyDeviceToDevice, s));
  for (int64_t it = 0; it < K; ++it) {
    if (it % every == 0) {
      DFT_HIP_CHECK(hipMemcpyAsync(cp[it / every], cur->grid->dev, bytes, hipMemcpyDeviceToDevice, s));
      if (g_verbose) {
        DFT_HIP_CHECK(hipStreamSynchronize(s));
        fprintf(stderr, "libdft: adjoint forward step %lld / %lld\n", (long long)it, (long long)K);
      }
    }
    dft_ot_propagate(otf, cur, vext, mu, dt);
  }
  seed(cur, psib, th);
  // reverse sweep, segment by segment
  for (int64_t j = ncp - 1; j >= 0; --j) {
    const int64_t n0 = j * every, n1 = std::min(K, n0 + every);
    DFT_HIP_CHECK(hipMemcpyAsync(cur->grid->dev, cp[j], bytes, hipMemcpyDeviceToDevice, s));
    for (int64_t it = n0; it < n1; ++it) {
      DFT_HIP_CHECK(hipMemcpyAsync(seg[it - n0], cur->grid->dev, bytes, hipMemcpyDeviceToDevice, s));
      if (it + 1 < n1) dft_ot_propagate(otf, cur, vext, mu, dt);
    }
    for (int64_t it = n1 - 1; it >= n0; --it)
      adj_step<T>(otf, seg[it - n0], psib, mass, wl, mu, dtr, dti, k0, w, th);
    if (g_verbose) fprintf(stderr, "libdft: adjoint reverse segment %lld / %lld done\n", (long long)(ncp - j), (long long)ncp);
  }
  cleanup();
  return 0;
}

// J = E_kin + E_OT + E_ext of psi_K and dJ/d(amp, width) of the well.
template <typename T>
int adjoint_well_impl(dft_ot_functional *otf, wf *gwf, const Well &wl, double mu, double dtr, double dti, int64_t K,
                      int64_t every, double *E, double grad[2]) {
  dft_ot_device *d = otf->device;
  hipStream_t s = lib_stream();
  const int64_t n = d->n;
  const bool BF = (otf->model & DFT_OT_BACKFLOW) != 0;
  AdjWork<T> w;
  cplx<T> *psib = nullptr;
  bool ok = w.alloc(n, (int64_t)d->nx * d->ny * d->PP, BF) && (psib = (cplx<T> *)dmalloc(sizeof(cplx<T>) * n));
  double th[2] = {0.0, 0.0};
  auto seed = [&](wf *cur, cplx<T> *pb, double t2[2]) {
    // the gradient of the total energy with respect to psi_K
    Plan3D<T> &p = plan_of<T>(d);
    const cplx<T> *psiK = (const cplx<T> *)cur->grid->dev;
    const double h3 = d->step * d->step * d->step;
    ew_density<T>(psiK, w.r[ARHO], n, s);
    gradient_impl<T>(otf, w.r[AGRAD], w.r[ARHO]);
    c2c_fwd<T>(p, psiK, w.c[0], s);
    const double cx = std::pow(2.0 * M_PI / (d->nx * d->step), 2) / (2.0 * gwf->mass);
    const double cy = std::pow(2.0 * M_PI / (d->ny * d->step), 2) / (2.0 * gwf->mass);
    const double cz = std::pow(2.0 * M_PI / (d->nz * d->step), 2) / (2.0 * gwf->mass);
    const double o[3] = {cur->grid->kx0 * d->nx * d->step / (2.0 * M_PI), cur->grid->ky0 * d->ny * d->step / (2.0 * M_PI),
                         cur->grid->kz0 * d->nz * d->step / (2.0 * M_PI)};
    hipLaunchKernelGGL(k_kin_mul<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, w.c[0], d->nx, d->ny, d->nz, cx, cy, cz,
                       o[0], o[1], o[2]);
    DFT_LAUNCH_CHECK();
    c2c_inv<T>(p, w.c[0], w.c[1], s);
    hipLaunchKernelGGL(k_adj_seed<T>, dim3((unsigned)red_blocks(n)), dim3(kBlock), 0, s, pb, (const cplx<T> *)w.c[1],
                       psiK, (const T *)w.r[AGRAD], wl, h3, w.pa, w.pw, n);
    DFT_LAUNCH_CHECK();
    t2[0] = h3 * sum_partials(w.pa, red_blocks(n), s);
    t2[1] = h3 * sum_partials(w.pw, red_blocks(n), s);
    // the backflow energy also depends on psi through the velocity
    if (BF) seed_bf<T>(otf, psiK, gwf->mass, pb, w);
    *E = grid_wf_energy(cur, nullptr) + dft_ot_energy(otf, cur) + wl.amp * t2[0];
  };
  const int rc = ok ? adjoint_sweep<T>(otf, gwf, wl, mu, dtr, dti, K, every, w, psib, th, seed) : -1;
  if (rc) fprintf(stderr, "libdft: dft_ot_adjoint_well: out of device memory\n");
  grad[0] = th[0];
  grad[1] = th[1];
  DFT_HIP_CHECK(hipStreamSynchronize(s));
  dfree(psib);
  w.release();
  return rc;
}

// psi_0 = sqrt(rho0 (1 + tanh((r - R0)/w)) / 2): bulk with a spherical cavity.
template <typename T> __global__ void k_cavity(cplx<T> *psi, Well g, double rho0, double R0, double wd, int64_t n) {
  GRID_STRIDE(c, n) {
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((int64_t)g.nz * g.ny));
    const double x = (i - g.nx / 2) * g.h, y = (j - g.ny / 2) * g.h, z = (k - g.nz / 2) * g.h;
    const double r = sqrt(x * x + y * y + z * z);
    psi[c] = cmk<cplx<T>>((T)sqrt(0.5 * rho0 * (1.0 + tanh((r - R0) / wd))), (T)0);
  }
}

// Window of the bubble measure: W(r) = (1 - tanh((r - Rc)/wc)) / 2.  The
// missing volume of the whole box is the conserved norm (N / rho0) and says
// nothing about where the deficit is; inside the window it is the bubble's.
// Also dW/dRc = (1 - tanh^2) / (2 wc).
__device__ inline double cavity_window(const Well &g, int64_t c, double Rc, double wc, double *dW = nullptr) {
  const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((int64_t)g.nz * g.ny));
  const double x = (i - g.nx / 2) * g.h, y = (j - g.ny / 2) * g.h, z = (k - g.nz / 2) * g.h;
  const double th = tanh((sqrt(x * x + y * y + z * z) - Rc) / wc);
  if (dW) *dW = (1.0 - th * th) / (2.0 * wc);
  return 0.5 * (1.0 - th);
}

// Per block: sum W (1 - |psi|^2 / rho0) (the missing volume in the window over
// h^3) and sum dW/dRc (1 - |psi|^2 / rho0).
template <typename T>
__global__ void k_cavity_volume(const cplx<T> *psi, Well g, double Rc, double wc, double rho0, double *pa, double *pb,
                                int64_t n) {
  double sa = 0.0, sb = 0.0;
  GRID_STRIDE(c, n) {
    const double pr = psi[c].x, pi = psi[c].y, m = 1.0 - (pr * pr + pi * pi) / rho0;
    double dW;
    sa += cavity_window(g, c, Rc, wc, &dW) * m;
    sb += dW * m;
  }
  block_sum2(sa, sb, pa, pb);
}

// psib = f W psi (f real).
template <typename T>
__global__ void k_cavity_seed(cplx<T> *psib, const cplx<T> *psi, Well g, double Rc, double wc, double f, int64_t n) {
  GRID_STRIDE(c, n) {
    const double fw = f * cavity_window(g, c, Rc, wc);
    psib[c] = cmk<cplx<T>>((T)(fw * (double)psi[c].x), (T)(fw * (double)psi[c].y));
  }
}

// Per block: sum Re(conj(psib) dpsi_0/dR0) with dpsi_0/dR0 = -psi_0 (1 - tanh s) / (2w), s = (r - R0)/w.
template <typename T>
__global__ void k_cavity_grad(const cplx<T> *psib, Well g, double rho0, double R0, double wd, double *pa, double *pb,
                              int64_t n) {
  double sa = 0.0;
  GRID_STRIDE(c, n) {
    const int k = (int)(c % g.nz), j = (int)((c / g.nz) % g.ny), i = (int)(c / ((int64_t)g.nz * g.ny));
    const double x = (i - g.nx / 2) * g.h, y = (j - g.ny / 2) * g.h, z = (k - g.nz / 2) * g.h;
    const double th = tanh((sqrt(x * x + y * y + z * z) - R0) / wd);
    const double p0 = sqrt(0.5 * rho0 * (1.0 + th));
    sa += (double)psib[c].x * (-p0 * (1.0 - th) / (2.0 * wd));
  }
  block_sum2(sa, 0.0, pa, pb);
}

// Released cavity: psi_0(R0) written to gwf, K steps without V_ext, J = the
// bubble radius (3 V_b / 4 pi)^(1/3), V_b = h^3 sum W (1 - rho / rho0) with the
// window W of radius Rc = 1.5 R0 and wall wc = 4 w, and dJ/dR0: the path
// through psi_0 (adjoint sweep) plus the window's own dependence on R0.
template <typename T>
int adjoint_cavity_impl(dft_ot_functional *otf, wf *gwf, double rho0, double R0, double wd, double mu, double dtr,
                        double dti, int64_t K, int64_t every, double *J, double *grad) {
  dft_ot_device *d = otf->device;
  hipStream_t s = lib_stream();
  const int64_t n = d->n;
  const double h3 = d->step * d->step * d->step;
  const Well geo{0.0, 1.0, 0.0, 0.0, 0.0, d->step, d->nx, d->ny, d->nz};
  const double Rc = 1.5 * R0, wc = 4.0 * wd;
  hipLaunchKernelGGL(k_cavity<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, (cplx<T> *)gwf->grid->dev, geo, rho0, R0, wd,
                     n);
  DFT_LAUNCH_CHECK();
  AdjWork<T> w;
  cplx<T> *psib = nullptr;
  bool ok = w.alloc(n, (int64_t)d->nx * d->ny * d->PP, (otf->model & DFT_OT_BACKFLOW) != 0) &&
            (psib = (cplx<T> *)dmalloc(sizeof(cplx<T>) * n));
  double th[2] = {0.0, 0.0}, explicit_term = 0.0;
  auto seed = [&](wf *cur, cplx<T> *pb, double *) {
    const cplx<T> *psiK = (const cplx<T> *)cur->grid->dev;
    hipLaunchKernelGGL(k_cavity_volume<T>, dim3((unsigned)red_blocks(n)), dim3(kBlock), 0, s, psiK, geo, Rc, wc, rho0,
                       w.pa, w.pw, n);
    DFT_LAUNCH_CHECK();
    const double V = h3 * sum_partials(w.pa, red_blocks(n), s);
    *J = std::cbrt(3.0 * V / (4.0 * M_PI));
    // dJ/dR0 at fixed psi_K: (J / 3V) h^3 sum dW/dRc (1 - rho/rho0) dRc/dR0
    explicit_term = (*J / (3.0 * V)) * h3 * sum_partials(w.pw, red_blocks(n), s) * 1.5;
    // dJ/dpsi = (J / 3V) dV/dpsi, dV/dpsi = -2 h^3 W psi / rho0
    hipLaunchKernelGGL(k_cavity_seed<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, pb, psiK, geo, Rc, wc,
                       (*J / (3.0 * V)) * (-2.0 * h3 / rho0), n);
    DFT_LAUNCH_CHECK();
  };
  int rc = ok ? adjoint_sweep<T>(otf, gwf, geo, mu, dtr, dti, K, every, w, psib, th, seed) : -1;
  if (rc) {
    fprintf(stderr, "libdft: dft_ot_adjoint_cavity: out of device memory\n");
  } else {
    hipLaunchKernelGGL(k_cavity_grad<T>, dim3((unsigned)red_blocks(n)), dim3(kBlock), 0, s, (const cplx<T> *)psib, geo,
                       rho0, R0, wd, w.pa, w.pw, n);
    DFT_LAUNCH_CHECK();
    *grad = sum_partials(w.pa, red_blocks(n), s) + explicit_term;
  }
  DFT_HIP_CHECK(hipStreamSynchronize(s));
  dfree(psib);
  w.release();
  return rc;
}

// Standalone reverse pass of V_OT at the density of gwf (for tests): rho_bar = (dV/drho)^T v_bar.
template <typename T> bool potential_vjp_impl(dft_ot_functional *otf, T *rho_bar, const T *v_bar, wf *gwf) {
  dft_ot_device *d = otf->device;
  AdjWork<T> w;
  if (!w.alloc(d->n, (int64_t)d->nx * d->ny * d->PP, false)) {
    fprintf(stderr, "libdft: dft_ot_potential_vjp: out of device memory\n");
    w.release();
    return false;
  }
  Well none{0.0, 1.0, 0.0, 0.0, 0.0, d->step, d->nx, d->ny, d->nz};
  adj_potential<T>(otf, (const cplx<T> *)gwf->grid->dev, gwf->mass, none, 0.0, w);
  vjp_ot<T>(otf, v_bar, w.r[ARHO], nullptr, nullptr, rho_bar, w);
  DFT_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  w.release();
  return true;
}

// re = Re c, im = Im c
template <typename T> __global__ void k_split(T *re, T *im, const cplx<T> *c, int64_t n) {
  GRID_STRIDE(i, n) {
    re[i] = c[i].x;
    if (im) im[i] = c[i].y;
  }
}

// Standalone reverse pass of the complex V_OT with respect to psi (every model):
// psi_bar = (dV/dpsi)^T v_bar, with Re v_bar the adjoint of Re V and Im v_bar
// that of Im V (backflow), in the convention dJ = Re sum conj(psi_bar) dpsi.
template <typename T> bool potential_vjp_psi_impl(dft_ot_functional *otf, cplx<T> *psib, const cplx<T> *vb, wf *gwf) {
  dft_ot_device *d = otf->device;
  hipStream_t s = lib_stream();
  const bool BF = (otf->model & DFT_OT_BACKFLOW) != 0;
  AdjWork<T> w;
  if (!w.alloc(d->n, (int64_t)d->nx * d->ny * d->PP, BF)) {
    fprintf(stderr, "libdft: dft_ot_potential_vjp_psi: out of device memory\n");
    w.release();
    return false;
  }
  const cplx<T> *psi = (const cplx<T> *)gwf->grid->dev;
  Well none{0.0, 1.0, 0.0, 0.0, 0.0, d->step, d->nx, d->ny, d->nz};
  adj_potential<T>(otf, psi, gwf->mass, none, 0.0, w);
  hipLaunchKernelGGL(k_split<T>, dim3(grid_for(d->n)), dim3(kBlock), 0, s, w.r[AVBAR], BF ? w.b[BVBI] : nullptr, vb,
                     d->n);
  DFT_LAUNCH_CHECK();
  DFT_HIP_CHECK(hipMemsetAsync(psib, 0, sizeof(cplx<T>) * d->n, s));
  vjp_ot<T>(otf, w.r[AVBAR], w.r[ARHO], psib, psi, nullptr, w);
  if (BF) vjp_bf<T>(otf, w.r[AVBAR], w.b[BVBI], w.r[ARHO], psib, psi, gwf->mass, w);
  DFT_HIP_CHECK(hipStreamSynchronize(s));
  w.release();
  return true;
}

}  // namespace

// ================================================================== C ABI ==
extern "C" {

// src/ot.c:1002-1271: the T = 0 parameter row (1026-1036), the rows of
// DFT_OT_T0MK (1038-1048: the thermal parametrisation at 0 K, which adds no
// ideal-gas term) and of T400MK ... T3000MK (1050-1216), and the common block
// (1234-1270).
void dft_ot_init_params(dft_ot_functional *otf, INT model) {
  const double a3 = GRID_AUTOANG * GRID_AUTOANG * GRID_AUTOANG;
  otf->beta = otf->rhom = otf->C = 0.0;
  otf->b = -718.99;
  otf->c2 = -2.411857E4;
  otf->c2_exp = 2;
  otf->c3 = 1.858496E6;
  otf->c3_exp = 3;
  otf->c4 = 0.0;
  otf->temp = 0.0;
  otf->rho0 = 0.0218360;
  otf->lj_params.h = 2.1903;
  // rows applied in the reference's order, each whose bit is set (the last one wins)
  static const struct {
    INT bit;
    double b, c2, c3, c4, temp, rho0, h;
  } rows[] = {
      {DFT_OT_T0MK, -719.2435, -24258.88, 1865257.0, 0.0, 0.0, 0.0218354, 2.19035},  // src/ot.c:1038-1048
      {32, -714.2174, -24566.29, 1873203.0, 0.98004, 0.4, 0.0218351, 2.18982},      // 1050-1060
      {64, -705.1319, -25124.17, 1887707.0, 0.99915, 0.6, 0.0218346, 2.18887},
      {128, -690.4745, -26027.12, 1911283.0, 0.99548, 0.8, 0.0218331, 2.18735},
      {256, -646.5135, -28582.81, 1973737.0, 0.99666, 1.2, 0.0218298, 2.18287},
      {512, -625.8123, -29434.03, 1984068.0, 0.99829, 1.4, 0.0218332, 2.18080},
      {1024, -605.9788, -30025.96, 1980898.0, 1.00087, 1.6, 0.0218453, 2.17885},
      {2048, -593.8289, -29807.56, 1945685.0, 1.00443, 1.8, 0.0218703, 2.17766},
      {4096, -600.8313, -27850.96, 1847407.0, 1.00919, 2.0, 0.0219153, 2.17834},
      {8192, -620.9129, -25418.15, 1747494.0, 1.01156, 2.1, 0.0219500, 2.18032},
      {16384, -619.2016, -25096.68, 1720802.0, 1.01436, 2.2, 0.0219859, 2.18015},
      {32768, -609.0757, -26009.98, 1747943.0, 1.02130, 2.4, 0.0218748, 2.17915},
      {65536, -634.0664, -23790.66, 1670707.0, 1.02770, 2.6, 0.0217135, 2.18162},
      {131072, -663.9942, -21046.37, 1574611.0, 1.03429, 2.8, 0.0215090, 2.18463},
      {262144, -673.6543, -20022.76, 1535887.0, 1.04271, 3.0, 0.0212593, 2.18562},  // 1206-1216
  };
  for (const auto &r : rows)
    if (model & r.bit) {
      otf->b = r.b;
      otf->c2 = r.c2;
      otf->c3 = r.c3;
      otf->c4 = r.c4;
      otf->temp = r.temp;
      otf->rho0 = r.rho0;
      otf->lj_params.h = r.h;
    }
  otf->mu0 = 0.0;
  otf->xi = 1E4 / a3;
  otf->rhobf = 0.033 * a3;
  if (model & (DFT_OT_HD | DFT_OT_HD2)) {  // high-density penalty, src/ot.c:1004-1007
    otf->beta = 40.0 / a3;
    otf->rhom = 0.37 * a3;
    otf->C = 0.1;
  }
  otf->b /= GRID_AUTOK * a3;
  otf->c2 /= GRID_AUTOK * std::pow(GRID_AUTOANG, 3.0 * otf->c2_exp);
  otf->c3 /= GRID_AUTOK * std::pow(GRID_AUTOANG, 3.0 * otf->c3_exp);
  otf->rho0 *= a3;
  otf->lj_params.h /= GRID_AUTOANG;
  otf->lj_params.sigma = 2.556 / GRID_AUTOANG;
  otf->lj_params.epsilon = 10.22 / GRID_AUTOK;
  otf->lj_params.cval = 0.0;
  otf->rho_0s = 0.04 * a3;
  otf->alpha_s = 54.31 / a3;
  otf->l_g = 1.0 / GRID_AUTOANG;
  otf->mass = 4.0026 / GRID_AUTOAMU;
  otf->bf_params.g11 = -19.7544;
  otf->bf_params.g12 = 12.5616 * GRID_AUTOANG * GRID_AUTOANG;
  otf->bf_params.a1 = 1.023 * GRID_AUTOANG * GRID_AUTOANG;
  otf->bf_params.g21 = -0.2395;
  otf->bf_params.g22 = 0.0312 * GRID_AUTOANG * GRID_AUTOANG;
  otf->bf_params.a2 = 0.14912 * GRID_AUTOANG * GRID_AUTOANG;
  otf->div_epsilon = DFT_EPS;
  otf->max_veloc = 999.0 / GRID_AUTOMPS;
  otf->max_bfpot = 10000.0 / GRID_AUTOK;
  otf->c_bfpot = 1.0;
  otf->kc_coef23 = 1.0;
}

// src/ot.c:66-258
dft_ot_functional *dft_ot_alloc(INT model, wf *gwf, INT min_substeps, INT max_substeps) {
  if (!supported_model(model)) {
    fprintf(stderr, "libdft: dft_ot_alloc: model %ld not supported (PLAIN, KC, BACKFLOW, HD, HD2 and T0MK ... T3000MK)\n",
            (long)model);
    return nullptr;
  }
  if (!gwf || !gwf->grid) {
    fprintf(stderr, "libdft: dft_ot_alloc: no wave function\n");
    return nullptr;
  }
  cgrid *g = gwf->grid;
  if (g->nx < 4 || g->ny < 4 || g->nz < 4) {
    fprintf(stderr, "libdft: dft_ot_alloc: 1-D and 2-D grids are not supported\n");
    return nullptr;
  }
  if (min_substeps < 1 || max_substeps < min_substeps) {
    fprintf(stderr, "libdft: dft_ot_alloc: bad substep range\n");
    return nullptr;
  }
  auto *otf = (dft_ot_functional *)calloc(1, sizeof(dft_ot_functional));
  if (!otf) {
    fprintf(stderr, "libdft: Error in dft_ot_alloc(): Could not allocate memory for dft_ot_functional.\n");
    return nullptr;
  }
  otf->model = model;
  fprintf(stderr, "libdft: Grid %ld x %ld x %ld with step %le Bohr (%s).\n", (long)g->nx, (long)g->ny, (long)g->nz,
          g->step, g->prec == GRID_FP32 ? "FP32" : "FP64");
  fprintf(stderr, "libdft: Functional: %ld.\n", (long)model);
  dft_ot_init_params(otf, model);

  auto *d = new dft_ot_device();
  memset(d, 0, sizeof(*d));
  otf->device = d;
  d->prec = g->prec;
  d->nx = (int)g->nx;
  d->ny = (int)g->ny;
  d->nz = (int)g->nz;
  d->P = d->nz / 2 + 1;
  d->PP = half_pitch(d->nz);
  d->n = (int64_t)d->nx * d->ny * d->nz;
  d->step = g->step;
  d->slot_elems = (int64_t)d->nx * d->ny * 2 * d->PP;
  const size_t es = g->prec == GRID_FP32 ? sizeof(float) : sizeof(double);
  const int64_t nh = (int64_t)d->nx * d->ny * d->PP;
  bool ok = true;
  for (int i = 0; i < NSLOT && ok; ++i) ok = (d->slot[i] = dmalloc(es * d->slot_elems)) != nullptr;
  for (int i = 0; i < 4 && ok; ++i) ok = (d->kern[i] = dmalloc(es * nh)) != nullptr;
  // the step's potential and the predictor's potential (complex fields), and
  // the x-middle programs of the fused schedule
  if (ok) ok = (d->phase = dmalloc(2 * es * d->n)) != nullptr;
  if (ok) ok = (d->vsave = dmalloc(2 * es * d->n)) != nullptr;
  if (ok)
    ok = (d->norm_part = (double *)dmalloc(sizeof(double) * (strang_partials(d->nx, d->ny, d->nz, d->prec) + 1))) !=
         nullptr;
  if (ok) ok = (d->mid_dev = (fused::MidProg *)dmalloc(3 * sizeof(fused::MidProg))) != nullptr;
  if (ok && has_ig(otf)) {
    double tab[3 * ig::kTerms];
    ig::fill_table(tab);
    d->ig_g1 = ig::series(1.0, tab + ig::kTerms);
    ok = (d->igtab = (double *)dmalloc(sizeof(tab))) != nullptr;
    if (ok) DFT_HIP_CHECK(hipMemcpy(d->igtab, tab, sizeof(tab), hipMemcpyHostToDevice));
  }
  if (ok) {
    if (d->prec == GRID_FP32) {
      auto *p = new Plan3D<float>();
      p->init(d->nx, d->ny, d->nz, d->step);
      d->plan = p;
      ok = build_kernels<float>(otf, (int)min_substeps, (int)max_substeps);
    } else {
      auto *p = new Plan3D<double>();
      p->init(d->nx, d->ny, d->nz, d->step);
      d->plan = p;
      ok = build_kernels<double>(otf, (int)min_substeps, (int)max_substeps);
    }
  }
  if (!ok) {
    fprintf(stderr, "libdft: Error in dft_ot_alloc(): device allocation failed.\n");
    free_device(d);
    free(otf);
    return nullptr;
  }
  otf->lennard_jones = spectral_handle(otf, K_LJ, "OT Lennard-Jones");
  otf->spherical_avg = spectral_handle(otf, K_SA, "OT Sph. average");
  if (model & DFT_OT_KC) otf->gaussian_tf = spectral_handle(otf, K_F, "OT KC Gauss TF");
  if (model & DFT_OT_BACKFLOW) otf->backflow_pot = spectral_handle(otf, K_U, "OT Backflow");
  // x/y/z derivatives of F are applied in k space (i sin(k h)/h F^), not stored
  otf->gaussian_x_tf = otf->gaussian_y_tf = otf->gaussian_z_tf = nullptr;

  int save = grid_get_precision();
  grid_set_precision(d->prec);
  char *nm = (char *)"OT Density";
  otf->density = rgrid_alloc(g->nx, g->ny, g->nz, g->step, RGRID_PERIODIC_BOUNDARY, nullptr, nm);
  ok = otf->density != nullptr;
  rgrid **ws[9] = {&otf->workspace1, &otf->workspace2, &otf->workspace3, &otf->workspace4, &otf->workspace5,
                   &otf->workspace6, &otf->workspace7, &otf->workspace8, &otf->workspace9};
  int nws = 3;
  if (model & (DFT_OT_KC | DFT_OT_BACKFLOW)) nws = 6;
  if (model & DFT_OT_BACKFLOW) nws = 9;
  for (int i = 0; i < nws; ++i) {
    char id[32];
    snprintf(id, sizeof(id), "OT Workspace %d", i + 1);
    *ws[i] = rgrid_alloc(g->nx, g->ny, g->nz, g->step, RGRID_PERIODIC_BOUNDARY, nullptr, id);
    ok = ok && *ws[i] != nullptr;
  }
  grid_set_precision(save);
  if (!ok) {
    fprintf(stderr, "libdft: Error in dft_ot_alloc(): device allocation failed.\n");
    dft_ot_free(otf);
    return nullptr;
  }
  return otf;
}

// src/ot.c:268-290
void dft_ot_free(dft_ot_functional *otf) {
  if (!otf) return;
  rgrid **ws[10] = {&otf->workspace1, &otf->workspace2, &otf->workspace3, &otf->workspace4, &otf->workspace5,
                    &otf->workspace6, &otf->workspace7, &otf->workspace8, &otf->workspace9, &otf->density};
  for (auto w : ws)
    if (*w) rgrid_free(*w);
  for (rgrid *h : {otf->lennard_jones, otf->spherical_avg, otf->gaussian_tf, otf->backflow_pot}) free(h);
  free_device(otf->device);
  free(otf);
}

void dft_ot_set_kc_mode(dft_ot_functional *otf, int exact) { otf->kc_coef23 = exact ? 1.0 / otf->rho_0s : 1.0; }

// src/ot.c:314-408
void dft_ot_potential(dft_ot_functional *otf, cgrid *potential, wf *gwf) {
  if (!check_wf(otf, gwf, "dft_ot_potential") || !check_c(otf, potential, "dft_ot_potential")) return;
  DFT_BY_PREC(otf->device->prec, fused_impl<T>(otf, (const cplx<T> *)gwf->grid->dev, gwf->mass,
                                               (T *)otf->density->dev, (cplx<T> *)potential->dev, nullptr));
}

// One split-operator step with the OT potential of the wave function itself:
// the result equals cgrid_zero(pot); [grid_real_to_complex_re(pot, vext)];
// dft_ot_potential(otf, pot, gwf); cgrid_add(pot, -mu); grid_wf_propagate(gwf, pot, time)
// (single-step driver pattern, doc/libdft-manual.tex:496-504).
void dft_ot_propagate(dft_ot_functional *otf, wf *gwf, rgrid *vext, REAL mu, REAL_COMPLEX time) {
  if (!check_wf(otf, gwf, "dft_ot_propagate") || (vext && !check_r(otf, vext, "dft_ot_propagate"))) return;
  dft_ot_device *d = otf->device;
  StepIn st{vext ? vext->dev : nullptr, mu, nullptr, nullptr, d->phase, gwf->grid->dev, gwf->grid->dev,
            __real__ time, __imag__ time};
  const double k0[3] = {gwf->grid->kx0, gwf->grid->ky0, gwf->grid->kz0};
  DFT_BY_PREC(d->prec, {
    fused_impl<T>(otf, (const cplx<T> *)gwf->grid->dev, gwf->mass, (T *)otf->density->dev, nullptr, &st);
    strang_tail<T>(plan_of<T>(d), (cplx<T> *)gwf->grid->dev, (const cplx<T> *)d->phase, __real__ time,
                   __imag__ time, gwf->mass, lib_stream(), k0);
  });
}

// Predict-correct step as in examples/imag_time/imag_time.c:231-239:
//   V_n = V_OT[psi_n] + V_ext - mu;  psi_p = Step(psi_n; V_n)
//   V_p = V_OT[psi_p] + V_ext - mu;  psi_n+1 = Step(psi_n; (V_n + V_p)/2)
// gwfp receives psi_p.
void dft_ot_propagate_pc(dft_ot_functional *otf, wf *gwf, wf *gwfp, rgrid *vext, REAL mu, REAL_COMPLEX time) {
  if (!check_wf(otf, gwf, "dft_ot_propagate_pc") || !check_wf(otf, gwfp, "dft_ot_propagate_pc") ||
      (vext && !check_r(otf, vext, "dft_ot_propagate_pc")))
    return;
  dft_ot_device *d = otf->device;
  const void *ve = vext ? vext->dev : nullptr;
  StepIn pred{ve, mu, nullptr, d->vsave, d->phase, gwf->grid->dev, gwfp->grid->dev, __real__ time, __imag__ time};
  StepIn corr{ve, mu, d->vsave, nullptr, d->phase, gwf->grid->dev, gwf->grid->dev, __real__ time, __imag__ time};
  const double k0[3] = {gwf->grid->kx0, gwf->grid->ky0, gwf->grid->kz0};
  DFT_BY_PREC(d->prec, {
    fused_impl<T>(otf, (const cplx<T> *)gwf->grid->dev, gwf->mass, (T *)otf->density->dev, nullptr, &pred);
    strang_tail<T>(plan_of<T>(d), (cplx<T> *)gwfp->grid->dev, (const cplx<T> *)d->phase, __real__ time,
                   __imag__ time, gwf->mass, lib_stream(), k0);
    fused_impl<T>(otf, (const cplx<T> *)gwfp->grid->dev, gwf->mass, (T *)otf->density->dev, nullptr, &corr);
    strang_tail<T>(plan_of<T>(d), (cplx<T> *)gwf->grid->dev, (const cplx<T> *)d->phase, __real__ time,
                   __imag__ time, gwf->mass, lib_stream(), k0);
  });
}

// dft_ot_propagate / dft_ot_propagate_pc followed by grid_wf_normalize
// (examples/imag_time_rot/imag_time.c:198-208), the normalisation fused into
// the step's closing z pass.
static int propagate_normalize(dft_ot_functional *otf, wf *gwf, wf *gwfp, rgrid *vext, REAL mu, REAL_COMPLEX time,
                               const char *who) {
  if (!check_wf(otf, gwf, who) || (gwfp && !check_wf(otf, gwfp, who)) || (vext && !check_r(otf, vext, who)))
    return -1;
  if (!(gwf->norm > 0.0)) {
    fprintf(stderr, "libdft: %s: the wave function's norm target is %g\n", who, (double)gwf->norm);
    return -1;
  }
  dft_ot_device *d = otf->device;
  const void *ve = vext ? vext->dev : nullptr;
  const double k0[3] = {gwf->grid->kx0, gwf->grid->ky0, gwf->grid->kz0};
  const Renorm rn{d->norm_part, d->step * d->step * d->step, gwf->norm};
  DFT_BY_PREC(d->prec, {
    if (gwfp) {
      StepIn pred{ve, mu, nullptr, d->vsave, d->phase, gwf->grid->dev, gwfp->grid->dev, __real__ time, __imag__ time};
      StepIn corr{ve, mu, d->vsave, nullptr, d->phase, gwf->grid->dev, gwf->grid->dev, __real__ time, __imag__ time};
      fused_impl<T>(otf, (const cplx<T> *)gwf->grid->dev, gwf->mass, (T *)otf->density->dev, nullptr, &pred);
      strang_tail<T>(plan_of<T>(d), (cplx<T> *)gwfp->grid->dev, (const cplx<T> *)d->phase, __real__ time,
                     __imag__ time, gwf->mass, lib_stream(), k0);
      fused_impl<T>(otf, (const cplx<T> *)gwfp->grid->dev, gwf->mass, (T *)otf->density->dev, nullptr, &corr);
    } else {
      StepIn st{ve, mu, nullptr, nullptr, d->phase, gwf->grid->dev, gwf->grid->dev, __real__ time, __imag__ time};
      fused_impl<T>(otf, (const cplx<T> *)gwf->grid->dev, gwf->mass, (T *)otf->density->dev, nullptr, &st);
    }
    strang_tail<T>(plan_of<T>(d), (cplx<T> *)gwf->grid->dev, (const cplx<T> *)d->phase, __real__ time,
                   __imag__ time, gwf->mass, lib_stream(), k0, &rn);
  });
  return 0;
}

int dft_ot_propagate_normalize(dft_ot_functional *otf, wf *gwf, rgrid *vext, REAL mu, REAL_COMPLEX time) {
  return propagate_normalize(otf, gwf, nullptr, vext, mu, time, "dft_ot_propagate_normalize");
}

int dft_ot_propagate_pc_normalize(dft_ot_functional *otf, wf *gwf, wf *gwfp, rgrid *vext, REAL mu,
                                  REAL_COMPLEX time) {
  if (!gwfp) {
    fprintf(stderr, "libdft: dft_ot_propagate_pc_normalize: no predictor wave function\n");
    return -1;
  }
  return propagate_normalize(otf, gwf, gwfp, vext, mu, time, "dft_ot_propagate_pc_normalize");
}

// src/ot.c:431-510
void dft_ot_potential2(dft_ot_functional *otf, cgrid *potential, rgrid *density, rgrid *vx, rgrid *vy, rgrid *vz) {
  if (!check_c(otf, potential, "dft_ot_potential2") || !check_r(otf, density, "dft_ot_potential2")) return;
  PotParts parts{true, (otf->model & DFT_OT_KC) != 0, (otf->model & DFT_OT_BACKFLOW) != 0};
  if (parts.bf)
    for (rgrid *v : {vx, vy, vz})
      if (!check_r(otf, v, "dft_ot_potential2")) return;
  DFT_BY_PREC(otf->device->prec,
              potential_impl<T>(otf, potential, nullptr, 0.0, (T *)density->dev, parts.bf ? (T *)vx->dev : nullptr,
                                parts.bf ? (T *)vy->dev : nullptr, parts.bf ? (T *)vz->dev : nullptr, parts, true));
}

// src/ot.c:524-531: potential += V_LJ * density.  The reference leaves FFT(density)
// in workspace1 for dft_ot_add_local_correlation_potential; here the workspaces
// are accepted and not written (that function transforms its density itself).
void dft_ot_add_lennard_jones_potential(dft_ot_functional *otf, cgrid *potential, rgrid *density, rgrid *, rgrid *) {
  if (!check_c(otf, potential, "dft_ot_add_lennard_jones_potential") ||
      !check_r(otf, density, "dft_ot_add_lennard_jones_potential"))
    return;
  PotParts parts{false, false, false, 1};
  DFT_BY_PREC(otf->device->prec, potential_impl<T>(otf, potential, nullptr, 0.0, (T *)density->dev, nullptr, nullptr,
                                                   nullptr, parts, false));
}

// src/ot.c:547-592: potential += c2/2 rhobar^2 + c3/3 rhobar^3 + Pi * ((c2 rhobar +
// c3 rhobar^2) rho), rhobar = Pi * rho, for the OT exponents (c2_exp = 2, c3_exp
// = 3).  rho_tf (FFT(rho) in the reference) and the workspaces are accepted and
// not read or written: rhobar is computed from rho.
void dft_ot_add_local_correlation_potential(dft_ot_functional *otf, cgrid *potential, rgrid *rho, rgrid *, rgrid *,
                                            rgrid *) {
  if (!check_c(otf, potential, "dft_ot_add_local_correlation_potential") ||
      !check_r(otf, rho, "dft_ot_add_local_correlation_potential"))
    return;
  PotParts parts{true, false, false, 0};
  DFT_BY_PREC(otf->device->prec, potential_impl<T>(otf, potential, nullptr, 0.0, (T *)rho->dev, nullptr, nullptr,
                                                   nullptr, parts, false));
}

// src/ot.c:840-991 (the workspace arguments are accepted and unused)
void dft_ot_backflow_potential(dft_ot_functional *otf, cgrid *potential, rgrid *density, rgrid *vx, rgrid *vy,
                               rgrid *vz, rgrid *, rgrid *, rgrid *, rgrid *, rgrid *, rgrid *) {
  if (!check_otf(otf, "dft_ot_backflow_potential")) return;
  if (!(otf->model & DFT_OT_BACKFLOW)) {
    fprintf(stderr, "libdft: dft_ot_backflow_potential: functional allocated without DFT_OT_BACKFLOW (no kernel U)\n");
    return;
  }
  if (!check_c(otf, potential, "dft_ot_backflow_potential")) return;
  for (rgrid *v : {density, vx, vy, vz})
    if (!check_r(otf, v, "dft_ot_backflow_potential")) return;
  PotParts parts{false, false, true};
  DFT_BY_PREC(otf->device->prec, potential_impl<T>(otf, potential, nullptr, 0.0, (T *)density->dev, (T *)vx->dev,
                                                   (T *)vy->dev, (T *)vz->dev, parts, true));
}

// src/ot-energy.c:38-110
void dft_ot_energy_density(dft_ot_functional *otf, rgrid *energy_density, wf *gwf) {
  if (!check_wf(otf, gwf, "dft_ot_energy_density") || !check_r(otf, energy_density, "dft_ot_energy_density")) return;
  PotParts parts{true, (otf->model & DFT_OT_KC) != 0, (otf->model & DFT_OT_BACKFLOW) != 0};
  DFT_BY_PREC(otf->device->prec, energy_impl<T>(otf, (T *)energy_density->dev, (const cplx<T> *)gwf->grid->dev,
                                                gwf->mass, (T *)otf->density->dev, false, parts, false));
}

// src/ot-energy.c:123-191 (adds the KC part)
void dft_ot_energy_density_kc(dft_ot_functional *otf, rgrid *energy_density, wf *gwf, rgrid *density) {
  if (!check_r(otf, energy_density, "dft_ot_energy_density_kc") || !check_r(otf, density, "dft_ot_energy_density_kc"))
    return;
  if (!(otf->model & DFT_OT_KC)) {
    fprintf(stderr, "libdft: dft_ot_energy_density_kc: functional allocated without DFT_OT_KC (no kernel F)\n");
    return;
  }
  PotParts parts{false, true, false};
  DFT_BY_PREC(otf->device->prec, energy_impl<T>(otf, (T *)energy_density->dev, nullptr, gwf ? gwf->mass : 0.0,
                                                (T *)density->dev, true, parts, true));
}

// src/ot-energy.c:202-288 (adds the BF part)
void dft_ot_energy_density_bf(dft_ot_functional *otf, rgrid *energy_density, wf *gwf, rgrid *density) {
  if (!check_wf(otf, gwf, "dft_ot_energy_density_bf") || !check_r(otf, energy_density, "dft_ot_energy_density_bf") ||
      !check_r(otf, density, "dft_ot_energy_density_bf"))
    return;
  if (!(otf->model & DFT_OT_BACKFLOW)) {
    fprintf(stderr, "libdft: dft_ot_energy_density_bf: functional allocated without DFT_OT_BACKFLOW (no kernel U)\n");
    return;
  }
  PotParts parts{false, false, true};
  DFT_BY_PREC(otf->device->prec, energy_impl<T>(otf, (T *)energy_density->dev, (const cplx<T> *)gwf->grid->dev,
                                                gwf->mass, (T *)density->dev, true, parts, true));
}

REAL dft_ot_energy(dft_ot_functional *otf, wf *gwf) {
  if (!check_wf(otf, gwf, "dft_ot_energy")) return NAN;
  dft_ot_energy_density(otf, otf->workspace1, gwf);
  return rgrid_integral(otf->workspace1);
}

void dft_ot_energy_gradient(dft_ot_functional *otf, rgrid *grad, rgrid *density) {
  if (!check_r(otf, grad, "dft_ot_energy_gradient") || !check_r(otf, density, "dft_ot_energy_gradient"))
    return;
  DFT_BY_PREC(otf->device->prec, gradient_impl<T>(otf, (T *)grad->dev, (const T *)density->dev));
}

REAL dft_ot_energy_rho(dft_ot_functional *otf, rgrid *density) {
  if (!check_r(otf, density, "dft_ot_energy_rho")) return NAN;
  PotParts parts{true, (otf->model & DFT_OT_KC) != 0, false};
  DFT_BY_PREC(otf->device->prec, energy_impl<T>(otf, (T *)otf->workspace1->dev, nullptr, 1.0, (T *)density->dev,
                                                true, parts, false));
  return rgrid_integral(otf->workspace1);
}


int dft_ot_adjoint_well(dft_ot_functional *otf, wf *gwf, REAL amp, REAL width, REAL x0, REAL y0, REAL z0, REAL mu,
                        REAL_COMPLEX time, INT nsteps, INT every, REAL *energy, REAL *grad) {
  if (!check_wf(otf, gwf, "dft_ot_adjoint_well")) return -1;
  if (nsteps < 1 || every < 1 || width <= 0.0 || !energy || !grad) {
    fprintf(stderr, "libdft: dft_ot_adjoint_well: invalid arguments\n");
    return -1;
  }
  dft_ot_device *d = otf->device;
  Well wl{amp, width, x0, y0, z0, d->step, d->nx, d->ny, d->nz};
  int rc = 0;
  double E = 0.0, g[2] = {0.0, 0.0};
  DFT_BY_PREC(d->prec, rc = adjoint_well_impl<T>(otf, gwf, wl, mu, __real__ time, __imag__ time, nsteps, every, &E, g));
  *energy = E;
  grad[0] = g[0];
  grad[1] = g[1];
  return rc;
}

void dft_set_verbose(int level) { g_verbose = level; }

int dft_ot_adjoint_cavity(dft_ot_functional *otf, wf *gwf, REAL rho0, REAL radius, REAL width, REAL mu,
                          REAL_COMPLEX time, INT nsteps, INT every, REAL *bubble_radius, REAL *grad) {
  if (!check_wf(otf, gwf, "dft_ot_adjoint_cavity")) return -1;
  if (nsteps < 1 || every < 1 || width <= 0.0 || rho0 <= 0.0 || !bubble_radius || !grad) {
    fprintf(stderr, "libdft: dft_ot_adjoint_cavity: invalid arguments\n");
    return -1;
  }
  int rc = 0;
  double J = 0.0, g = 0.0;
  DFT_BY_PREC(otf->device->prec, rc = adjoint_cavity_impl<T>(otf, gwf, rho0, radius, width, mu, __real__ time,
                                                             __imag__ time, nsteps, every, &J, &g));
  *bubble_radius = J;
  *grad = g;
  return rc;
}

int dft_ot_potential_vjp(dft_ot_functional *otf, rgrid *rho_bar, rgrid *v_bar, wf *gwf) {
  if (!check_wf(otf, gwf, "dft_ot_potential_vjp") || !check_r(otf, rho_bar, "dft_ot_potential_vjp") ||
      !check_r(otf, v_bar, "dft_ot_potential_vjp"))
    return -1;
  if (otf->model & DFT_OT_BACKFLOW) {
    fprintf(stderr, "libdft: dft_ot_potential_vjp: V_OT of a backflow model depends on the phase of psi; use dft_ot_potential_vjp_psi\n");
    return -1;
  }
  bool ok = false;
  DFT_BY_PREC(otf->device->prec, ok = potential_vjp_impl<T>(otf, (T *)rho_bar->dev, (const T *)v_bar->dev, gwf));
  return ok ? 0 : -1;
}

int dft_ot_potential_vjp_psi(dft_ot_functional *otf, cgrid *psi_bar, cgrid *v_bar, wf *gwf) {
  if (!check_wf(otf, gwf, "dft_ot_potential_vjp_psi") || !check_c(otf, psi_bar, "dft_ot_potential_vjp_psi") ||
      !check_c(otf, v_bar, "dft_ot_potential_vjp_psi"))
    return -1;
  bool ok = false;
  DFT_BY_PREC(otf->device->prec, ok = potential_vjp_psi_impl<T>(otf, (cplx<T> *)psi_bar->dev,
                                                                (const cplx<T> *)v_bar->dev, gwf));
  return ok ? 0 : -1;
}

}  // extern "C"
