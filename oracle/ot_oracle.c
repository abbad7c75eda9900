/*
 * FP64 CPU oracle for the Orsay-Trento hot path.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, by the smoke check and as the CPU baseline in bench.py;
 * never linked into or called by the product library.
 *
 * Each function cites the reference code it restates.  libgrid (absent from
 * the reference tree) is restated with these fixed semantics:
 *   - rgrid_fft / rgrid_inverse_fft: unnormalised forward (e^{-i}) / backward
 *     (e^{+i}) transforms; rgrid_inverse_fft_norm2 multiplies by h^3/N.
 *   - kernels are sampled with r = 0 at index n/2 and rgrid_fft_convolute
 *     multiplies the spectra by (-1)^(i+j+k).
 *   - rgrid_gradient_*: periodic 2nd-order central difference.
 *   - grid_wf_velocity(psi, eps): v = Im(psi* d psi) / (M (|psi|^2 + eps)).
 *   - rgrid_division_eps(a, b, eps) = a / (b + eps).
 *   - rgrid_adaptive_map: cell average over an ns^3 midpoint sub-grid, ns
 *     doubling from min to max until successive averages differ by < tol.
 *   - grid_wf_propagate (WF_2ND_ORDER_FFT): Strang V/2 - T - V/2.
 * Spectra are held as full complex arrays (mathematically identical to the
 * half spectra of an r2c transform).
 */
#define _POSIX_C_SOURCE 199309L
#include "ot_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Planted mistakes (tests only, default 0): each number breaks one term, so
   that the tests can show the checks that pin the oracle would catch it.
   1 Re V_BF and eps_BF: -2 v.B -> +2 v.B      2 Re V_BF and eps_BF: C dropped
   3 eps_BF: m/4 -> m/2 (potential unchanged)  4 Re V_BF and eps_BF both x2
   5 Im V_BF: (d rho/(rho+eps)) W term negated 6 Im V_BF: 1/2 d W -> d W
   7 W = v A - B -> v A + B                     8 KC term 1: rho_st -> 1
   9 KC term 1 negated                         10 KC term 2 negated
  11 KC term 3 dropped                        12 HD switch: Re V_BF without d(G rho)/d rho
  13 HD switch: gradient half of Im V_BF takes rho, not G rho
  14 HD switch: divergence half of Im V_BF without G
  15 HD switch: eps_BF with rho, not G rho */
static int g_plant = 0;
void orc_set_plant(int which) { g_plant = which; }

double orc_wall_time(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ FFT -- */

static int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

/* In-place iterative radix-2 FFT of one line, sign = -1 forward, +1 backward. */
static void fft_line(double complex *a, int n, const double complex *tw) {
  int lg = ilog2(n);
  for (int i = 0; i < n; i++) {
    int r = 0;
    for (int b = 0; b < lg; b++) if (i & (1 << b)) r |= 1 << (lg - 1 - b);
    if (r > i) { double complex t = a[i]; a[i] = a[r]; a[r] = t; }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, tstep = n / len;
    for (int s = 0; s < n; s += len)
      for (int j = 0; j < half; j++) {
        double complex u = a[s + j], v = a[s + j + half] * tw[j * tstep];
        a[s + j] = u + v;
        a[s + j + half] = u - v;
      }
  }
}

static double complex *twiddles(int n, int sign) {
  double complex *tw = malloc(sizeof(double complex) * (size_t) (n / 2 > 0 ? n / 2 : 1));
  for (int k = 0; k < n / 2; k++) {
    double ang = sign * 2.0 * M_PI * (double) k / (double) n;
    tw[k] = cos(ang) + I * sin(ang);
  }
  return tw;
}

/* 3-D transform along every axis with n > 1; sign = -1 forward, +1 backward (unnormalised). */
void orc_fft3d(double complex *a, int nx, int ny, int nz, int sign) {
  const int LB = 8;
  if (nz > 1) {
    double complex *tw = twiddles(nz, sign);
    #pragma omp parallel for schedule(static)
    for (long l = 0; l < (long) nx * ny; l++) fft_line(a + l * nz, nz, tw);
    free(tw);
  }
  /* y and x: gather LB adjacent k-lines at a time so reads use whole cache lines */
  for (int axis = 1; axis >= 0; axis--) {
    int n = axis ? ny : nx;
    if (n <= 1) continue;
    long stride = axis ? nz : (long) ny * nz;
    long outer = axis ? nx : 1;            /* planes */
    long inner = axis ? 1 : ny;            /* y for the x pass */
    double complex *tw = twiddles(n, sign);
    long nblk = (nz + LB - 1) / LB;
    #pragma omp parallel
    {
      double complex *buf = malloc(sizeof(double complex) * (size_t) n * LB);
      #pragma omp for schedule(static) collapse(2)
      for (long o = 0; o < outer * inner; o++)
        for (long kb = 0; kb < nblk; kb++) {
          long base = axis ? (o * (long) ny * nz) : (o * (long) nz);
          int k0 = (int) (kb * LB), kc = nz - k0 < LB ? nz - k0 : LB;
          for (int e = 0; e < n; e++)
            for (int q = 0; q < kc; q++) buf[q * n + e] = a[base + e * stride + k0 + q];
          for (int q = 0; q < kc; q++) fft_line(buf + q * n, n, tw);
          for (int e = 0; e < n; e++)
            for (int q = 0; q < kc; q++) a[base + e * stride + k0 + q] = buf[q * n + e];
        }
      free(buf);
    }
    free(tw);
  }
}

/* -------------------------------------------------- libgrid primitives -- */

typedef struct { int nx, ny, nz; long n; double step; } geom;

static geom G(const orc_ot *o) { geom g = {o->nx, o->ny, o->nz, (long) o->nx * o->ny * o->nz, o->step}; return g; }

/* rgrid_fft: full complex spectrum of a real grid. */
static void r_fft(const geom *g, const double *in, double complex *out) {
  #pragma omp parallel for
  for (long i = 0; i < g->n; i++) out[i] = in[i];
  orc_fft3d(out, g->nx, g->ny, g->nz, -1);
}

/* rgrid_fft_convolute: dst = a * b * (-1)^(i+j+k) (centred-kernel convention). */
static void r_convolute(const geom *g, double complex *dst, const double complex *a, const double complex *b) {
  #pragma omp parallel for
  for (long l = 0; l < g->n; l++) {
    int k = (int) (l % g->nz), j = (int) ((l / g->nz) % g->ny), i = (int) (l / ((long) g->nz * g->ny));
    double s = ((i + j + k) & 1) ? -1.0 : 1.0;
    dst[l] = a[l] * b[l] * s;
  }
}

/* rgrid_inverse_fft_norm2: out = Re(F^-1[spec]) * h^3 / N.  spec is destroyed. */
static void r_inv_norm2(const geom *g, double complex *spec, double *out) {
  orc_fft3d(spec, g->nx, g->ny, g->nz, +1);
  double s = g->step * g->step * g->step / (double) g->n;
  #pragma omp parallel for
  for (long i = 0; i < g->n; i++) out[i] = creal(spec[i]) * s;
}

/* grid_wf_density */
void orc_wf_density(const double complex *psi, double *rho, int64_t n) {
  #pragma omp parallel for
  for (long i = 0; i < n; i++) rho[i] = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
}

static inline long idx3(int i, int j, int k, int ny, int nz) { return ((long) i * ny + j) * nz + k; }
static inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* rgrid_gradient_{x,y,z}: periodic central difference. */
static void grad_dir(const double *f, double *g, int nx, int ny, int nz, double step, int dir) {
  double inv = 0.5 / step;
  int n[3] = {nx, ny, nz};
  if (n[dir] == 1) { memset(g, 0, sizeof(double) * (size_t) nx * ny * nz); return; }
  #pragma omp parallel for collapse(2)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        int p[3] = {i, j, k}, m[3] = {i, j, k};
        p[dir] = wrap(p[dir] + 1, n[dir]);
        m[dir] = wrap(m[dir] - 1, n[dir]);
        g[idx3(i, j, k, ny, nz)] = inv * (f[idx3(p[0], p[1], p[2], ny, nz)] - f[idx3(m[0], m[1], m[2], ny, nz)]);
      }
}

void orc_gradient(const double *f, double *gx, double *gy, double *gz, int nx, int ny, int nz, double step) {
  if (gx) grad_dir(f, gx, nx, ny, nz, step, 0);
  if (gy) grad_dir(f, gy, nx, ny, nz, step, 1);
  if (gz) grad_dir(f, gz, nx, ny, nz, step, 2);
}

/* grid_wf_velocity: v_d = Im(psi* (psi_{+d} - psi_{-d}) / 2h) / (M (|psi|^2 + eps)). */
void orc_wf_velocity(const double complex *psi, double *vx, double *vy, double *vz,
                     int nx, int ny, int nz, double step, double mass, double eps) {
  int n[3] = {nx, ny, nz};
  double *out[3] = {vx, vy, vz};
  double inv = 0.5 / step;
  for (int dir = 0; dir < 3; dir++) {
    if (!out[dir]) continue;
    #pragma omp parallel for collapse(2)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          long c = idx3(i, j, k, ny, nz);
          if (n[dir] == 1) { out[dir][c] = 0.0; continue; }
          int p[3] = {i, j, k}, m[3] = {i, j, k};
          p[dir] = wrap(p[dir] + 1, n[dir]);
          m[dir] = wrap(m[dir] - 1, n[dir]);
          double complex d = (psi[idx3(p[0], p[1], p[2], ny, nz)] - psi[idx3(m[0], m[1], m[2], ny, nz)]) * inv;
          double complex s = psi[c];
          double num = creal(s) * cimag(d) - cimag(s) * creal(d);
          double rho = creal(s) * creal(s) + cimag(s) * cimag(s);
          out[dir][c] = num / (mass * (rho + eps));
        }
  }
}

/* rgrid_threshold_clear(d, s, ul, ll, uv, lv) */
static void r_clamp(double *f, long n, double ul, double ll, double uv, double lv) {
  #pragma omp parallel for
  for (long i = 0; i < n; i++) {
    if (f[i] > ul) f[i] = uv;
    if (f[i] < ll) f[i] = lv;
  }
}

static void add_re(double complex *pot, const double *f, long n) {
  #pragma omp parallel for
  for (long i = 0; i < n; i++) pot[i] += f[i];
}

static void add_im(double complex *pot, const double *f, long n) {
  #pragma omp parallel for
  for (long i = 0; i < n; i++) pot[i] += I * f[i];
}

/* ------------------------------------------------------ parameters (a1) -- */

/* src/ot.c:1002-1271: the T = 0 row (1026-1036) for model < DFT_OT_T0MK, the
   rows of DFT_OT_T0MK (1038-1048; the thermal parametrisation at 0 K, with no
   ideal-gas term) and of the thermal models T400MK ... T3000MK (1050-1216),
   and the common block. */
static void init_params(orc_ot *o, int64_t model) {
  double a3 = ORC_AUTOANG * ORC_AUTOANG * ORC_AUTOANG;
  o->b = -718.99; o->c2 = -2.411857e4; o->c2_exp = 2; o->c3 = 1.858496e6; o->c3_exp = 3;
  o->c4 = 0.0; o->temp = 0.0; o->rho0 = 0.0218360; o->lj_h = 2.1903;
  /* rows applied in the reference's order, each whose bit is set (the last one wins) */
  static const struct { int64_t bit; double b, c2, c3, c4, temp, rho0, h; } rows[] = {
    {16, -719.2435, -24258.88, 1865257.0, 0.0, 0.0, 0.0218354, 2.19035},        /* src/ot.c:1038-1048 */
    {32, -714.2174, -24566.29, 1873203.0, 0.98004, 0.4, 0.0218351, 2.18982},    /* 1050-1060 */
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
    {262144, -673.6543, -20022.76, 1535887.0, 1.04271, 3.0, 0.0212593, 2.18562},   /* 1206-1216 */
  };
  for (size_t i = 0; i < sizeof rows / sizeof rows[0]; i++)
    if (model & rows[i].bit) {
      o->b = rows[i].b; o->c2 = rows[i].c2; o->c2_exp = 2; o->c3 = rows[i].c3; o->c3_exp = 3;
      o->c4 = rows[i].c4; o->temp = rows[i].temp; o->rho0 = rows[i].rho0; o->lj_h = rows[i].h;
    }
  o->bf_xi = 1e4 / a3; o->bf_rhobf = 0.033 * a3;
  if (model & (ORC_OT_HD | ORC_OT_HD2)) {       /* src/ot.c:1004-1007 */
    o->hd_beta = 40.0 / a3; o->hd_rhom = 0.37 * a3; o->hd_C = 0.1;
  }
  o->b /= ORC_AUTOK * a3;
  o->c2 /= ORC_AUTOK * pow(ORC_AUTOANG, 3.0 * o->c2_exp);
  o->c3 /= ORC_AUTOK * pow(ORC_AUTOANG, 3.0 * o->c3_exp);
  o->rho0 *= a3;
  o->lj_h /= ORC_AUTOANG;
  o->sa_h = (model & ORC_OT_HD2) ? o->lj_h * 1.065 : o->lj_h;
  o->lj_sigma = 2.556 / ORC_AUTOANG;
  o->lj_epsilon = 10.22 / ORC_AUTOK;
  o->lj_cval = 0.0;
  o->rho_0s = 0.04 * a3;
  o->alpha_s = 54.31 / a3;
  o->l_g = 1.0 / ORC_AUTOANG;
  o->mass = 4.0026 / ORC_AUTOAMU;
  o->g11 = -19.7544;
  o->g12 = 12.5616 * ORC_AUTOANG * ORC_AUTOANG;
  o->a1 = 1.023 * ORC_AUTOANG * ORC_AUTOANG;
  o->g21 = -0.2395;
  o->g22 = 0.0312 * ORC_AUTOANG * ORC_AUTOANG;
  o->a2 = 0.14912 * ORC_AUTOANG * ORC_AUTOANG;
  o->div_epsilon = ORC_DFT_EPS;
  o->max_veloc = 999.0 / ORC_AUTOMPS;
  o->max_bfpot = 10000.0 / ORC_AUTOK;
  o->c_bfpot = 1.0;
  o->kc_coef23 = 1.0;
}

/* ------------------------------------------------- kernel functions (a2) -- */

/* src/common.c:31-66 (LJ with zero core), 131-138 (spherical average),
   193-200 (gaussian), src/ot-private.h:11-22 (backflow U_J). */
static inline double kfun(const orc_ot *o, int which, double x, double y, double z) {
  double r2 = x * x + y * y + z * z;
  switch (which) {
  case 0: {
    if (r2 <= o->lj_h * o->lj_h) return o->lj_cval;
    double s = o->lj_sigma * o->lj_sigma / r2;
    s = s * s * s;
    return 4.0 * o->lj_epsilon * s * (s - 1.0);
  }
  case 1: {
    double h = o->sa_h;
    return (r2 <= h * h) ? 3.0 / (4.0 * M_PI * h * h * h) : 0.0;
  }
  case 2: {
    double iw = 1.0 / o->l_g;
    double nrm = 0.5 * (2.0 / sqrt(M_PI)) * iw;
    nrm = nrm * nrm * nrm;
    return nrm * exp(-r2 * iw * iw);
  }
  default:
    return (o->g11 + o->g12 * r2) * exp(-o->a1 * r2) + (o->g21 + o->g22 * r2) * exp(-o->a2 * r2);
  }
}

void orc_kernel_map(orc_ot *o, int which, double *out, int min_sub, int max_sub) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double h = o->step, tol = 0.01 / ORC_AUTOK;
  #pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        double x = (double) (i - nx / 2) * h, y = (double) (j - ny / 2) * h, z = (double) (k - nz / 2) * h;
        double prev = 0.0;
        for (int ns = min_sub; ns <= max_sub; ns *= 2) {
          double s = 0.0;
          for (int a = 0; a < ns; a++) {
            double xs = x + (((double) a + 0.5) / (double) ns - 0.5) * h;
            for (int b = 0; b < ns; b++) {
              double ys = y + (((double) b + 0.5) / (double) ns - 0.5) * h;
              for (int c = 0; c < ns; c++) {
                double zs = z + (((double) c + 0.5) / (double) ns - 0.5) * h;
                s += kfun(o, which, xs, ys, zs);
              }
            }
          }
          double cur = s / ((double) ns * (double) ns * (double) ns);
          int done = (ns > min_sub) && fabs(cur - prev) < tol;
          prev = cur;
          if (done) break;
        }
        out[idx3(i, j, k, ny, nz)] = prev;
      }
}

/* ------------------------------------------------------------ alloc (a2) -- */

static double complex *cbuf(long n) { return calloc((size_t) n, sizeof(double complex)); }
static double *rbuf(long n) { return calloc((size_t) n, sizeof(double)); }

/* src/ot.c:66-258 (3-D, T = 0, no DR/GP; HD2's spherical-average radius). */
orc_ot *orc_ot_alloc(int64_t model, int nx, int ny, int nz, double step, int min_sub, int max_sub) {
  orc_ot *o = calloc(1, sizeof(orc_ot));
  if (!o) return NULL;
  o->model = model; o->nx = nx; o->ny = ny; o->nz = nz; o->step = step;
  init_params(o, model);
  geom g = G(o);
  double *tmp = rbuf(g.n);
  double h3 = step * step * step;

  o->lj_tf = cbuf(g.n);
  orc_kernel_map(o, 0, tmp, min_sub, max_sub);
  r_fft(&g, tmp, o->lj_tf);
  double s = o->b / (h3 * creal(o->lj_tf[0]));
  for (long i = 0; i < g.n; i++) o->lj_tf[i] *= s;

  o->sa_tf = cbuf(g.n);
  orc_kernel_map(o, 1, tmp, min_sub, max_sub);
  r_fft(&g, tmp, o->sa_tf);
  s = 1.0 / (h3 * creal(o->sa_tf[0]));
  for (long i = 0; i < g.n; i++) o->sa_tf[i] *= s;

  if (model & ORC_OT_KC) {
    double *gx = rbuf(g.n), *gy = rbuf(g.n), *gz = rbuf(g.n);
    o->g_tf = cbuf(g.n); o->gx_tf = cbuf(g.n); o->gy_tf = cbuf(g.n); o->gz_tf = cbuf(g.n);
    orc_kernel_map(o, 2, tmp, min_sub, max_sub);
    orc_gradient(tmp, gx, gy, gz, nx, ny, nz, step);
    r_fft(&g, gx, o->gx_tf);
    r_fft(&g, gy, o->gy_tf);
    r_fft(&g, gz, o->gz_tf);
    r_fft(&g, tmp, o->g_tf);
    free(gx); free(gy); free(gz);
  }
  if (model & ORC_OT_BACKFLOW) {
    o->bf_tf = cbuf(g.n);
    orc_kernel_map(o, 3, tmp, min_sub, max_sub);
    r_fft(&g, tmp, o->bf_tf);
  }
  free(tmp);
  return o;
}

void orc_ot_free(orc_ot *o) {
  if (!o) return;
  free(o->lj_tf); free(o->sa_tf); free(o->g_tf); free(o->gx_tf); free(o->gy_tf); free(o->gz_tf); free(o->bf_tf);
  free(o);
}

void orc_ot_set_kc_mode(orc_ot *o, int exact) { o->kc_coef23 = exact ? 1.0 / o->rho_0s : 1.0; }

/* -------------------------------------------------------- potential (a3) -- */

/* High-density penalty (src/ot.c:374-378, 814-819; src/ot-energy.c:96-99),
   whose formulas live in libgrid (grid_func4/5) and are restated from
   Ancilotto, Barranco, Caupin, Mayol and Pi, PRB 72, 214522 (2005):
   eps = C rho (1 + tanh(beta (rho - rho_m))), V = d eps / d rho.
   1 + tanh(x) = 2 / (1 + e^{-2x}) and 1 - tanh(x) = 2 / (1 + e^{2x}). */
static inline double hd_s(const orc_ot *o, double r) { return 2.0 / (1.0 + exp(-2.0 * o->hd_beta * (r - o->hd_rhom))); }
static inline double hd_t(const orc_ot *o, double r) { return 2.0 / (1.0 + exp(2.0 * o->hd_beta * (r - o->hd_rhom))); }
static inline double hd_pot(const orc_ot *o, double r) {
  return o->hd_C * hd_s(o, r) * (1.0 + o->hd_beta * r * hd_t(o, r));
}
static inline double hd_eps(const orc_ot *o, double r) { return o->hd_C * r * hd_s(o, r); }
static inline int has_hd(const orc_ot *o) { return (o->model & (ORC_OT_HD | ORC_OT_HD2)) != 0; }

/* Backflow density switch of HD / HD2 with backflow (src/ot.c:843-982,
   src/ot-energy.c:221-227).  The reference passes rho through libgrid's
   grid_func2 (G(rho) rho) wherever the backflow uses the density: in A, B, C,
   the energy weight and the gradient half of Im V_BF; it multiplies Re V_BF
   by grid_func3 (d(G rho)/d rho) and the divergence half of Im V_BF by
   grid_func1 (G).  libgrid is absent, so G is restated from the published
   form (Ancilotto et al., Int. Rev. Phys. Chem. 36, 621 (2017)):
   G = (1 - tanh(xi (rho - rho_bf))) / 2 = 1 / (1 + e^{2x}), x = xi (rho - rho_bf),
   formed from e = e^{-2|x|}; G = 1 to below 1e-40 at liquid densities. */
static inline int has_bfhd(const orc_ot *o) { return has_hd(o) && (o->model & ORC_OT_BACKFLOW); }
static inline double bf_g(const orc_ot *o, double r) {
  double x = o->bf_xi * (r - o->bf_rhobf), e = exp(-2.0 * fabs(x));
  return x > 0.0 ? e / (1.0 + e) : 1.0 / (1.0 + e);
}
static inline double bf_dg(const orc_ot *o, double r) {   /* d(G rho)/d rho = G + rho G' */
  double x = o->bf_xi * (r - o->bf_rhobf), e = exp(-2.0 * fabs(x));
  return bf_g(o, r) - 2.0 * o->bf_xi * r * e / ((1.0 + e) * (1.0 + e));
}
/* Ideal Bose gas of the thermal models, as the reference's src/common.c states it
   (the bulk closed forms call it, src/helium-ot-bulk.c:30-31, 54-55; libgrid's
   grid_func6a/6b, which the potential and energy call with the same mass,
   temperature and c4, src/ot.c:804-808, src/ot-energy.c:101-104, are absent):
     lambda^3 = (2 pi / (M k_B T))^{3/2}                      (common.c:253-260)
     g_s(z) = sum_{k=1}^{256} z^k / k^s                      (common.c:293-306)
     z(x): the golden-section minimum of (x - g_{3/2}(z))^2 on [0, 1] to a
           bracket of 1e-6, or 1 when x >= g_{3/2}(1)       (common.c:400-437)
     eps = c4 k_B T (rho ln z - g_{5/2}(z) / lambda^3)        (common.c:489-497)
     V = c4 k_B T (ln z + x / g_{1/2}(z) - g_{3/2}(z) / g_{1/2}(z)), x = rho lambda^3,
         or -c4 k_B (T / lambda^3) g_{3/2}(1) when x >= g_{3/2}(1)  (common.c:510-530). */
static int has_ig(const orc_ot *o) { return o->model >= ORC_OT_T400MK && o->model < ORC_GP; }
static double ig_series(double z, double s) {
  double v = 0.0, zk = 1.0;
  for (int k = 1; k <= 256; k++) {
    zk *= z;
    v += zk / pow((double) k, s);
  }
  return v;
}
static double ig_lambda3(const orc_ot *o) {
  double l = sqrt(2.0 * M_PI / (o->mass * (o->temp / ORC_AUTOK)));
  return l * l * l;
}
static double ig_z(double x) {
  const double gr = (sqrt(5.0) + 1.0) / 2.0;
  if (x >= ig_series(1.0, 1.5)) return 1.0;
  double a = 0.0, b = 1.0, c = b - (b - a) / gr, d = a + (b - a) / gr;
  while (fabs(c - d) > 1e-6) {
    double tc = x - ig_series(c, 1.5), td = x - ig_series(d, 1.5);
    if (tc * tc < td * td) b = d; else a = c;
    c = b - (b - a) / gr;
    d = a + (b - a) / gr;
  }
  return (b + a) / 2.0;
}
double orc_ig_energy(const orc_ot *o, double rho) {
  double l3 = ig_lambda3(o), z = ig_z(rho * l3);
  return o->c4 * (o->temp / ORC_AUTOK) * (rho * log(z) - ig_series(z, 2.5) / l3);
}
/* the derivative of that energy per volume: the search's z is constant between
   the steps of its branch decisions, so it is c4 k_B T ln z below saturation and
   0 above (unlike the potential, which keeps the reference's saturated constant) */
double orc_ig_dEdRho_energy(const orc_ot *o, double rho) {
  double l3 = ig_lambda3(o), x = rho * l3;
  if (x >= ig_series(1.0, 1.5)) return 0.0;
  return o->c4 * (o->temp / ORC_AUTOK) * log(ig_z(x));
}
double orc_ig_dEdRho(const orc_ot *o, double rho) {
  double l3 = ig_lambda3(o), x = rho * l3, kt = o->temp / ORC_AUTOK, g1 = ig_series(1.0, 1.5);
  if (x >= g1) return -o->c4 * kt / l3 * g1;
  double z = ig_z(x), g12 = ig_series(z, 0.5), g32 = ig_series(z, 1.5);
  return o->c4 * kt * (log(z) + x / g12 - g32 / g12);
}

/* the density the backflow uses: rho, or G(rho) rho with the switch */
static void bf_density(const orc_ot *o, const double *rho, double *out, long n) {
  #pragma omp parallel for
  for (long i = 0; i < n; i++) out[i] = has_bfhd(o) ? bf_g(o, rho[i]) * rho[i] : rho[i];
}

/* src/ot.c:524-532 and 547-586: LJ + local correlation.  rho_tf is FFT(rho). */
static void add_lj_local(orc_ot *o, const geom *g, double complex *pot, const double *rho,
                         const double complex *rho_tf, double complex *cw, double *w1, double *w2) {
  long n = g->n;
  r_convolute(g, cw, rho_tf, o->lj_tf);
  r_inv_norm2(g, cw, w2);
  add_re(pot, w2, n);

  r_convolute(g, cw, rho_tf, o->sa_tf);
  r_inv_norm2(g, cw, w1);                       /* w1 = rho bar */
  #pragma omp parallel for
  for (long i = 0; i < n; i++) w2[i] = w1[i] * w1[i] * (o->c2 / 2.0);
  add_re(pot, w2, n);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) w2[i] = w1[i] * w1[i] * w1[i] * (o->c3 / 3.0);
  add_re(pot, w2, n);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) {
    double t2 = w1[i] * (o->c2 * o->c2_exp / 2.0);
    double t1 = w1[i] * w1[i] * (o->c3 * o->c3_exp / 3.0);
    w2[i] = (t2 + t1) * rho[i];
  }
  r_fft(g, w2, cw);
  r_convolute(g, cw, cw, o->sa_tf);
  r_inv_norm2(g, cw, w2);
  add_re(pot, w2, n);
}

/* src/ot.c:604-797: non-local (kinetic) correlation, x/y/z in turn. */
static void add_kc(orc_ot *o, const geom *g, double complex *pot, const double *rho,
                   const double complex *rho_tf, double complex *cw1, double complex *cw2,
                   double *rho_st, double *w1, double *w2, double *w3, double *w4) {
  long n = g->n;
  double c = o->alpha_s / (2.0 * o->mass), k23 = o->kc_coef23;
  r_convolute(g, cw1, o->g_tf, rho_tf);
  r_inv_norm2(g, cw1, rho_st);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) rho_st[i] = rho_st[i] * (-1.0 / o->rho_0s) + 1.0;

  const double complex *gd_tf[3] = {o->gx_tf, o->gy_tf, o->gz_tf};
  int dims[3] = {g->nx, g->ny, g->nz};
  for (int d = 0; d < 3; d++) {
    if (d < 2 && dims[d] <= 1) continue;
    grad_dir(rho, w1, g->nx, g->ny, g->nz, g->step, d);      /* w1 = d rho */
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w2[i] = w1[i] * rho_st[i];  /* G */
    r_fft(g, w2, cw2);                                       /* cw2 = FFT(G) */
    /* term 1: c rho_st (dF * G) */
    r_convolute(g, cw1, gd_tf[d], cw2);
    r_inv_norm2(g, cw1, w3);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w3[i] = w3[i] * (g_plant == 8 ? 1.0 : rho_st[i]) * (g_plant == 9 ? -c : c);
    add_re(pot, w3, n);
    /* term 2: c F * (d rho J), J = F * G */
    r_convolute(g, cw1, o->g_tf, cw2);
    r_inv_norm2(g, cw1, w3);                                 /* w3 = J */
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w4[i] = w1[i] * w3[i];
    r_fft(g, w4, cw1);
    r_convolute(g, cw1, o->g_tf, cw1);
    r_inv_norm2(g, cw1, w4);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w4[i] = w4[i] * (g_plant == 10 ? -c * k23 : c * k23);
    add_re(pot, w4, n);
    /* term 3: -c J (dF * rho) */
    r_convolute(g, cw1, gd_tf[d], rho_tf);
    r_inv_norm2(g, cw1, w2);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w2[i] = w2[i] * w3[i] * (g_plant == 11 ? 0.0 : -c * k23);
    add_re(pot, w2, n);
  }
}

/* src/ot.c:840-991: adds Re and Im backflow potential, with the density
   switch of HD / HD2 (rt = G rho; rt = rho without it).  vx, vy, vz are
   overwritten with v_d A - B_d as in the reference. */
static void add_bf(orc_ot *o, const geom *g, double complex *pot, const double *rho,
                   double *vx, double *vy, double *vz, double complex *cw,
                   double *A, double *C, double *Bx, double *By, double *Bz, double *w6) {
  long n = g->n;
  const int hb = has_bfhd(o);
  double *rt = rbuf(n);
  bf_density(o, rho, rt, n);
  r_fft(g, rt, cw);
  r_convolute(g, cw, cw, o->bf_tf);
  r_inv_norm2(g, cw, A);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) C[i] = (vz[i] * vz[i] + vy[i] * vy[i] + vx[i] * vx[i]) * rt[i];
  r_fft(g, C, cw); r_convolute(g, cw, cw, o->bf_tf); r_inv_norm2(g, cw, C);
  double *B[3] = {Bx, By, Bz}, *v[3] = {vx, vy, vz};
  for (int d = 0; d < 3; d++) {
    #pragma omp parallel for
    for (long i = 0; i < n; i++) B[d][i] = v[d][i] * rt[i];
    r_fft(g, B[d], cw); r_convolute(g, cw, cw, o->bf_tf); r_inv_norm2(g, cw, B[d]);
  }
  /* real part */
  #pragma omp parallel for
  for (long i = 0; i < n; i++) {
    double t = vz[i] * vz[i] + vy[i] * vy[i] + vx[i] * vx[i];
    const double mb = g_plant == 1 ? 2.0 : -2.0;
    t = t * A[i];
    t += mb * vx[i] * Bx[i];
    t += mb * vy[i] * By[i];
    t += mb * vz[i] * Bz[i];
    if (g_plant != 2) t = t + C[i];
    if (hb && g_plant != 12) t = t * bf_dg(o, rho[i]);
    w6[i] = t * (g_plant == 4 ? -o->mass : -0.5 * o->mass);
  }
  if (o->max_bfpot < 10000.0 / ORC_AUTOK) r_clamp(w6, n, o->max_bfpot, -o->max_bfpot, o->max_bfpot, -o->max_bfpot);
  if (o->c_bfpot != 1.0) for (long i = 0; i < n; i++) w6[i] *= o->c_bfpot;
  add_re(pot, w6, n);
  /* imaginary part */
  memset(w6, 0, sizeof(double) * (size_t) n);
  for (int d = 0; d < 3; d++) {
    #pragma omp parallel for
    for (long i = 0; i < n; i++) v[d][i] = v[d][i] * A[i] + (g_plant == 7 ? B[d][i] : -B[d][i]);
  }
  for (int d = 0; d < 3; d++) {
    grad_dir(g_plant == 13 ? rho : rt, C, g->nx, g->ny, g->nz, g->step, d);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w6[i] += (g_plant == 5 ? -0.5 : 0.5) * (C[i] / (rho[i] + o->div_epsilon)) * v[d][i];
  }
  for (int d = 0; d < 3; d++) {
    grad_dir(v[d], C, g->nx, g->ny, g->nz, g->step, d);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) {
      double dw = hb && g_plant != 14 ? C[i] * bf_g(o, rho[i]) : C[i];
      w6[i] += (g_plant == 6 ? 1.0 : 0.5) * dw;
    }
  }
  if (o->max_bfpot < 10000.0 / ORC_AUTOK) r_clamp(w6, n, o->max_bfpot, -o->max_bfpot, o->max_bfpot, -o->max_bfpot);
  if (o->c_bfpot != 1.0) for (long i = 0; i < n; i++) w6[i] *= o->c_bfpot;
  add_im(pot, w6, n);
  free(rt);
}

/* src/ot.c:431-510 (dft_ot_potential2) with the src/ot.c:314-408 body shared. */
static void potential_body(orc_ot *o, double complex *pot, const double *rho, double *vx, double *vy, double *vz) {
  geom g = G(o);
  long n = g.n;
  double complex *rho_tf = cbuf(n), *cw1 = cbuf(n), *cw2 = cbuf(n);
  double *w[8];
  for (int i = 0; i < 8; i++) w[i] = rbuf(n);
  r_fft(&g, rho, rho_tf);
  add_lj_local(o, &g, pot, rho, rho_tf, cw1, w[0], w[1]);
  if (o->model & ORC_OT_KC) add_kc(o, &g, pot, rho, rho_tf, cw1, cw2, w[0], w[1], w[2], w[3], w[4]);
  if (has_hd(o))                               /* src/ot.c:374-378 */
    for (long i = 0; i < n; i++) pot[i] += hd_pot(o, rho[i]);
  if (o->model & ORC_OT_BACKFLOW) {
    if (o->max_veloc < 999.0 / ORC_AUTOMPS) {
      r_clamp(vx, n, o->max_veloc, -o->max_veloc, o->max_veloc, -o->max_veloc);
      r_clamp(vy, n, o->max_veloc, -o->max_veloc, o->max_veloc, -o->max_veloc);
      r_clamp(vz, n, o->max_veloc, -o->max_veloc, o->max_veloc, -o->max_veloc);
    }
    add_bf(o, &g, pot, rho, vx, vy, vz, cw1, w[0], w[1], w[2], w[3], w[4], w[5]);
  }
  if (has_ig(o))                               /* src/ot.c:402-407 */
    #pragma omp parallel for
    for (long i = 0; i < n; i++) pot[i] += orc_ig_dEdRho(o, rho[i]);
  free(rho_tf); free(cw1); free(cw2);
  for (int i = 0; i < 8; i++) free(w[i]);
}

void orc_ot_potential(orc_ot *o, double complex *pot, const double complex *psi, double wf_mass, double *rho_out) {
  geom g = G(o);
  double *rho = rbuf(g.n), *vx = NULL, *vy = NULL, *vz = NULL;
  orc_wf_density(psi, rho, g.n);
  if (o->model & ORC_OT_BACKFLOW) {
    vx = rbuf(g.n); vy = rbuf(g.n); vz = rbuf(g.n);
    orc_wf_velocity(psi, vx, vy, vz, g.nx, g.ny, g.nz, g.step, wf_mass, o->div_epsilon);
  }
  potential_body(o, pot, rho, vx, vy, vz);
  if (rho_out) memcpy(rho_out, rho, sizeof(double) * (size_t) g.n);
  free(rho); free(vx); free(vy); free(vz);
}

void orc_ot_potential2(orc_ot *o, double complex *pot, const double *rho, double *vx, double *vy, double *vz) {
  potential_body(o, pot, rho, vx, vy, vz);
}

/* ------------------------------------------------- energy density (a10) -- */

/* src/ot-energy.c:38-110, 123-191, 202-288 (HD penalty, no thermal), with the
   density rho and, for backflow, the velocity vel given. */
static void energy_body(orc_ot *o, double *eps, const double *rho, double *const *vel) {
  geom g = G(o);
  long n = g.n;
  double *w1 = rbuf(n), *w2 = rbuf(n);
  double complex *cw1 = cbuf(n), *cw2 = cbuf(n);
  memset(eps, 0, sizeof(double) * (size_t) n);
  r_fft(&g, rho, cw1);
  r_convolute(&g, cw2, cw1, o->lj_tf);
  r_inv_norm2(&g, cw2, w2);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) eps[i] += 0.5 * rho[i] * w2[i];
  r_convolute(&g, cw1, cw1, o->sa_tf);
  r_inv_norm2(&g, cw1, w1);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) eps[i] += (o->c2 / 2.0) * (w1[i] * w1[i] * rho[i]);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) eps[i] += (o->c3 / 3.0) * (w1[i] * w1[i] * w1[i] * rho[i]);
  if (has_hd(o))                               /* src/ot-energy.c:96-99 */
    for (long i = 0; i < n; i++) eps[i] += hd_eps(o, rho[i]);
  if (has_ig(o))                               /* src/ot-energy.c:101-104 */
    #pragma omp parallel for
    for (long i = 0; i < n; i++) eps[i] += orc_ig_energy(o, rho[i]);

  if (o->model & ORC_OT_KC) {
    double *gr[3], *w[3];
    for (int d = 0; d < 3; d++) { gr[d] = rbuf(n); w[d] = rbuf(n); }
    r_fft(&g, rho, cw2);
    r_convolute(&g, cw1, cw2, o->g_tf);
    r_inv_norm2(&g, cw1, w1);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w1[i] = w1[i] * (-1.0 / o->rho_0s) + 1.0;
    orc_gradient(rho, gr[0], gr[1], gr[2], g.nx, g.ny, g.nz, g.step);
    for (int d = 0; d < 3; d++) {
      #pragma omp parallel for
      for (long i = 0; i < n; i++) w[d][i] = gr[d][i] * w1[i];
      r_fft(&g, w[d], cw1);
      r_convolute(&g, cw1, cw1, o->g_tf);
      r_inv_norm2(&g, cw1, w[d]);
      #pragma omp parallel for
      for (long i = 0; i < n; i++) w[d][i] = w[d][i] * gr[d][i] * w1[i];
    }
    double s = -o->alpha_s / (4.0 * o->mass);
    for (int d = 0; d < 3; d++) {
      #pragma omp parallel for
      for (long i = 0; i < n; i++) eps[i] += s * w[d][i];
    }
    for (int d = 0; d < 3; d++) { free(gr[d]); free(w[d]); }
  }

  if (o->model & ORC_OT_BACKFLOW) {
    /* rho is G rho throughout with the HD / HD2 switch (src/ot-energy.c:221-227) */
    double *const *v = vel, *v2 = rbuf(n), *rt = rbuf(n);
    bf_density(o, rho, rt, n);
    if (g_plant == 15) memcpy(rt, rho, sizeof(double) * (size_t) n);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) v2[i] = (v[0][i] * v[0][i] + v[1][i] * v[1][i]) + v[2][i] * v[2][i];
    double mq = o->c_bfpot * o->mass / (g_plant == 3 || g_plant == 4 ? 2.0 : 4.0), mh = o->c_bfpot * o->mass / 2.0;
    if (g_plant == 1) mh = -mh;
    if (g_plant == 4) mh *= 2.0;
    r_fft(&g, rt, cw1);
    r_convolute(&g, cw1, o->bf_tf, cw1);
    r_inv_norm2(&g, cw1, w1);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) eps[i] += -mq * (w1[i] * v2[i] * rt[i]);
    for (int d = 0; d < 3; d++) {
      #pragma omp parallel for
      for (long i = 0; i < n; i++) w2[i] = rt[i] * v[d][i];
      r_fft(&g, w2, cw1);
      r_convolute(&g, cw1, o->bf_tf, cw1);
      r_inv_norm2(&g, cw1, w1);
      #pragma omp parallel for
      for (long i = 0; i < n; i++) eps[i] += mh * (w1[i] * rt[i] * v[d][i]);
    }
    #pragma omp parallel for
    for (long i = 0; i < n; i++) w2[i] = rt[i] * v2[i];
    r_fft(&g, w2, cw1);
    r_convolute(&g, cw1, o->bf_tf, cw1);
    r_inv_norm2(&g, cw1, w1);
    #pragma omp parallel for
    for (long i = 0; i < n; i++) eps[i] += g_plant == 2 ? 0.0 : -mq * (w1[i] * rt[i]);
    free(v2); free(rt);
  }
  free(w1); free(w2); free(cw1); free(cw2);
}

void orc_ot_energy_density(orc_ot *o, double *eps, const double complex *psi, double wf_mass) {
  geom g = G(o);
  double *rho = rbuf(g.n), *v[3] = {NULL, NULL, NULL};
  orc_wf_density(psi, rho, g.n);
  if (o->model & ORC_OT_BACKFLOW) {
    for (int d = 0; d < 3; d++) v[d] = rbuf(g.n);
    orc_wf_velocity(psi, v[0], v[1], v[2], g.nx, g.ny, g.nz, g.step, wf_mass, ORC_DFT_EPS);
  }
  energy_body(o, eps, rho, v);
  free(rho);
  for (int d = 0; d < 3; d++) free(v[d]);
}

void orc_ot_energy_density_rv(orc_ot *o, double *eps, const double *rho, const double *vx, const double *vy,
                              const double *vz) {
  double *v[3] = {(double *) vx, (double *) vy, (double *) vz};
  energy_body(o, eps, rho, v);
}

/* ------------------------------------------------- backward of E[rho] -- */

/* Exact discrete functional derivative dE/drho of the density parts of
   src/ot-energy.c (LJ, c2/c3, KC), derived by reverse-mode differentiation of
   the discrete energy (central differences are antisymmetric, the kernels
   symmetric):
     dE/drho = V_LJ*rho + c2/2 rb^2 + c3/3 rb^3 + Pi*[(c2 rb + c3 rb^2) rho]
             + c sum_d D_d(rho_st J_d) + (c/rho_0s) F*(sum_d (D_d rho) J_d),
   c = alpha_s/(2M), J_d = F*((D_d rho) rho_st), rho_st = 1 - F*rho/rho_0s.
   The backflow energy depends on the current and is not part of E[rho]. */
void orc_ot_energy_gradient(orc_ot *o, double *grad, const double *rho) {
  geom g = G(o);
  long n = g.n;
  double complex *rho_tf = cbuf(n), *cw = cbuf(n);
  double *w1 = rbuf(n), *w2 = rbuf(n);
  memset(grad, 0, sizeof(double) * (size_t) n);
  double complex *pot = cbuf(n);
  r_fft(&g, rho, rho_tf);
  add_lj_local(o, &g, pot, rho, rho_tf, cw, w1, w2);
  for (long i = 0; i < n; i++) grad[i] = creal(pot[i]);
  if (has_hd(o))
    for (long i = 0; i < n; i++) grad[i] += hd_pot(o, rho[i]);
  if (o->model & ORC_OT_KC) {
    double c = o->alpha_s / (2.0 * o->mass);
    double *rho_st = rbuf(n), *gd = rbuf(n), *J = rbuf(n), *H = rbuf(n), *M = rbuf(n);
    r_convolute(&g, cw, o->g_tf, rho_tf);
    r_inv_norm2(&g, cw, rho_st);
    for (long i = 0; i < n; i++) rho_st[i] = rho_st[i] * (-1.0 / o->rho_0s) + 1.0;
    for (int d = 0; d < 3; d++) {
      grad_dir(rho, gd, g.nx, g.ny, g.nz, g.step, d);
      for (long i = 0; i < n; i++) w1[i] = gd[i] * rho_st[i];
      r_fft(&g, w1, cw);
      r_convolute(&g, cw, o->g_tf, cw);
      r_inv_norm2(&g, cw, J);
      for (long i = 0; i < n; i++) {
        H[i] += gd[i] * J[i];
        M[i] = rho_st[i] * J[i];
      }
      grad_dir(M, w2, g.nx, g.ny, g.nz, g.step, d);
      for (long i = 0; i < n; i++) grad[i] += c * w2[i];
    }
    r_fft(&g, H, cw);
    r_convolute(&g, cw, o->g_tf, cw);
    r_inv_norm2(&g, cw, w1);
    for (long i = 0; i < n; i++) grad[i] += (c / o->rho_0s) * w1[i];
    free(rho_st); free(gd); free(J); free(H); free(M);
  }
  if (has_ig(o))   /* the ideal gas: c4 k_B T ln z below saturation, 0 above */
    for (long i = 0; i < n; i++) grad[i] += orc_ig_dEdRho_energy(o, rho[i]);
  free(rho_tf); free(cw); free(w1); free(w2); free(pot);
}

/* E[rho] of the density parts (LJ, c2/c3, KC): h^3 sum eps. */
double orc_ot_energy_rho(orc_ot *o, const double *rho) {
  geom g = G(o);
  long n = g.n;
  double complex *psi = cbuf(n);
  double *eps = rbuf(n);
  for (long i = 0; i < n; i++) psi[i] = sqrt(rho[i] > 0 ? rho[i] : 0.0);
  int64_t save = o->model;
  o->model &= ~(int64_t) ORC_OT_BACKFLOW;
  orc_ot_energy_density(o, eps, psi, ORC_HELIUM_MASS);
  o->model = save;
  double e = orc_integral(eps, n, g.step);
  free(psi); free(eps);
  return e;
}

/* ------------------------------------------------- propagation (a11) -- */

static inline double kcomp(int i, int n, double step) {
  double L = (double) n * step;
  return 2.0 * M_PI * (double) (i < n / 2 ? i : i - n) / L;
}

/* grid_wf_propagate, WF_2ND_ORDER_FFT (libgrid; Strang split V/2 - T - V/2). */
void orc_wf_propagate(double complex *psi, const double complex *pot, double complex dt,
                      int nx, int ny, int nz, double step, double mass) {
  long n = (long) nx * ny * nz;
  #pragma omp parallel for
  for (long l = 0; l < n; l++) psi[l] *= cexp(-I * pot[l] * dt * 0.5);
  orc_fft3d(psi, nx, ny, nz, -1);
  double inv = 1.0 / (double) n;
  #pragma omp parallel for
  for (long l = 0; l < n; l++) {
    int k = (int) (l % nz), j = (int) ((l / nz) % ny), i = (int) (l / ((long) nz * ny));
    double kx = nx > 1 ? kcomp(i, nx, step) : 0.0, ky = ny > 1 ? kcomp(j, ny, step) : 0.0, kz = nz > 1 ? kcomp(k, nz, step) : 0.0;
    double k2 = kx * kx + ky * ky + kz * kz;
    psi[l] *= cexp(-I * (k2 / (2.0 * mass)) * dt) * inv;
  }
  orc_fft3d(psi, nx, ny, nz, +1);
  #pragma omp parallel for
  for (long l = 0; l < n; l++) psi[l] *= cexp(-I * pot[l] * dt * 0.5);
}

/* ------------------------------------------------- reductions (a12) -- */

double orc_wf_norm(const double complex *psi, int64_t n, double step) {
  double s = 0.0;
  #pragma omp parallel for reduction(+ : s)
  for (long i = 0; i < n; i++) s += creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
  return s * step * step * step;
}

double orc_integral(const double *f, int64_t n, double step) {
  double s = 0.0;
  #pragma omp parallel for reduction(+ : s)
  for (long i = 0; i < n; i++) s += f[i];
  return s * step * step * step;
}

/* grid_wf_energy(gwf, NULL): h^3/N sum_k (k^2/2M) |psi_k|^2. */
double orc_wf_kinetic_energy(const double complex *psi, int nx, int ny, int nz, double step, double mass) {
  long n = (long) nx * ny * nz;
  double complex *t = cbuf(n);
  memcpy(t, psi, sizeof(double complex) * (size_t) n);
  orc_fft3d(t, nx, ny, nz, -1);
  double s = 0.0;
  #pragma omp parallel for reduction(+ : s)
  for (long l = 0; l < n; l++) {
    int k = (int) (l % nz), j = (int) ((l / nz) % ny), i = (int) (l / ((long) nz * ny));
    double kx = nx > 1 ? kcomp(i, nx, step) : 0.0, ky = ny > 1 ? kcomp(j, ny, step) : 0.0, kz = nz > 1 ? kcomp(k, nz, step) : 0.0;
    double a = cabs(t[l]);
    s += (kx * kx + ky * ky + kz * kz) / (2.0 * mass) * a * a;
  }
  free(t);
  return s * step * step * step / (double) n;
}

/* ------------------------------------------------- drivers (CS2, CS3) -- */

static void set_vext(double complex *pot, const double *vext, long n, int add) {
  #pragma omp parallel for
  for (long i = 0; i < n; i++) {
    double v = vext ? vext[i] : 0.0;
    pot[i] = add ? pot[i] + v : v;
  }
}

/* examples/imag_time/imag_time.c:231-239; predict = Step(psi; V_n), correct = Step(psi; (V_n + V_p)/2). */
void orc_pc_step(orc_ot *o, double complex *psi, const double *vext, double mu0,
                 double complex dt, double wf_mass, double complex *work) {
  geom g = G(o);
  long n = g.n;
  double complex *pot = work, *psip = work + n;
  set_vext(pot, vext, n, 0);
  orc_ot_potential(o, pot, psi, wf_mass, NULL);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) pot[i] -= mu0;
  memcpy(psip, psi, sizeof(double complex) * (size_t) n);
  orc_wf_propagate(psip, pot, dt, g.nx, g.ny, g.nz, g.step, wf_mass);
  set_vext(pot, vext, n, 1);
  orc_ot_potential(o, pot, psip, wf_mass, NULL);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) pot[i] = (pot[i] - mu0) * 0.5;
  orc_wf_propagate(psi, pot, dt, g.nx, g.ny, g.nz, g.step, wf_mass);
}

void orc_single_step(orc_ot *o, double complex *psi, const double *vext, double mu0,
                     double complex dt, double wf_mass, double complex *work) {
  geom g = G(o);
  long n = g.n;
  double complex *pot = work;
  set_vext(pot, vext, n, 0);
  orc_ot_potential(o, pot, psi, wf_mass, NULL);
  #pragma omp parallel for
  for (long i = 0; i < n; i++) pot[i] -= mu0;
  orc_wf_propagate(psi, pot, dt, g.nx, g.ny, g.nz, g.step, wf_mass);
}

/* examples/imag_time/imag_time.c:251-257 plus mu = <psi|H|psi>/N. */
void orc_observables(orc_ot *o, const double complex *psi, const double *vext, double wf_mass, double *out) {
  geom g = G(o);
  long n = g.n;
  double h3 = g.step * g.step * g.step;
  double *eps = rbuf(n), *rho = rbuf(n);
  double complex *pot = cbuf(n);
  orc_ot_energy_density(o, eps, psi, wf_mass);
  orc_ot_potential(o, pot, psi, wf_mass, rho);
  double ekin = orc_wf_kinetic_energy(psi, g.nx, g.ny, g.nz, g.step, wf_mass);
  double eot = orc_integral(eps, n, g.step);
  double eext = 0.0, ev = 0.0;
  #pragma omp parallel for reduction(+ : eext, ev)
  for (long i = 0; i < n; i++) {
    if (vext) eext += rho[i] * vext[i];
    ev += rho[i] * creal(pot[i]);
  }
  eext *= h3; ev *= h3;
  double nrm = orc_wf_norm(psi, n, g.step);
  out[0] = ekin + eot + eext;
  out[1] = nrm;
  out[2] = ekin;
  out[3] = eot;
  out[4] = eext;
  out[5] = (ekin + ev + eext) / nrm;
  free(eps); free(rho); free(pot);
}

/* ------------------------------------------------- bulk forms (a13) -- */

/* src/helium-ot-bulk.c:21-33 */
double orc_bulk_energy(const orc_ot *o, double rho) {
  double e = 0.5 * o->b * rho * rho + 0.5 * o->c2 * rho * rho * rho + (1.0 / 3.0) * o->c3 * rho * rho * rho * rho;
  if (o->c4 != 0.0) e += orc_ig_energy(o, rho);
  return e;
}
/* src/helium-ot-bulk.c:45-57 */
double orc_bulk_dEdRho(const orc_ot *o, double rho) {
  double d = o->b * rho + 1.5 * o->c2 * rho * rho + (4.0 / 3.0) * o->c3 * rho * rho * rho;
  if (o->c4 != 0.0) d += orc_ig_dEdRho(o, rho);
  return d;
}
/* src/helium-ot-bulk.c:143-146 */
double orc_bulk_pressure(const orc_ot *o, double rho) { return rho * orc_bulk_dEdRho(o, rho) - orc_bulk_energy(o, rho); }
/* src/helium-ot-bulk.c:157-172 (central FD, eps 1e-7) */
double orc_bulk_dPdRho(const orc_ot *o, double rho) {
  const double e = 1e-7;
  return (orc_bulk_pressure(o, rho + e) - orc_bulk_pressure(o, rho - e)) / (2.0 * e);
}
/* src/helium-ot-bulk.c:185-211 (Newton from rho = 1, |misP| < 1e-12, 1000 iterations) */
double orc_bulk_density_pressurized(const orc_ot *o, double pressure) {
  double rho0 = 1.0, misP = orc_bulk_pressure(o, rho0) - pressure;
  for (int i = 0; i < 1000; i++) {
    if (fabs(misP) < 1e-12) return rho0;
    rho0 -= misP / orc_bulk_dPdRho(o, rho0);
    misP = orc_bulk_pressure(o, rho0) - pressure;
  }
  return NAN;
}
double orc_bulk_chempot_pressurized(const orc_ot *o, double pressure) {
  return orc_bulk_dEdRho(o, orc_bulk_density_pressurized(o, pressure));
}
double orc_bulk_compressibility(const orc_ot *o, double rho) { return 1.0 / (rho * orc_bulk_dPdRho(o, rho)); }
double orc_bulk_sound_speed(const orc_ot *o, double rho) { return 1.0 / sqrt(o->mass * rho * orc_bulk_compressibility(o, rho)); }

/* src/helium-ot-bulk.c:379-409 */
static double ft_lj(const orc_ot *o, double k) {
  double sigma = o->lj_sigma, eps = o->lj_epsilon, h = o->lj_h, val = 0.0, ks = k * sigma;
  for (double x = h / sigma; x < 1e3; x += 1e-2) val += sin(ks * x) * (pow(x, -11.0) - pow(x, -5.0));
  return val * 1e-2 * 16.0 * M_PI * eps * sigma * sigma * sigma / ks;
}
static double ft_pi(const orc_ot *o, double k) {
  double kh = k * o->lj_h;
  return (3.0 / (kh * kh * kh)) * (sin(kh) - kh * cos(kh));
}
static double ft_vj(const orc_ot *o, double k) {
  double k2 = k * k, m = pow(M_PI, 1.5);
  double e1 = exp(-k2 / (4.0 * o->a1)), e2 = exp(-k2 / (4.0 * o->a2));
  return (o->g11 * pow(M_PI / o->a1, 1.5) + o->g12 * (6.0 * o->a1 - k2) * m / (4.0 * pow(o->a1, 3.5))) * e1
       + (o->g21 * pow(M_PI / o->a2, 1.5) + o->g22 * (6.0 * o->a2 - k2) * m / (4.0 * pow(o->a2, 3.5))) * e2;
}
/* src/helium-ot-bulk.c:445-463 */
double orc_bulk_istatic(const orc_ot *o, double k, double rho0) {
  if (rho0 < 0.0) return 0.0;
  if (k < 1e-2) return 1e-2;
  double lj = ft_lj(o, k), pi = ft_pi(o, k);
  double ikai = k * k / (4.0 * o->mass) + rho0 * lj + o->c2 * (2.0 * pi + pi * pi) * rho0 * rho0
              + 2.0 * o->c3 * (pi + pi * pi) * rho0 * rho0 * rho0;
  if (o->model & ORC_OT_KC)
    ikai += -(1.0 / (2.0 * o->mass)) * o->alpha_s * rho0 * (1.0 - rho0 / o->rho_0s) * (1.0 - rho0 / o->rho_0s)
            * exp(-k * k * o->l_g * o->l_g / 4.0);
  return ikai;
}
/* src/helium-ot-bulk.c:411-434 */
double orc_bulk_dispersion(const orc_ot *o, double k, double rho0) {
  if (rho0 < 0.0 || k < 1e-3) return 0.0;
  double lj = ft_lj(o, k), pi = ft_pi(o, k), tmp = k * k / o->mass;
  double ikai = tmp / 4.0 + rho0 * lj + o->c2 * (2.0 * pi + pi * pi) * rho0 * rho0
              + 2.0 * o->c3 * (pi + pi * pi) * rho0 * rho0 * rho0;
  if (o->model & ORC_OT_KC)
    ikai += -(tmp / 2.0) * o->alpha_s * rho0 * (1.0 - rho0 / o->rho_0s) * (1.0 - rho0 / o->rho_0s)
            * exp(-k * k * o->l_g * o->l_g / 4.0);
  if (o->model & ORC_OT_BACKFLOW) return sqrt(tmp * ikai * (1.0 - rho0 * (ft_vj(o, 0.0) - ft_vj(o, k))));
  return sqrt(tmp * ikai);
}
