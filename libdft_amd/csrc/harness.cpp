// Host-side harnesses over the C ABI.
#include <cmath>
#include <cstdio>

#include "../../include/libdft_mi355x.h"

namespace {

struct PlaneWave {
  double kx, ky, kz, a, rho;
};

// sqrt(rho) (1 + a cos(k.r)) (src/common.c:1214-1224)
REAL_COMPLEX plane_wave(void *arg, REAL x, REAL y, REAL z) {
  const PlaneWave *p = (const PlaneWave *)arg;
  REAL_COMPLEX v;
  __real__ v = std::sqrt(p->rho) * (1.0 + p->a * std::cos(p->kx * x + p->ky * y + p->kz * z));
  __imag__ v = 0.0;
  return v;
}

}  // namespace

extern "C" {

// Excitation energy of the density wave k along axis dir from the dynamics
// (src/helium-ot-bulk.c:274-361): psi = sqrt(rho0)(1 + amp cos(k r)) with k
// rounded to the grid (and returned in *k), real-time steps ts (single steps,
// or predict-correct with -mu0 when pred), until the running sum of
// rho(centre) - rho0 turns negative after l steps, i.e. half a period: the
// frequency 1 / (2 l ts), returned as an energy (hartree).  Returns 0 for
// k = 0 and -1 (with a message) for an invalid direction or when no half
// period is found within 10^7 steps.
REAL dft_ot_dispersion(wf *gwf, dft_ot_functional *otf, REAL ts, REAL *k, REAL amp, char pred, char dir) {
  if (dir < 0 || dir > 2) {
    fprintf(stderr, "libdft: dft_ot_dispersion: the direction must be 0, 1 or 2\n");
    return -1.0;
  }
  const INT n[3] = {gwf->grid->nx, gwf->grid->ny, gwf->grid->nz};
  const REAL step = gwf->grid->step;
  if (n[(int)dir] < 2) {
    fprintf(stderr, "libdft: dft_ot_dispersion: one point along the wave direction\n");
    return -1.0;
  }
  const REAL dk = 2.0 * M_PI / ((REAL)n[(int)dir] * step);
  const REAL kq = (REAL)(INT)(0.5 + *k / dk) * dk;
  *k = kq;
  if (kq == 0.0) return 0.0;
  PlaneWave pw{0.0, 0.0, 0.0, amp, otf->rho0};
  (dir == 0 ? pw.kx : dir == 1 ? pw.ky : pw.kz) = kq;
  grid_wf_map(gwf, plane_wave, &pw);
  wf *gwfp = pred ? grid_wf_clone(gwf, (char *)"dft_ot_dispersion predictor") : nullptr;
  const REAL mu0 = dft_ot_bulk_chempot2(otf);
  REAL_COMPLEX dt;
  __real__ dt = ts;
  __imag__ dt = 0.0;
  double sum = 0.0;
  INT l = 0;
  const INT cap = 10000000;
  for (;; ++l) {
    if (l >= cap) {
      fprintf(stderr, "libdft: dft_ot_dispersion: no half period within %lld steps\n", (long long)cap);
      if (gwfp) grid_wf_free(gwfp);
      return -1.0;
    }
    if (pred) dft_ot_propagate_pc(otf, gwf, gwfp, nullptr, mu0, dt);
    else dft_ot_propagate(otf, gwf, nullptr, 0.0, dt);
    const REAL_COMPLEX c = cgrid_value_at_index(gwf->grid, n[0] / 2, n[1] / 2, n[2] / 2);
    sum += __real__ c * __real__ c + __imag__ c * __imag__ c - otf->rho0;
    if (sum < 0.0) {
      --l;
      break;
    }
  }
  if (gwfp) grid_wf_free(gwfp);
  const double nu = 1.0 / (2.0 * (double)l * ts);  // per atomic unit of time
  return (nu / GRID_AUTOS) * GRID_HZTOCM1 * 1.439 / GRID_AUTOK;
}

}  // extern "C"
