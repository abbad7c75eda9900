// Grid files (.grd) for checkpoints and restarts.  The layout is this
// library's own (libgrid's is not part of the reference):
//   8 bytes  "LDFTGRD1"
//   int64    kind: 1 real field, 2 complex field, 3 real half spectrum
//   int64    nx, ny, nz
//   double   step, x0, y0, z0
//   data     FP64 row-major [i][j][k], k fastest; complex as (re, im); a half
//            spectrum (a functional's kernel grid: the kernels are even, so
//            their spectra are real) as [nx][ny][nz/2+1] real values
// Values are written in FP64 whatever the device precision of the grid and
// rounded to it on reading.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/libdft_mi355x.h"

namespace {

const char kMagic[8] = {'L', 'D', 'F', 'T', 'G', 'R', 'D', '1'};
enum : int64_t { kReal = 1, kComplex = 2, kHalf = 3 };

struct Header {
  char magic[8];
  int64_t kind, nx, ny, nz;
  double step, x0, y0, z0;
};

int64_t points(INT nx, INT ny, INT nz) { return (int64_t)nx * ny * nz; }
int64_t half_points(INT nx, INT ny, INT nz) { return (int64_t)nx * ny * (nz / 2 + 1); }

bool read_header(FILE *fp, Header &h, const char *who) {
  if (fread(&h, sizeof h, 1, fp) != 1 || memcmp(h.magic, kMagic, sizeof kMagic) != 0) {
    fprintf(stderr, "libdft: %s: not a grid file of this library\n", who);
    return false;
  }
  return true;
}

bool same_shape(const Header &h, INT nx, INT ny, INT nz, const char *who) {
  if (h.nx != nx || h.ny != ny || h.nz != nz) {
    fprintf(stderr, "libdft: %s: file grid %lld x %lld x %lld, grid %lld x %lld x %lld\n", who, (long long)h.nx,
            (long long)h.ny, (long long)h.nz, (long long)nx, (long long)ny, (long long)nz);
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

void rgrid_write(rgrid *grid, FILE *fp) {
  const int64_t n = grid->spectral ? half_points(grid->nx, grid->ny, grid->nz) : points(grid->nx, grid->ny, grid->nz);
  std::vector<REAL> v(n);
  rgrid_copy_to_host(grid, v.data());
  Header h;
  memcpy(h.magic, kMagic, sizeof kMagic);
  h.kind = grid->spectral ? kHalf : kReal;
  h.nx = grid->nx, h.ny = grid->ny, h.nz = grid->nz;
  h.step = grid->step, h.x0 = grid->x0, h.y0 = grid->y0, h.z0 = grid->z0;
  if (fwrite(&h, sizeof h, 1, fp) != 1 || fwrite(v.data(), sizeof(REAL), n, fp) != (size_t)n)
    fprintf(stderr, "libdft: rgrid_write: write failed\n");
}

rgrid *rgrid_read(rgrid *grid, FILE *fp) {
  Header h;
  if (!read_header(fp, h, "rgrid_read")) return nullptr;
  if (h.kind != kReal && h.kind != kHalf) {
    fprintf(stderr, "libdft: rgrid_read: the file holds a complex grid\n");
    return nullptr;
  }
  const bool own = grid == nullptr;
  if (own) {
    // a new grid has room for a real field or a half spectrum
    grid = rgrid_alloc(h.nx, h.ny, h.nz, h.step, nullptr, nullptr, (char *)"rgrid_read");
    if (!grid) return nullptr;
    rgrid_set_origin(grid, h.x0, h.y0, h.z0);
    grid->spectral = h.kind == kHalf;
  } else if (!same_shape(h, grid->nx, grid->ny, grid->nz, "rgrid_read")) {
    return nullptr;
  } else if ((h.kind == kHalf) != (grid->spectral != 0)) {
    fprintf(stderr, "libdft: rgrid_read: the file holds a %s, grid %s a %s\n",
            h.kind == kHalf ? "half spectrum" : "real field", grid->id, grid->spectral ? "half spectrum" : "real field");
    return nullptr;
  }
  const int64_t n = h.kind == kHalf ? half_points(h.nx, h.ny, h.nz) : points(h.nx, h.ny, h.nz);
  std::vector<REAL> v(n);
  if (fread(v.data(), sizeof(REAL), n, fp) != (size_t)n) {
    fprintf(stderr, "libdft: rgrid_read: file too short\n");
    if (own) rgrid_free(grid);
    return nullptr;
  }
  rgrid_copy_from_host(grid, v.data());
  return grid;
}

void cgrid_write(cgrid *grid, FILE *fp) {
  const int64_t n = points(grid->nx, grid->ny, grid->nz);
  std::vector<REAL_COMPLEX> v(n);
  cgrid_copy_to_host(grid, v.data());
  Header h;
  memcpy(h.magic, kMagic, sizeof kMagic);
  h.kind = kComplex;
  h.nx = grid->nx, h.ny = grid->ny, h.nz = grid->nz;
  h.step = grid->step, h.x0 = grid->x0, h.y0 = grid->y0, h.z0 = grid->z0;
  if (fwrite(&h, sizeof h, 1, fp) != 1 || fwrite(v.data(), sizeof(REAL_COMPLEX), n, fp) != (size_t)n)
    fprintf(stderr, "libdft: cgrid_write: write failed\n");
}

cgrid *cgrid_read(cgrid *grid, FILE *fp) {
  Header h;
  if (!read_header(fp, h, "cgrid_read")) return nullptr;
  if (h.kind != kComplex) {
    fprintf(stderr, "libdft: cgrid_read: the file holds a real grid\n");
    return nullptr;
  }
  const bool own = grid == nullptr;
  if (own) {
    grid = cgrid_alloc(h.nx, h.ny, h.nz, h.step, nullptr, nullptr, (char *)"cgrid_read");
    if (!grid) return nullptr;
    cgrid_set_origin(grid, h.x0, h.y0, h.z0);
  } else if (!same_shape(h, grid->nx, grid->ny, grid->nz, "cgrid_read")) {
    return nullptr;
  }
  const int64_t n = points(h.nx, h.ny, h.nz);
  std::vector<REAL_COMPLEX> v(n);
  if (fread(v.data(), sizeof(REAL_COMPLEX), n, fp) != (size_t)n) {
    fprintf(stderr, "libdft: cgrid_read: file too short\n");
    if (own) cgrid_free(grid);
    return nullptr;
  }
  cgrid_copy_from_host(grid, v.data());
  return grid;
}

// <base>.grd
void rgrid_write_grid(char *base, rgrid *grid) {
  const std::string name = std::string(base) + ".grd";
  FILE *fp = fopen(name.c_str(), "wb");
  if (!fp) {
    fprintf(stderr, "libdft: rgrid_write_grid: cannot open %s\n", name.c_str());
    return;
  }
  rgrid_write(grid, fp);
  fclose(fp);
}

void cgrid_write_grid(char *base, cgrid *grid) {
  const std::string name = std::string(base) + ".grd";
  FILE *fp = fopen(name.c_str(), "wb");
  if (!fp) {
    fprintf(stderr, "libdft: cgrid_write_grid: cannot open %s\n", name.c_str());
    return;
  }
  cgrid_write(grid, fp);
  fclose(fp);
}

void rgrid_read_grid(rgrid *grid, char *file) {
  FILE *fp = fopen(file, "rb");
  if (!fp) {
    fprintf(stderr, "libdft: rgrid_read_grid: cannot open %s\n", file);
    return;
  }
  rgrid_read(grid, fp);
  fclose(fp);
}

void cgrid_read_grid(cgrid *grid, char *file) {
  FILE *fp = fopen(file, "rb");
  if (!fp) {
    fprintf(stderr, "libdft: cgrid_read_grid: cannot open %s\n", file);
    return;
  }
  cgrid_read(grid, fp);
  fclose(fp);
}

}  // extern "C"
