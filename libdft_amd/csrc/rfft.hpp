// Register-resident line FFTs for gfx950.
//
// A line of N points is held by TPL = N / E threads, E points each, in the
// decimated layout
//
//     register i of thread t  <->  point t + TPL * i        (i < E, t < TPL).
//
// A Stockham autosort transform with radices R_0..R_{S-1} (each dividing E)
// runs stage s as E / R_s radix-R_s butterflies per thread, entirely in VGPRs.
// Between stages the points are exchanged through LDS (S - 1 exchanges per
// transform: one for N = 64..256, two for N = 512); the last stage leaves its
// outputs in the same decimated layout, in natural frequency order.  So a
// pass loads a line from HBM straight into registers, transforms, applies its
// k-space or pointwise work in registers and stores: no LDS round trip for the
// data itself, and the forward and inverse transforms share one layout.
//
// Exchanges are synchronised either per wavefront (all TPL threads of a line
// in one wave: z passes, where lines are contiguous and t-fastest lanes are
// already coalesced) or per block (column-tile passes along x and y, where a
// column-fastest lane order gives 128-byte row segments and a line's threads
// span several waves).
#pragma once
#include <type_traits>

#include "fft.hpp"

namespace dft {
namespace rf {

// Radix plans.  EMAX bounds the points per thread: 16 for passes that hold one
// field per thread (more work per exchange), 8 for fused stages that keep
// several fields live at once (half the registers per field, one more
// exchange for N = 128 and 256).
template <int N, int EMAX> struct Radix;
template <int EMAX> struct Radix<4, EMAX> { static constexpr int S = 1, R0 = 4, R1 = 1, R2 = 1; };
template <int EMAX> struct Radix<8, EMAX> { static constexpr int S = 1, R0 = 8, R1 = 1, R2 = 1; };
template <> struct Radix<16, 16> { static constexpr int S = 1, R0 = 16, R1 = 1, R2 = 1; };
template <> struct Radix<16, 8> { static constexpr int S = 2, R0 = 4, R1 = 4, R2 = 1; };
template <int EMAX> struct Radix<32, EMAX> { static constexpr int S = 2, R0 = 8, R1 = 4, R2 = 1; };
template <int EMAX> struct Radix<64, EMAX> { static constexpr int S = 2, R0 = 8, R1 = 8, R2 = 1; };
template <> struct Radix<128, 16> { static constexpr int S = 2, R0 = 16, R1 = 8, R2 = 1; };
template <> struct Radix<128, 8> { static constexpr int S = 3, R0 = 8, R1 = 8, R2 = 2; };
template <> struct Radix<256, 16> { static constexpr int S = 2, R0 = 16, R1 = 16, R2 = 1; };
template <> struct Radix<256, 8> { static constexpr int S = 3, R0 = 8, R1 = 8, R2 = 4; };
template <int EMAX> struct Radix<512, EMAX> { static constexpr int S = 3, R0 = 8, R1 = 8, R2 = 8; };
// 4 points per thread (N <= 256): the fused stages with the most live fields.
template <> struct Radix<4, 4> { static constexpr int S = 1, R0 = 4, R1 = 1, R2 = 1, R3 = 1; };
template <> struct Radix<8, 4> { static constexpr int S = 2, R0 = 4, R1 = 2, R2 = 1, R3 = 1; };
template <> struct Radix<16, 4> { static constexpr int S = 2, R0 = 4, R1 = 4, R2 = 1, R3 = 1; };
template <> struct Radix<32, 4> { static constexpr int S = 3, R0 = 4, R1 = 4, R2 = 2, R3 = 1; };
template <> struct Radix<64, 4> { static constexpr int S = 3, R0 = 4, R1 = 4, R2 = 4, R3 = 1; };
template <> struct Radix<128, 4> { static constexpr int S = 4, R0 = 4, R1 = 4, R2 = 4, R3 = 2; };
template <> struct Radix<256, 4> { static constexpr int S = 4, R0 = 4, R1 = 4, R2 = 4, R3 = 4; };
// N = 512 with 4 points per thread: 128 threads (two wavefronts) per line.
template <> struct Radix<512, 4> { static constexpr int S = 5, R0 = 4, R1 = 4, R2 = 4, R3 = 4, R4 = 2; };

template <typename P, typename = void> struct R3of { static constexpr int v = 1; };
template <typename P> struct R3of<P, decltype((void)P::R3)> { static constexpr int v = P::R3; };
template <typename P, typename = void> struct R4of { static constexpr int v = 1; };
template <typename P> struct R4of<P, decltype((void)P::R4)> { static constexpr int v = P::R4; };
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int N, int EMAX = 16> struct RP {
  using P = Radix<N, EMAX>;
  static constexpr int S = P::S, R0 = P::R0, R1 = P::R1, R2 = P::R2, R3 = R3of<P>::v, R4 = R4of<P>::v;
  static constexpr int E = cmax(cmax(cmax(R0, R1), cmax(R2, R3)), R4);
  static constexpr int TPL = N / E;
  static_assert(E % R0 == 0 && E % R1 == 0 && E % R2 == 0 && E % R3 == 0 && E % R4 == 0, "radices must divide E");
};

// The registers of one thread's share of a line.
template <typename T, int N, int EM> using Regs = cplx<T>[RP<N, EM>::E];

struct BlockSync {
  __device__ __forceinline__ static void sync() { __syncthreads(); }
};
// Lines whose threads are consecutive lanes (as under WaveSync) but span
// several wavefronts: the block barrier orders the exchange.
struct LineBlockSync {
  __device__ __forceinline__ static void sync() { __syncthreads(); }
};
// All threads of a line are lanes of one wavefront: LDS operations of a wave
// complete in issue order, so only the compiler has to be kept from moving
// them across the exchange.
struct WaveSync {
  __device__ __forceinline__ static void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// A zero the compiler cannot see through, produced where it is called.  Loads
// from memory that the kernel has not yet written are treated as invariant
// and would otherwise be hoisted to the kernel entry (all stages' loads live
// at once); adding this to a stage's offsets keeps its loads in that stage.
__device__ __forceinline__ uint32_t opaque_zero() {
  uint32_t z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}

// LDS layout of one line's exchange buffer: point p sits at element pad(p),
// LP elements per line.  LDS serves 128 bytes per cycle over 32 four-byte
// banks, i.e. 16 lanes of 8-byte (FP32 complex) or 8 lanes of 16-byte (FP64
// complex) accesses; the layout is chosen per plan so that both the stage
// writes (points strided by the earlier radices) and the reads (consecutive
// points) of those lanes fall in distinct banks.
//  - wave-synchronised lines (a line is one wavefront's lanes, t fastest),
//    N >= 64 with 4 or 8 points per thread, or N >= 256: XOR swizzle inside
//    each 16-element block, p ^ (5 * (p / 16) mod 16), no padding;
//  - column tiles (lanes column-fastest, one line per column), N >= 64: one
//    pad element every 8, and a line pitch that offsets neighbouring columns
//    by an odd number of elements (by 2 mod 4 for 8 points per thread and
//    N >= 256 in FP32);
//  - otherwise one pad element every 16.
template <int N, int EM, bool WAVE, int W> struct Lay {
  static constexpr bool XOR = WAVE && N >= 64 && (EM <= 8 || N >= 256);
  static constexpr bool ADD3 = !WAVE && N >= 64;
  static constexpr int MAXP = XOR ? N - 1 : (ADD3 ? (N - 1) + (N - 1) / 8 : (N - 1) + (N - 1) / 16);
  static constexpr int LP = XOR ? N
                            : WAVE ? N + N / 16 + 2
                                   : (ADD3 && RP<N, EM>::E == 8 && N >= 256 && W == 8 ? N + N / 8 + 2
                                                                                       : ((MAXP + 1) | 1));
  static_assert(LP > MAXP, "line pitch too small for the layout");
  __device__ __forceinline__ static int pad(int p) {
    if constexpr (XOR) return p ^ (((p >> 4) * 5) & 15);
    else if constexpr (ADD3) return p + (p >> 3);
    else return p + (p >> 4);
  }
};
// line-major lanes (the layout choice), whatever the barrier
template <class SYNC>
constexpr bool kWave = std::is_same<SYNC, WaveSync>::value || std::is_same<SYNC, LineBlockSync>::value;
template <typename T, int N, int EM, class SYNC> using LayOf = Lay<N, EM, kWave<SYNC>, (int)sizeof(cplx<T>)>;

// Inter-stage twiddles in LDS, stage by stage: stage s >= 1 (radix R_s, NS_s =
// R_0 ... R_{s-1}) uses w^(k r N / (NS_s R_s)) for r = 1..R_s-1, k < NS_s, at
// entry (NS_s - R_0) + (r - 1) NS_s + k, so the lanes of a line (consecutive k)
// read consecutive entries.  N - R_0 entries in all; the FP32 rounding
// residuals follow at offset N.
template <int N, int EM> __device__ __forceinline__ int tw_entry(int NS, int r, int k) {
  return (NS - RP<N, EM>::R0) + (r - 1) * NS + k;
}

// a * w^(DIR) with w = tw[idx] (forward table; its FP32 rounding residual at
// tw[N + idx]).
template <typename T, int N, int DIR>
__device__ __forceinline__ cplx<T> twmul(cplx<T> a, const cplx<T> *tw, int idx) {
  cplx<T> w = tw[idx];
  if (DIR > 0) w.y = -w.y;
  if constexpr (sizeof(T) == 4) {
    cplx<T> wl = tw[N + idx];
    if (DIR > 0) wl.y = -wl.y;
    return fft::cmul_hl(a, w, wl);
  } else {
    return cmul(a, w);
  }
}

// One radix-R Stockham stage; NS = product of the earlier radices.
template <typename T, int N, int EM, int R, int NS, bool LAST, int DIR, class SYNC>
__device__ __forceinline__ void rstage(Regs<T, N, EM> &v, cplx<T> *X, const cplx<T> *tw, int t) {
  using C = cplx<T>;
  constexpr int E = RP<N, EM>::E, TPL = RP<N, EM>::TPL, NB = E / R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + TPL * b;
    C u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[b + r * NB];
    if constexpr (NS > 1) {
      const int k = j & (NS - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) u[r] = twmul<T, N, DIR>(u[r], tw, tw_entry<N, EM>(NS, r, k));
    }
    fft::dft<R, DIR>(u);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * NB] = u[r];
    } else {
      const int d = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
      for (int r = 0; r < R; ++r) X[LayOf<T, N, EM, SYNC>::pad(d + r * NS)] = u[r];
    }
  }
  if constexpr (!LAST) {
    SYNC::sync();
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = X[LayOf<T, N, EM, SYNC>::pad(t + TPL * i)];
    SYNC::sync();
  }
}

// Unnormalised transform of one line (DIR -1 forward, +1 inverse) in the
// decimated layout of plan RP<N, EM>.  X: this line's LDS exchange buffer
// (LayOf::LP elements); tw: twiddle table of load_tw.  Every thread of the sync
// group must call it.
template <typename T, int N, int EM, int DIR, class SYNC>
__device__ __forceinline__ void rfft(Regs<T, N, EM> &v, cplx<T> *X, const cplx<T> *tw_, int t) {
  using P = RP<N, EM>;
  // Twiddles are re-read from LDS by every transform: the table is never
  // written after set-up, so the compiler would otherwise keep the values of
  // the first transform live in registers through all later ones.
  const cplx<T> *tw = tw_ + opaque_zero();
  static_assert(!std::is_same<SYNC, WaveSync>::value || P::TPL <= 64, "a wave-synchronised line must fit one wavefront");
  if constexpr (P::S == 1) {
    rstage<T, N, EM, P::R0, 1, true, DIR, SYNC>(v, X, tw, t);
  } else if constexpr (P::S == 2) {
    rstage<T, N, EM, P::R0, 1, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R1, P::R0, true, DIR, SYNC>(v, X, tw, t);
  } else if constexpr (P::S == 3) {
    rstage<T, N, EM, P::R0, 1, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R1, P::R0, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R2, P::R0 * P::R1, true, DIR, SYNC>(v, X, tw, t);
  } else if constexpr (P::S == 4) {
    rstage<T, N, EM, P::R0, 1, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R1, P::R0, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R2, P::R0 * P::R1, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R3, P::R0 * P::R1 * P::R2, true, DIR, SYNC>(v, X, tw, t);
  } else {
    static_assert(P::S == 5, "at most five stages");
    rstage<T, N, EM, P::R0, 1, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R1, P::R0, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R2, P::R0 * P::R1, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R3, P::R0 * P::R1 * P::R2, false, DIR, SYNC>(v, X, tw, t);
    rstage<T, N, EM, P::R4, P::R0 * P::R1 * P::R2 * P::R3, true, DIR, SYNC>(v, X, tw, t);
  }
}

// LDS entries of a twiddle table: the N - R0 stage-ordered twiddles, and for
// FP32 their rounding residuals at offset N.
template <typename T, int N> constexpr int tw_len() { return sizeof(T) == 4 ? 2 * N : N; }

// Gather the stage-ordered twiddles of plan RP<N, EM> (see tw_entry) from the
// natural table g_tw (w^m at m, FP32 residual at N + m) into LDS; callers
// synchronise before use.
template <typename T, int N, int EM, int BS>
__device__ __forceinline__ void load_tw(cplx<T> *s_tw, const cplx<T> *g_tw) {
  using P = RP<N, EM>;
  constexpr int NS1 = P::R0, NS2 = NS1 * P::R1, NS3 = NS2 * P::R2, NS4 = NS3 * P::R3;
  for (int e = threadIdx.x; e < N - P::R0; e += BS) {
    const int q = e + P::R0;
    int ns, R;
    if (P::S >= 5 && q >= NS4) ns = NS4, R = P::R4;
    else if (P::S >= 4 && q >= NS3) ns = NS3, R = P::R3;
    else if (P::S >= 3 && q >= NS2) ns = NS2, R = P::R2;
    else ns = NS1, R = P::R1;
    const int o = q - ns, r = o / ns + 1, k = o - (r - 1) * ns;
    const int m = k * r * (N / (ns * R));
    s_tw[e] = g_tw[m];
    if constexpr (sizeof(T) == 4) s_tw[N + e] = g_tw[N + m];
  }
}

// Global accesses as base pointer + 32-bit per-thread byte offset + constant.
// Every field is below 4 GiB, so the per-thread part is one VGPR; a base that
// is uniform across the wave (kernel argument, optionally advanced by a
// uniform stride) stays in SGPRs and the constant folds into the
// instruction's immediate offset, so a thread addresses all its points
// through one or two VGPRs instead of one 64-bit address per point.
template <typename X> __device__ __forceinline__ X gld(const void *base, uint32_t off, int imm = 0) {
  return *(const X *)((const char *)base + off + imm);
}
// A load of data that this kernel reads exactly once (DFT_NT_LOAD), marked
// non-temporal.  Used where it measured faster (the spectra the z passes
// read); the column passes keep ordinary loads, which measured faster there.
#ifndef DFT_NT_LOAD
#define DFT_NT_LOAD 1
#endif
template <typename X> __device__ __forceinline__ X gld_once(const void *base, uint32_t off, int imm = 0) {
  const X *p = (const X *)((const char *)base + off + imm);
  if constexpr (DFT_NT_LOAD && (sizeof(X) == 4 || sizeof(X) == 8 || sizeof(X) == 16)) {
    using V = uint32_t __attribute__((ext_vector_type(sizeof(X) / 4)));
    return __builtin_bit_cast(X, __builtin_nontemporal_load((const V *)p));
  } else {
    return *p;
  }
}
// Stores are marked non-temporal (DFT_NT_STORE): every pass writes fields
// that only a later kernel reads, so they need not displace the data (stencil
// neighbours, re-read inputs) that the writing kernel itself still reads.
#ifndef DFT_NT_STORE
#define DFT_NT_STORE 1
#endif
template <typename X> __device__ __forceinline__ void gst(void *base, uint32_t off, X v, int imm = 0) {
  X *p = (X *)((char *)base + off + imm);
  if constexpr (DFT_NT_STORE && (sizeof(X) == 4 || sizeof(X) == 8 || sizeof(X) == 16)) {
    using V = uint32_t __attribute__((ext_vector_type(sizeof(X) / 4)));
    __builtin_nontemporal_store(__builtin_bit_cast(V, v), (V *)p);
  } else {
    *p = v;
  }
}

// Logical block index with consecutive indices on one XCD: blocks are
// dispatched round-robin over the 8 XCDs, each with its own L2, so with this
// order the column tiles of one transform line (consecutive indices) share an
// L2.  Identity when the grid is not a multiple of 8.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, n = gridDim.x;
  return (n & 7) == 0 ? (b & 7) * (n >> 3) + (b >> 3) : b;
}

// Element-indexed forms: element `e` (per thread) plus constant element delta `d`.
template <typename X> __device__ __forceinline__ X lde(const void *f, uint32_t e, int d = 0) {
  return gld<X>(f, e * (uint32_t)sizeof(X), d * (int)sizeof(X));
}
template <typename X> __device__ __forceinline__ X lde_once(const void *f, uint32_t e, int d = 0) {
  return gld_once<X>(f, e * (uint32_t)sizeof(X), d * (int)sizeof(X));
}
template <typename X> __device__ __forceinline__ void ste(void *f, uint32_t e, X v, int d = 0) {
  gst<X>(f, e * (uint32_t)sizeof(X), v, d * (int)sizeof(X));
}

// Column-tile geometry for passes along a strided axis: COLS columns per
// block, column-fastest lanes (plan with up to 16 points per thread).
// At N = 512 in FP32 (8 points per thread, 64 threads per column) 16 columns
// (1024 threads) give 128-byte row segments.
// Small grids: blocks of DFT_COL_BS64 / DFT_COL_BS128 threads at N <= 64 / 128
// (instead of 256) and at least DFT_COL_MINC_SMALL columns, so that a pass over
// a small grid still has blocks for every CU (64^3 plain OT: 13.0k -> 15.3k
// steps/s; 128^3 full OT FP32: 2.48k -> 2.61k).  SMALL = false keeps a pass at
// 256 threads (the single-term x-middle outputs at N = 128, whose 7 outputs
// already fill the device).
#ifndef DFT_COL_BS64
#define DFT_COL_BS64 64
#endif
#ifndef DFT_COL_BS128
#define DFT_COL_BS128 128
#endif
#ifndef DFT_COL_MINC_SMALL
#define DFT_COL_MINC_SMALL 4
#endif
template <typename T, int N, int EMAX = 16, int MINCOLS = (N >= 512 && sizeof(T) == 4 ? 16 : 8), bool SMALL = true>
struct ColCfg {
  static constexpr int EM = EMAX;
  static constexpr int E = RP<N, EM>::E, TPL = RP<N, EM>::TPL, S = RP<N, EM>::S;
  static constexpr int LP = Lay<N, EM, false, (int)sizeof(cplx<T>)>::LP;
  static constexpr int TBS = !SMALL ? 256 : (N <= 64 ? DFT_COL_BS64 : (N <= 128 ? DFT_COL_BS128 : 256));
  static constexpr int MINC = SMALL && N <= 128 && DFT_COL_MINC_SMALL < MINCOLS ? DFT_COL_MINC_SMALL : MINCOLS;
  static constexpr int COLS = TBS / TPL < MINC ? MINC : TBS / TPL;
  static constexpr int BS = COLS * TPL;
  static_assert(BS <= 1024, "column block larger than 1024 threads");
};

// Line geometry for passes along the contiguous axis: one line (or line pair)
// per TPL consecutive lanes, 256 threads per block (DFT_LINE_BS64 at N <= 64,
// so that the z passes of a small grid have blocks for every CU).
#ifndef DFT_LINE_BS64
#define DFT_LINE_BS64 64
#endif
template <typename T, int N, int EMAX = 16> struct LineCfg {
  static constexpr int EM = EMAX;
  static constexpr int E = RP<N, EM>::E, TPL = RP<N, EM>::TPL, S = RP<N, EM>::S;
  static constexpr int LP = Lay<N, EM, true, (int)sizeof(cplx<T>)>::LP;
  static constexpr int BS = N <= 64 && DFT_LINE_BS64 >= TPL ? DFT_LINE_BS64 : 256;
  static constexpr int LINES = BS / TPL;
};

}  // namespace rf
}  // namespace dft
