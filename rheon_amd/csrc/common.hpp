// Shared device/host definitions for the rheon_amd HIP kernels (gfx950).
//
// Layout: every cell array is x-fastest, index i + nx*(j + ny*k); face array
// of axis a has counts n + e_a and the same x-fastest order
// (src/geometry.rs:250-258).  Kernels walk "rows" (fixed j,k) with 64 lanes
// along x, so every wave load is one contiguous, coalesced x segment.
//
// All code is built with -ffp-contract=off: each multiply and add rounds
// separately, as in the reference, so the f64 paths are bit-comparable.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/rheon_gpu.h"

namespace rh {

constexpr int kLanesX = 64;        // lanes along x (one wavefront)
constexpr int kRowsPerBlock = 4;   // rows (j,k) per block: 4 waves = 256 threads
constexpr int kBlock = kLanesX * kRowsPerBlock;
// Launch shape.  Each block takes kGroupsPerBlock consecutive groups of
// kRowsPerBlock rows (or of kBlock four-cell quads), up to kMaxBlocks
// blocks (a 256^3 box and its face lattices stay under the cap; larger
// lattices give each block a longer contiguous run).  Blocks dispatched in
// order keep the active window of the whole chip on one compact region of
// memory, the access pattern HBM serves fastest: a 7-point sweep at 256^3
// ran at 4.2-4.9 TB/s this way against 3.2 TB/s with persistent blocks
// owning per-XCD slabs.  Two groups per block (rather than one) halve the
// per-block start-up latency (control-block load, reduction) that short
// kernels pay once per residency round.
#ifndef RH_MAX_BLOCKS
#define RH_MAX_BLOCKS 16384
#endif
constexpr int kMaxBlocks = RH_MAX_BLOCKS;
#ifndef RH_GROUPS_PER_BLOCK
#define RH_GROUPS_PER_BLOCK 2
#endif
constexpr int kGroupsPerBlock = RH_GROUPS_PER_BLOCK;

// floor(n / d) for 0 <= n < 2^31 and a divisor d >= 1 fixed per launch,
// as a multiply-high, an add and a shift (Granlund and Montgomery, "Division
// by invariant integers using multiplication", 1994, the round-up method):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, then
// floor(n / d) = (mulhi(m, n) + n) >> l; the sum stays below 2^32 because
// mulhi(m, n) <= n < 2^31.  A runtime integer division costs ~18 VALU
// instructions; kernels that map a flat index to (row, chunk) or (j, k)
// every iteration take this form instead.
struct FastDiv {
  unsigned m;
  int l;
};
inline FastDiv make_fast_div(long long d) {
  FastDiv f{0u, 0};
  if (d <= 0 || d > 0x7fffffffll) return f;  // never used for such a divisor
  while ((1ll << f.l) < d) ++f.l;
  f.m = (unsigned)(((1ull << 32) * ((1ull << f.l) - (unsigned long long)d)) / (unsigned long long)d + 1ull);
  return f;
}
__device__ __forceinline__ int fast_div(int n, FastDiv f) {
  return (int)((__umulhi(f.m, (unsigned)n) + (unsigned)n) >> f.l);
}

// A chunk of one fluid region's cells in the region-sorted cell list.
struct CompChunk {
  int comp, begin, end, pad;
};

// Validated box geometry and operator constants, passed by value.
struct DevGrid {
  int nx, ny, nz;
  // Nonzero when every face and cell-centre position of the lattice maps
  // back to its exact lattice coordinate: (o + (p + s) h - o) / h == p + s
  // for every p in 0..n and s in {0, 1/2} on each axis (true for a
  // power-of-two h and a modest origin).  Sampling at such a position has
  // weights in {0, 1/2, 1} only.
  int lattice;
  double h[3];
  double origin[3];
  double upper[3];
  double volume;
  double area[3];  // face areas (src/operator.rs:63)
  double acc[3];   // 1/(rho h_d)   (src/operator.rs:64)
  double w[3];     // area*acc      (src/operator.rs:65)
  double diag[64]; // cached boundary diagonals by neighbour mask (src/operator.rs:81-91)
  double inv_h[3]; // 1/h, used only where h is a power of two (exact)
  int pow2[3];     // h is a power of two: x/h == x*inv_h bit for bit
  int gauge;       // eliminated gauge cell (0 in the sealed box)
  // Optional per-cell open-face mask for solid obstacles: bits 0..5 = faces
  // x-, x+, y-, y+, z-, z+ open (the neighbour bits of the cached diagonal
  // table), bit 6 = fluid cell, bit 7 = the gauge of a fluid region.
  // nullptr = the plain box.
  const unsigned char* mask;
  // Divisors of the four-cell chunk loops: nx / 4 chunks per row, ny rows
  // per plane; of the flat face loops: nx and nx + 1 faces per row, ny + 1
  // rows per plane (set by make_dev_grid).
  FastDiv div_q, div_y;
  FastDiv div_x, div_x1, div_y1;
  // Obstacles that split the fluid into ncomp > 1 regions: one gauge per
  // region (kGaugeBit in the mask), the region label of every cell (255 =
  // solid), each region's gauge cell, a slot per region for the pressure
  // captured at its gauge, and the region-sorted fluid cell list in chunks
  // for the per-region compatibility sums.  ncomp <= 1: all unused.
  int ncomp, comp_nchunks;
  const unsigned char* comp_label;
  const int* comp_gauge;
  double* comp_shift;
  const int* comp_cells;
  const CompChunk* comp_chunks;
  const int* comp_chunk_first;  // ncomp + 1 chunk offsets
};

constexpr unsigned char kFluidBit = 64;
constexpr unsigned char kGaugeBit = 128;  // the gauge cell of a fluid region


// Every byte offset into each face array of g (elements of `elem` bytes)
// fits in 32 bits, so kernels may address them from a uniform base; and
// the flat index (i + fx (j + fy k)) of every face can be formed with
// 24-bit multiplies (fx < 2^24 and fy fz <= 2^24 for every face lattice).
inline bool face_offsets_32(const DevGrid& g, size_t elem) {
  const unsigned long long x = g.nx, y = g.ny, z = g.nz;
  unsigned long long m = (x + 1) * y * z;
  if (x * (y + 1) * z > m) m = x * (y + 1) * z;
  if (x * y * (z + 1) > m) m = x * y * (z + 1);
  const unsigned long long yz = (y + 1) * (z + 1);
  return m * elem <= (1ull << 32) && x + 1 < (1ull << 24) && yz <= (1ull << 24);
}

// The eliminated rows: the gauge cell, and with a split fluid every region's
// gauge (src/pressure.rs:534-541; src/obstacle_pressure.rs:385-388).
__device__ __forceinline__ bool is_gauge(const DevGrid& g, int c) {
  return c == g.gauge || (g.ncomp > 1 && (g.mask[c] & kGaugeBit));
}

// Interior face of axis a with face lattice index (i, j, k): its upper cell
// is (i, j, k), and the face is open when both cells are fluid (the lower-
// face bit of the upper cell).  Always open in the plain box.
__device__ __forceinline__ bool face_open(const DevGrid& g, int a, int i, int j, int k) {
  if (!g.mask) return true;
  return (g.mask[i + g.nx * (j + g.ny * k)] >> (2 * a)) & 1;
}

// Double-double accumulator.  Summing rounded products with error-free
// TwoSum gives (almost always) the correctly rounded sum, which is also what
// the reference's serial Neumaier loop returns (src/pressure.rs:135-161).
struct DD {
  double hi, lo;
};

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dd_acc(DD& x, double v) {
  double s, e;
  two_sum(x.hi, v, s, e);
  x.hi = s;
  x.lo += e;
}
__device__ __forceinline__ DD dd_add(DD x, DD y) {
  double s, e;
  two_sum(x.hi, y.hi, s, e);
  e += x.lo + y.lo;
  double h = s + e;
  return DD{h, e - (h - s)};
}
__device__ __forceinline__ double dd_value(DD x) { return x.hi + x.lo; }

__device__ __forceinline__ DD shfl_xor_dd(DD v, int m) {
  return DD{__shfl_xor(v.hi, m, 64), __shfl_xor(v.lo, m, 64)};
}

// Reduce NV double-double values across the block (fixed order: butterfly in
// the wave, then waves 0..3 in order).  Result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_reduce_dd(DD (&v)[NV]) {
  __shared__ DD lds[NV][kRowsPerBlock];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = dd_add(v[q], shfl_xor_dd(v[q], m));
  const int lane = threadIdx.x, wave = threadIdx.y;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[q][wave] = v[q];
  __syncthreads();
  if (lane == 0 && wave == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      DD a = lds[q][0];
      for (int w = 1; w < kRowsPerBlock; ++w) a = dd_add(a, lds[q][w]);
      v[q] = a;
    }
  }
  __syncthreads();
}

template <int NV>
__device__ __forceinline__ void block_reduce_max(double (&v)[NV]) {
  __shared__ double lds[NV][kRowsPerBlock];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = fmax(v[q], __shfl_xor(v[q], m, 64));
  const int lane = threadIdx.x, wave = threadIdx.y;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[q][wave] = v[q];
  __syncthreads();
  if (lane == 0 && wave == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double a = lds[q][0];
      for (int w = 1; w < kRowsPerBlock; ++w) a = fmax(a, lds[q][w]);
      v[q] = a;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ int flat_tid() { return threadIdx.y * kLanesX + threadIdx.x; }

// Grid reductions are two kernels: every block of the producing kernel
// publishes its partials with plain stores (block_publish), and a one-block
// finisher kernel (k_finish) combines them in block-index order -- a fixed
// order, so results are deterministic for a launch shape -- and runs the
// decision logic `fin(sums, maxima)` in thread 0.  (A single-kernel
// "last block" scheme needs a device-scope release fence per block, which on
// gfx950 writes back and invalidates the whole L2 and destroys stencil reuse.)
// ND double-double sums and NM maxima (both >= 1; pass dummies if unused).
// A kernel body runs as block b of nb: the hardware block of a one-scene
// launch, or a block of one scene's share of a batched launch.
struct VBlock {
  int b, nb;
};
__device__ __forceinline__ VBlock hw_block() { return VBlock{(int)blockIdx.x, (int)gridDim.x}; }

template <int ND, int NM>
__device__ __forceinline__ void block_publish(DD (&dd)[ND], double (&mx)[NM], DD* __restrict__ pdd,
                                              double* __restrict__ pmx, VBlock vb) {
  block_reduce_dd<ND>(dd);
  block_reduce_max<NM>(mx);
  if (flat_tid() == 0) {
#pragma unroll
    for (int q = 0; q < ND; ++q) pdd[q * kMaxBlocks + vb.b] = dd[q];
#pragma unroll
    for (int q = 0; q < NM; ++q) pmx[q * kMaxBlocks + vb.b] = mx[q];
  }
}
template <int ND, int NM>
__device__ __forceinline__ void block_publish(DD (&dd)[ND], double (&mx)[NM], DD* __restrict__ pdd,
                                              double* __restrict__ pmx) {
  block_publish<ND, NM>(dd, mx, pdd, pmx, hw_block());
}

// `Fin` provides `bool active() const` (false when the producer returned
// early without publishing) and `void operator()(const double* sums,
// const double* maxima) const`, both __device__.
// kU: partials loaded per round (0: the largest that fits, below); a
// finisher run inside another kernel takes a small one to keep that
// kernel's register budget.
template <int ND, int NM, typename Fin, int kU = 0>
__device__ __forceinline__ void finish_body(const DD* __restrict__ pdd, const double* __restrict__ pmx,
                                            int nb, const Fin& fin) {
  if (!fin.active()) return;
  const int t = flat_tid();
  DD a[ND];
  double m[NM];
#pragma unroll
  for (int q = 0; q < ND; ++q) a[q] = DD{0.0, 0.0};
#pragma unroll
  for (int q = 0; q < NM; ++q) m[q] = 0.0;
  // kUnroll partials per thread are loaded before any is added, so their
  // loads are in flight together (the partials come from other XCDs' L2s):
  // 32 (8192 partials in one round) when they fit 256 registers, else 16.
  // Each thread adds the partials b = t (mod kBlock) in ascending order
  // whatever kUnroll is, so the sums do not depend on it.
  constexpr int kUnroll = kU > 0 ? kU : ((4 * ND + 2 * NM) * 32 <= 256 ? 32 : 16);
  for (int b0 = t; b0 < nb; b0 += kUnroll * kBlock) {
    DD v[kUnroll][ND];
    double w[kUnroll][NM];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int b = b0 + u * kBlock;
      const bool in = b < nb;
#pragma unroll
      for (int q = 0; q < ND; ++q) v[u][q] = in ? pdd[q * kMaxBlocks + b] : DD{0.0, 0.0};
#pragma unroll
      for (int q = 0; q < NM; ++q) w[u][q] = in ? pmx[q * kMaxBlocks + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int q = 0; q < ND; ++q) a[q] = dd_add(a[q], v[u][q]);
#pragma unroll
      for (int q = 0; q < NM; ++q) m[q] = fmax(m[q], w[u][q]);
    }
  }
  block_reduce_dd<ND>(a);
  block_reduce_max<NM>(m);
  if (t == 0) {
    double sums[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) sums[q] = dd_value(a[q]);
    fin(sums, m);
  }
}

template <int ND, int NM, typename Fin>
__global__ void __launch_bounds__(kBlock)
k_finish(const DD* __restrict__ pdd, const double* __restrict__ pmx, int nb, Fin fin) {
  finish_body<ND, NM, Fin>(pdd, pmx, nb, fin);
}

template <int ND, int NM, typename Fin>
inline void launch_finish(const DD* pdd, const double* pmx, int nb, const Fin& fin, hipStream_t s) {
  k_finish<ND, NM, Fin><<<1, dim3(kLanesX, kRowsPerBlock), 0, s>>>(pdd, pmx, nb, fin);
}

// Writes sums/maxima to an output array (stateless reductions).
struct FinStore {
  double* out;
  int mode;  // 0: out[0] = max0; 1: velocity rate; 2: out[0..2] = sum0, max0, max1;
             // 3: out[0] = max0 and the velocity rate of max1..3 into out[1..4],
             //    out[5] = sum0; 4: out[0] = sum0
  double h0, h1, h2;
  __device__ bool active() const { return true; }
  __device__ void operator()(const double* sums, const double* mx) const {
    if (mode == 0) {
      out[0] = mx[0];
    } else if (mode == 1) {
      double rate = 0.0;
      rate += mx[0] / h0;
      rate += mx[1] / h1;
      rate += mx[2] / h2;
      out[0] = rate;
      out[1] = mx[0];
      out[2] = mx[1];
      out[3] = mx[2];
    } else if (mode == 4) {
      out[0] = sums[0];
    } else if (mode == 3) {
      out[0] = mx[0];
      FinStore{out + 1, 1, h0, h1, h2}(sums, mx + 1);
      out[5] = sums[0];
    } else {
      out[0] = sums[0];
      out[1] = mx[0];
      out[2] = mx[1];
    }
  }
};

// Iterate the cells of rows [row_begin, row_end) of an (nx, ny*nz) lattice.
// Each wave owns one row at a time; lanes stride along x.
struct RowRange {
  int begin, end;
};

// Balanced contiguous partition of `units` work units over the nb blocks
// in block order: block b owns [b units / nb, (b + 1) units / nb).
__device__ __forceinline__ RowRange block_span(int units, VBlock vb) {
  RowRange r;
  r.begin = (int)((long long)vb.b * units / vb.nb);
  r.end = (int)((long long)(vb.b + 1) * units / vb.nb);
  return r;
}

// A copy of a record that no kernel writes (a batch's per-scene grid
// record), read through the constant address space so that its fields are
// loaded into scalar registers once per wave; read through a generic
// pointer after stores, the compiler keeps them in vector registers (the
// batched advection then ran at 2 or 3 waves per SIMD against 3 or 4 for
// its one-scene form).  Only for kernels that do not index the record's
// tables per lane: such an index would put the copy in scratch memory.
template <typename R>
__device__ __forceinline__ R load_record(const R* p) {
  static_assert(sizeof(R) % 8 == 0 && alignof(R) >= 8, "records are copied in 8-byte words");
  typedef const __attribute__((address_space(4))) unsigned long long* CWords;
  const CWords src = (CWords)p;
  R r;
  unsigned long long* dst = reinterpret_cast<unsigned long long*>(&r);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(R) / 8); ++i) dst[i] = src[i];
  return r;
}

// Contiguous partition of `units` into nb chunks of a whole number of
// kBlock units (the last chunk short, trailing blocks possibly empty), so
// that every pass of a block's threads is full except in the last chunk.
__device__ __forceinline__ RowRange chunk_span(int units, VBlock vb) {
  const long long per = ((long long)units + vb.nb - 1) / vb.nb;
  const long long chunk = (per + kBlock - 1) / kBlock * kBlock;
  const long long b = (long long)vb.b * chunk, e = b + chunk;
  RowRange r;
  r.begin = (int)(b < units ? b : units);
  r.end = (int)(e < units ? e : units);
  return r;
}

// Row range of block vb: a balanced share of the kRowsPerBlock-row groups
// (kGroupsPerBlock groups per block when nb = grid_blocks(rows)).
__device__ __forceinline__ RowRange block_rows(int rows, VBlock vb) {
  const int groups = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
  const RowRange g = block_span(groups, vb);
  RowRange r;
  r.begin = g.begin * kRowsPerBlock;
  r.end = g.end * kRowsPerBlock;
  if (r.end > rows) r.end = rows;
  if (r.begin > r.end) r.begin = r.end;
  return r;
}
__device__ __forceinline__ RowRange block_rows(int rows) { return block_rows(rows, hw_block()); }

// Scene of block b in a batched launch: prefix[s] <= b < prefix[s + 1]
// (prefix has n + 1 entries; scenes with no blocks are skipped).
__device__ __forceinline__ int batch_scene(const int* __restrict__ prefix, int n, int b) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ VBlock batch_block(const int* __restrict__ prefix, int s) {
  return VBlock{(int)blockIdx.x - prefix[s], prefix[s + 1] - prefix[s]};
}

// Host helpers ---------------------------------------------------------------
int grid_blocks(int rows);               // blocks of a row launch (kRowsPerBlock rows each)
int quad_blocks(int quads_per_row, int rows);  // blocks of a four-cell-per-lane launch
int level_blocks(long long cells);             // blocks of a flat one-cell-per-lane launch
DevGrid make_dev_grid(const rh_grid& g, double density);
rh_status validate_grid(const rh_grid* g, size_t* cells_out);

}  // namespace rh

#define RH_HIP(call)                                  \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) {                           \
      rh::set_last_error(hipGetErrorString(e_), #call); \
      return RH_DEVICE_ERROR;                         \
    }                                                 \
  } while (0)

namespace rh {
void set_last_error(const char* what, const char* where);
}
