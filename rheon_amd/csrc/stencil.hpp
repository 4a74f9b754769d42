// Cell-lattice iteration and the 7-point stencil helpers shared by the PCG
// and multigrid kernels (box operator of the reference, optionally with
// solid obstacles through DevGrid::mask).
#pragma once
#include "common.hpp"

namespace rh {

// Loop over this block's cells; defines (row, k, j, i, c) for the body.
#define FOR_CELLS(g) FOR_CELLS_VB(g, hw_block())
#define FOR_CELLS_VB(g, VB)                                                          \
  {                                                                                  \
    const RowRange rr_ = block_rows((g).ny * (g).nz, (VB));                          \
    for (int row = rr_.begin + threadIdx.y; row < rr_.end; row += kRowsPerBlock) {   \
      const int k = row / (g).ny, j = row - k * (g).ny;                              \
      (void)k;                                                                       \
      (void)j;                                                                       \
      for (int i = threadIdx.x; i < (g).nx; i += kLanesX) {                          \
        const int c = i + (g).nx * row;
#define END_CELLS \
  }               \
  }               \
  }

template <typename T>
__device__ __forceinline__ double ld(const T* __restrict__ x, int i) {
  return (double)x[i];
}

// Open-face bits of a cell in the order of the cached diagonal table
// (src/operator.rs:816-820): x-, x+, y-, y+, z-, z+.  In the plain box a face
// is open when the neighbour exists; with obstacles it comes from the mask.
__device__ __forceinline__ int box_mask6(int nx, int ny, int nz, int i, int j, int k) {
  return (i > 0) | ((i + 1 < nx) << 1) | ((j > 0) << 2) | ((j + 1 < ny) << 3) | ((k > 0) << 4) |
         ((k + 1 < nz) << 5);
}
__device__ __forceinline__ int cell_mask6(const DevGrid& g, int c, int i, int j, int k) {
  return g.mask ? (g.mask[c] & 63) : box_mask6(g.nx, g.ny, g.nz, i, j, k);
}
__device__ __forceinline__ bool is_fluid(const DevGrid& g, int c) {
  return !g.mask || (g.mask[c] & kFluidBit);
}

// Neighbour offsets across open faces, 0 across closed ones.  Branch-free
// stencils load every neighbour at once (a clamped address is the cell
// itself): a closed face then contributes w*(x_c - x_c) = +0.0 to the
// running sum, which leaves every partial sum bit-identical to the
// reference's skipped term (the sum starts at +0.0 and never becomes -0.0),
// so the operation order of src/operator.rs:400-416 is kept.
struct Nbr {
  int xm, xp, ym, yp, zm, zp;
};
__device__ __forceinline__ Nbr nbr_from_mask(int m, int sy, int sz) {
  Nbr n;
  n.xm = (m & 1) ? -1 : 0;
  n.xp = (m & 2) ? 1 : 0;
  n.ym = (m & 4) ? -sy : 0;
  n.yp = (m & 8) ? sy : 0;
  n.zm = (m & 16) ? -sz : 0;
  n.zp = (m & 32) ? sz : 0;
  return n;
}
__device__ __forceinline__ Nbr nbr_offsets(int nx, int ny, int nz, int i, int j, int k) {
  return nbr_from_mask(box_mask6(nx, ny, nz, i, j, k), nx, nx * ny);
}

// apply_full (src/operator.rs:400-416): axis 0..2, lower then upper.
// The six neighbours are loaded from box-clamped addresses, independent of
// the obstacle mask (one round of loads in flight); a closed obstacle face
// then selects x_c itself, so its term is the exact +0.0 of the box wall.
template <typename T>
__device__ __forceinline__ double apply_at(const DevGrid& g, const T* __restrict__ x, int c, int i,
                                           int j, int k) {
  const Nbr n = nbr_offsets(g.nx, g.ny, g.nz, i, j, k);
  const double xc = ld(x, c);
  double a = ld(x, c + n.xm), b = ld(x, c + n.xp), cy = ld(x, c + n.ym), d = ld(x, c + n.yp),
         e = ld(x, c + n.zm), f = ld(x, c + n.zp);
  if (g.mask) {
    const int m = g.mask[c];
    a = (m & 1) ? a : xc;
    b = (m & 2) ? b : xc;
    cy = (m & 4) ? cy : xc;
    d = (m & 8) ? d : xc;
    e = (m & 16) ? e : xc;
    f = (m & 32) ? f : xc;
  }
  double v = 0.0;
  v += g.w[0] * (xc - a);
  v += g.w[0] * (xc - b);
  v += g.w[1] * (xc - cy);
  v += g.w[1] * (xc - d);
  v += g.w[2] * (xc - e);
  v += g.w[2] * (xc - f);
  return v;
}

// diagonal(): cached sum of open-face weights (src/operator.rs:813-822); 0
// for a solid cell.
__device__ __forceinline__ double diag_at(const DevGrid& g, int c, int i, int j, int k) {
  return g.diag[cell_mask6(g, c, i, j, k)];
}

// Weighted neighbour sum sum_nbr w_d * x_nbr, axis 0..2, lower then upper.
// A closed face contributes w * (+0.0) = +0.0 (same argument as above; the
// red-black kernels' `init` is never -0.0 unless the reference's is).
template <typename T>
__device__ __forceinline__ double nbr_sum(const DevGrid& g, const T* __restrict__ x, double init,
                                          int c, int i, int j, int k) {
  const Nbr n = nbr_offsets(g.nx, g.ny, g.nz, i, j, k);
  // unconditional loads (box-clamped addresses), then select
  const double la = ld(x, c + n.xm), lb = ld(x, c + n.xp), lc = ld(x, c + n.ym),
               ldd = ld(x, c + n.yp), le = ld(x, c + n.zm), lf = ld(x, c + n.zp);
  const int m = cell_mask6(g, c, i, j, k);
  const double a = (m & 1) ? la : 0.0;
  const double b = (m & 2) ? lb : 0.0;
  const double cy = (m & 4) ? lc : 0.0;
  const double d = (m & 8) ? ldd : 0.0;
  const double e = (m & 16) ? le : 0.0;
  const double f = (m & 32) ? lf : 0.0;
  double v = init;
  v += g.w[0] * a;
  v += g.w[0] * b;
  v += g.w[1] * cy;
  v += g.w[1] * d;
  v += g.w[2] * e;
  v += g.w[2] * f;
  return v;
}

static inline dim3 cells_grid(const DevGrid& g) { return dim3(grid_blocks(g.ny * g.nz)); }
static inline dim3 block_shape() { return dim3(kLanesX, kRowsPerBlock); }

}  // namespace rh
