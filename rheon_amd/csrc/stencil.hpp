// Cell-lattice iteration and the 7-point stencil helpers shared by the PCG
// and multigrid kernels (box fast path of the reference operator).
#pragma once
#include "common.hpp"

namespace rh {

// Loop over this block's cells; defines (row, k, j, i, c) for the body.
#define FOR_CELLS(g)                                                                 \
  {                                                                                  \
    const RowRange rr_ = block_rows((g).ny * (g).nz);                                \
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

// apply_full box fast path (src/operator.rs:400-416): axis 0..2, lower then upper.
template <typename T>
__device__ __forceinline__ double apply_at(const DevGrid& g, const T* __restrict__ x, int c, int i,
                                           int j, int k) {
  const double xc = ld(x, c);
  const int sy = g.nx, sz = g.nx * g.ny;
  double v = 0.0;
  if (i > 0) v += g.w[0] * (xc - ld(x, c - 1));
  if (i + 1 < g.nx) v += g.w[0] * (xc - ld(x, c + 1));
  if (j > 0) v += g.w[1] * (xc - ld(x, c - sy));
  if (j + 1 < g.ny) v += g.w[1] * (xc - ld(x, c + sy));
  if (k > 0) v += g.w[2] * (xc - ld(x, c - sz));
  if (k + 1 < g.nz) v += g.w[2] * (xc - ld(x, c + sz));
  return v;
}

// diagonal(): cached sum of existing neighbour weights (src/operator.rs:813-822).
__device__ __forceinline__ double diag_at(const DevGrid& g, int i, int j, int k) {
  const int mask = (i > 0) | ((i + 1 < g.nx) << 1) | ((j > 0) << 2) | ((j + 1 < g.ny) << 3) |
                   ((k > 0) << 4) | ((k + 1 < g.nz) << 5);
  return g.diag[mask];
}

// Weighted neighbour sum sum_nbr w_d * x_nbr, axis 0..2, lower then upper.
template <typename T>
__device__ __forceinline__ double nbr_sum(const DevGrid& g, const T* __restrict__ x, double init,
                                          int c, int i, int j, int k) {
  const int sy = g.nx, sz = g.nx * g.ny;
  double v = init;
  if (i > 0) v += g.w[0] * ld(x, c - 1);
  if (i + 1 < g.nx) v += g.w[0] * ld(x, c + 1);
  if (j > 0) v += g.w[1] * ld(x, c - sy);
  if (j + 1 < g.ny) v += g.w[1] * ld(x, c + sy);
  if (k > 0) v += g.w[2] * ld(x, c - sz);
  if (k + 1 < g.nz) v += g.w[2] * ld(x, c + sz);
  return v;
}

static inline dim3 cells_grid(const DevGrid& g) { return dim3(grid_blocks(g.ny * g.nz)); }
static inline dim3 block_shape() { return dim3(kLanesX, kRowsPerBlock); }

}  // namespace rh
