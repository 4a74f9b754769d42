// Preconditioned conjugate gradient on the device: the reference's
// "jacobi-pcg-v1" (src/pressure.rs:577-740) with every scalar decision taken
// on the device, plus a symmetric red-black Gauss-Seidel preconditioner.
//
// One iteration `it` is a fixed sequence of kernels; each begins by reading
// the control block and returns at once when the solve has finished or the
// kernel's branch is not taken, so the host can enqueue many iterations
// ahead and poll `done` only every few iterations.  A kernel with a dot
// product publishes per-block partials and is followed by a one-block
// finisher (k_finish) that combines them in fixed order and takes the
// decision:
//   apply_dir   q = K d, q[0] = 0, curvature = d.q, alpha        (pressure.rs:657-668)
//   update      p += alpha d, r -= alpha q (i >= 1), r.r, z, r.z (:669-681, :534-575)
//   [sgs black, sgs red]                                        (RB-SGS only)
//   true        if candidate: product = b - K p, l2, max, gates  (:686-709, :432-456)
//   replace     if replaced: r = product, z, r.z                 (:705-707, :716)
//   [sgs black, sgs red]                                        (RB-SGS replace path)
//   direction   d = z + beta d (or d = z after a replacement)    (:716-733)
// Dot products are double-double sums of the rounded products (see DD in
// common.hpp), so they equal the reference's Neumaier sums except in
// vanishingly rare midpoint cases.  Vector storage T is double (reference
// arithmetic) or float (halved traffic; scalars and arithmetic stay f64).
#include "common.hpp"
#include "kernels.hpp"
#include "batch.hpp"
#include "profiler.hpp"
#include "stencil.hpp"

namespace rh {

// z = r / D; a solid cell (D = 0, r = 0) gets z = 0.  For every fluid cell
// this is the reference's division (src/pressure.rs:567).
__device__ __forceinline__ double jdiv(double r, double d) { return d > 0.0 ? r / d : 0.0; }


// Which path a red-black stage belongs to: the speculative one right after
// `update` (0) or the replacement / initial one after `replace` (1).
__device__ __forceinline__ bool sgs_active(const PcgControl* ctl, int it, int mode) {
  if (ctl->done) return false;
  if (mode == 0) return it > 0;
  return it == 0 || ctl->replaced;
}

// -------------------------------------------------------------- finishers


// curvature checks and alpha (src/pressure.rs:661-668).
struct FinApplyDir {
  PcgControl* ctl;
  int it;
  __device__ bool active() const { return !ctl->done; }
  __device__ void operator()(const double* sum, const double* mx) const {
    if (mx[0] != 0.0) return fail(ctl, RH_ARITHMETIC_FAILURE, it);
    const double curvature = sum[0];
    const double rz = ctl->rz[it & 1];
    if (!isfinite(curvature) || curvature <= 0.0 || !isfinite(rz) || rz <= 0.0)
      return fail(ctl, RH_BREAKDOWN, it);
    const double alpha = rz / curvature;
    if (!isfinite(alpha)) return fail(ctl, RH_BREAKDOWN, it);
    ctl->alpha = alpha;
  }
};

// recursive residual and candidate test (src/pressure.rs:680-686).
struct FinUpdate {
  PcgControl* ctl;
  int it;
  int jacobi;
  __device__ bool active() const { return !ctl->done; }
  __device__ void operator()(const double* sum, const double*) const {
    ctl->rr = sum[0];
    const double recursive_l2 = sqrt(sum[0]);
    ctl->candidate = (recursive_l2 <= ctl->threshold) || it == ctl->max_iterations;
    ctl->replaced = 0;
    if (jacobi) ctl->rz[(it + 1) & 1] = sum[1];
    ctl->sum_r = sum[2];
  }
};

// true_report and the decisions of solve_loaded (src/pressure.rs:432-456,
// :638-647 at it = 0, :686-715 otherwise).
struct FinTrue {
  PcgControl* ctl;
  int it;
  __device__ bool active() const { return !ctl->done && (it == 0 || ctl->candidate); }
  __device__ void operator()(const double* sum, const double* mx) const {
    if (mx[1] != 0.0) return fail(ctl, RH_ARITHMETIC_FAILURE, it);
    const double squared = sum[0], maximum = mx[0];
    const double divergence = maximum * ctl->dt / ctl->volume;
    if (!isfinite(squared) || !isfinite(divergence)) return fail(ctl, RH_BREAKDOWN, it);
    const double l2 = sqrt(squared);
    ctl->true_l2 = l2;
    ctl->true_max = maximum;
    ctl->pred_div = divergence;
    const bool accepted = l2 <= ctl->threshold && divergence <= ctl->div_limit;
    if (accepted) {
      ctl->iteration = it;
      ctl->done = 1;
      ctl->status = RH_OK;
      return;
    }
    // a seeded solve replaces on the residual without the eliminated gauge
    // rows (src/pressure.rs:689-707)
    const double replacement_l2 = ctl->seeded ? sqrt(sum[1]) : l2;
    if (it > 0 && replacement_l2 > ctl->threshold) ctl->replaced = 1;
    if (it == ctl->max_iterations) {
      ctl->error_value = maximum;
      ctl->iteration = it;
      return fail(ctl, RH_ITERATION_LIMIT, it);
    }
  }
};

// next rz after the (initial or replacement) Jacobi preconditioning.
struct FinReplace {
  PcgControl* ctl;
  int it;
  __device__ bool active() const { return !ctl->done && (it == 0 || ctl->replaced); }
  __device__ void operator()(const double* sum, const double*) const { ctl->rz[(it + 1) & 1] = sum[0]; }
};

// Jacobi PCG, iterations >= 1, with the replacement folded into the next
// iteration's apply (the launch sequence is apply -> finish -> update ->
// finish -> direction + true residual -> finish): when the previous
// iteration was replaced, its new rz (partial r.z published beside d.q) is
// stored and breakdown-checked for that iteration first, as the separate
// replace, finish and direction kernels did (src/pressure.rs:705-719).
struct FinApplyDirJ {
  PcgControl* ctl;
  int it;
  __device__ bool active() const { return !ctl->done; }
  __device__ void operator()(const double* sum, const double* mx) const {
    if (ctl->replaced) {
      const double next_rz = sum[1];
      ctl->rz[it & 1] = next_rz;
      if (!isfinite(next_rz) || next_rz <= 0.0) return fail(ctl, RH_BREAKDOWN, it - 1);
      ctl->iteration = it - 1;
    }
    FinApplyDir{ctl, it}(sum, mx);
  }
};

// FinTrue for the folded sequence: the direction kernel ran before this
// decision, so an iteration that is neither accepted, replaced nor at the
// limit takes the next-rz breakdown check here (replaced ones take it in
// FinApplyDirJ on their replacement rz).
struct FinTrueJ {
  PcgControl* ctl;
  int it;
  __device__ bool active() const { return !ctl->done && ctl->candidate; }
  __device__ void operator()(const double* sum, const double* mx) const {
    FinTrue{ctl, it}(sum, mx);
    if (ctl->done || ctl->replaced) return;
    const double next_rz = ctl->rz[(it + 1) & 1];
    if (!isfinite(next_rz) || next_rz <= 0.0) return fail(ctl, RH_BREAKDOWN, it);
    ctl->iteration = it;
  }
};

// next rz after a red-black preconditioner application.
struct FinSgsRed {
  PcgControl* ctl;
  int it;
  int mode;
  __device__ bool active() const { return sgs_active(ctl, it, mode); }
  __device__ void operator()(const double* sum, const double*) const { ctl->rz[(it + 1) & 1] = sum[0]; }
};

// ---------------------------------------------------------------- kernels

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill_zero(T* __restrict__ x, int n) {
  for (int i = blockIdx.x * kBlock + flat_tid(); i < n; i += gridDim.x * kBlock) x[i] = (T)0.0;
}

// sum(b), sum(|b|), b.b and validate_cells (finite, src/operator.rs:840-848).
__device__ __forceinline__ void b_rhs_sums(VBlock vb, const DevGrid& g, const double* __restrict__ b,
                                           DD* pdd, double* pmx) {
  DD s[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  double bad[1] = {0.0};
  FOR_CELLS_VB(g, vb)
  const double v = b[c];
  if (!isfinite(v)) bad[0] = 1.0;
  dd_acc(s[0], v);
  dd_acc(s[1], fabs(v));
  dd_acc(s[2], v * v);
  END_CELLS
  block_publish<3, 1>(s, bad, pdd, pmx, vb);
}
__global__ void __launch_bounds__(kBlock)
k_rhs_sums(DevGrid g, const double* __restrict__ b, DD* pdd, double* pmx) {
  b_rhs_sums(hw_block(), g, b, pdd, pmx);
}

// A caller-supplied pressure guess (src/pressure.rs:615-623): p = seed,
// every value finite (OperatorError::NonFiniteInput) and 0 at the gauge
// (PressureError::InvalidInitialPressure; with a split fluid at every
// region's gauge).  One pass copies and flags; the finisher decides.
__global__ void __launch_bounds__(kBlock)
k_seed_copy(DevGrid g, const double* seed, double* p, const PcgControl* ctl, DD* pdd,  // seed may be p
            double* pmx) {
  if (ctl->done) return;
  DD none[1] = {{0.0, 0.0}};
  double bad[1] = {0.0};
  FOR_CELLS(g)
  const double v = seed[c];
  if (!isfinite(v)) bad[0] = 1.0;
  p[c] = v;
  END_CELLS
  block_publish<1, 1>(none, bad, pdd, pmx);
}
struct FinSeed {
  PcgControl* ctl;
  const double* seed;
  int gauge, ncomp;
  const int* comp_gauge;
  __device__ bool active() const { return !ctl->done; }
  __device__ void operator()(const double*, const double* mx) const {
    if (mx[0] != 0.0) return fail(ctl, RH_NONFINITE_INPUT, 0);
    bool zero = seed[gauge] == 0.0;
    if (ncomp > 1)
      for (int k = 0; k < ncomp; ++k) zero = zero && seed[comp_gauge[k]] == 0.0;
    if (!zero) return fail(ctl, RH_INVALID_INITIAL_PRESSURE, 0);
  }
};
void launch_seed(const DevGrid& g, const double* seed, double* p, PcgControl* ctl, DD* pdd, double* pmx,
                 hipStream_t s) {
  const dim3 grid = cells_grid(g);
  k_seed_copy<<<grid, block_shape(), 0, s>>>(g, seed, p, ctl, pdd, pmx);
  launch_finish<1, 1>(pdd, pmx, grid.x, FinSeed{ctl, seed, g.gauge, g.ncomp, g.comp_gauge}, s);
}

// Compatibility of a split fluid (src/obstacle_pressure.rs:455-472): every
// region's right-hand side sums to zero within 64 eps of its absolute sum.
// Chunk sums over the region-sorted cell list (one block per chunk, chunks
// never span two regions), then one block adds each region's chunk partials
// in chunk order and checks the regions in label order; the first failing
// region reports its sum.
__device__ __forceinline__ void b_comp_sums(const DevGrid& g, const double* __restrict__ b, VBlock vb, DD* pdd,
                                            double* pmx) {
  const CompChunk ch = g.comp_chunks[vb.b];
  DD s[2] = {{0.0, 0.0}, {0.0, 0.0}};
  double none[1] = {0.0};
  for (int t = ch.begin + flat_tid(); t < ch.end; t += kBlock) {
    const double v = b[g.comp_cells[t]];
    dd_acc(s[0], v);
    dd_acc(s[1], fabs(v));
  }
  block_publish<2, 1>(s, none, pdd, pmx, vb);
}
__device__ __forceinline__ void b_comp_check(const DevGrid& g, const DD* __restrict__ pdd, PcgControl* ctl) {
  __shared__ int stop;
  const int t = flat_tid();
  for (int k = 0; k < g.ncomp; ++k) {
    DD v[2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int q = g.comp_chunk_first[k] + t; q < g.comp_chunk_first[k + 1]; q += kBlock) {
      v[0] = dd_add(v[0], pdd[q]);
      v[1] = dd_add(v[1], pdd[kMaxBlocks + q]);
    }
    block_reduce_dd<2>(v);
    if (t == 0) {
      stop = 0;
      const double sum = dd_value(v[0]), sum_abs = dd_value(v[1]);
      if (!isfinite(sum) || !isfinite(sum_abs)) {
        fail(ctl, RH_BREAKDOWN, 0);
        stop = 1;
      } else if (fabs(sum) > (64.0 * 2.220446049250313e-16) * sum_abs) {
        ctl->error_value = sum;
        fail(ctl, RH_INCOMPATIBLE_RHS, 0);
        stop = 1;
      }
    }
    __syncthreads();
    if (stop) return;
    __syncthreads();
  }
}
__global__ void __launch_bounds__(kBlock)
k_comp_sums(DevGrid g, const double* __restrict__ b, const PcgControl* ctl, DD* pdd, double* pmx) {
  if (ctl->done) return;
  b_comp_sums(g, b, hw_block(), pdd, pmx);
}
__global__ void __launch_bounds__(kBlock) k_comp_check(DevGrid g, const DD* __restrict__ pdd, PcgControl* ctl) {
  if (ctl->done) return;
  b_comp_check(g, pdd, ctl);
}

// After the whole-domain compatibility sums: the per-region check of a
// split fluid (nothing for one region).
void launch_comp_check(const DevGrid& g, const double* b, PcgControl* ctl, DD* pdd, double* pmx, hipStream_t s) {
  if (g.ncomp <= 1) return;
  k_comp_sums<<<g.comp_nchunks, block_shape(), 0, s>>>(g, b, ctl, pdd, pmx);
  k_comp_check<<<1, block_shape(), 0, s>>>(g, pdd, ctl);
}

// q = K d, q[0] = 0 (gauge), partial d.q.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_apply_dir(DevGrid g, const T* __restrict__ d, T* __restrict__ q, const PcgControl* ctl, DD* pdd,
            double* pmx) {
  if (ctl->done) return;
  DD s[1] = {{0.0, 0.0}};
  double bad[1] = {0.0};
  FOR_CELLS(g)
  double v = apply_at<T>(g, d, c, i, j, k);
  if (!isfinite(v)) bad[0] = 1.0;
  if (is_gauge(g, c)) v = 0.0;
  const T qs = (T)v;
  q[c] = qs;
  dd_acc(s[0], ld(d, c) * (double)qs);
  END_CELLS
  block_publish<1, 1>(s, bad, pdd, pmx);
}

// p += alpha d, r -= alpha q for i >= 1; partial r.r and sum(r).
// JACOBI: z = r/D on the fly and partial r.z (speculative next rz).
// RB-SGS: first (red) half-sweep y_red = r/D stored in z.
template <typename T, bool SGS>
__global__ void __launch_bounds__(kBlock)
k_update(DevGrid g, double* __restrict__ p, T* __restrict__ r, const T* __restrict__ d,
         const T* __restrict__ q, T* __restrict__ z, const PcgControl* ctl, DD* pdd, double* pmx) {
  if (ctl->done) return;
  const double alpha = ctl->alpha;
  DD s[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  double none[1] = {0.0};
  FOR_CELLS(g)
  if (is_gauge(g, c)) {
    if (SGS) z[c] = (T)0.0;
    continue;
  }
  const double pn = p[c] + alpha * ld(d, c);  // the solution accumulates in f64
  const T rn = (T)(ld(r, c) - alpha * ld(q, c));
  p[c] = pn;
  r[c] = rn;
  const double rv = (double)rn;
  dd_acc(s[0], rv * rv);
  dd_acc(s[2], rv);
  if (SGS) {
    if (((i + j + k) & 1) == 0) z[c] = (T)jdiv(rv, diag_at(g, c, i, j, k));
  } else {
    const double zv = jdiv(rv, diag_at(g, c, i, j, k));
    dd_acc(s[1], rv * zv);
  }
  END_CELLS
  block_publish<3, 1>(s, none, pdd, pmx);
}

// product = b - K p over every cell (gauge row included): partial sum of
// squares, max|.|, nonfinite flag.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_true_residual(DevGrid g, const double* __restrict__ p, const double* __restrict__ b,
                T* __restrict__ out, const PcgControl* ctl, DD* pdd, double* pmx, int it) {
  if (ctl->done) return;
  if (it > 0 && !ctl->candidate) return;
  DD s[2] = {{0.0, 0.0}, {0.0, 0.0}};
  double m[2] = {0.0, 0.0};
  FOR_CELLS(g)
  const double v = apply_at<double>(g, p, c, i, j, k);
  if (!isfinite(v)) m[1] = 1.0;
  const double prod = b[c] - v;
  out[c] = (T)prod;
  dd_acc(s[0], prod * prod);
  if (!is_gauge(g, c)) dd_acc(s[1], prod * prod);  // without the gauge rows
  m[0] = fmax(m[0], fabs(prod));
  END_CELLS
  block_publish<2, 2>(s, m, pdd, pmx);
}

// Replacement r = product (r[0] = 0) and the first preconditioner stage.
// it == 0 is the initial precondition before the loop (src/pressure.rs:648-650).
// JACOBI: partial r.z with z = r/D.  RB-SGS: y_red = r/D into z.
template <typename T, bool SGS>
__global__ void __launch_bounds__(kBlock)
k_replace(DevGrid g, T* r, const T* prod, T* __restrict__ z, const PcgControl* ctl, DD* pdd,
          double* pmx, int it) {
  if (ctl->done) return;
  if (it > 0 && !ctl->replaced) return;
  DD s[1] = {{0.0, 0.0}};
  double none[1] = {0.0};
  FOR_CELLS(g)
  if (is_gauge(g, c)) {
    r[c] = (T)0.0;
    if (SGS) z[c] = (T)0.0;
    continue;
  }
  const T rn = prod[c];
  r[c] = rn;
  const double rv = (double)rn;
  if (SGS) {
    if (((i + j + k) & 1) == 0) z[c] = (T)jdiv(rv, diag_at(g, c, i, j, k));
  } else {
    const double zv = jdiv(rv, diag_at(g, c, i, j, k));
    dd_acc(s[0], rv * zv);
  }
  END_CELLS
  if (!SGS) block_publish<1, 1>(s, none, pdd, pmx);
}

// Black forward half-sweep: y_b = (r_b + sum_red w y) / D (neighbours are red).
// The black reverse step z_b = y_b is the identity (no later unknowns).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_sgs_black(DevGrid g, const T* __restrict__ r, T* __restrict__ z, const PcgControl* ctl, int it,
            int mode) {
  if (!sgs_active(ctl, it, mode)) return;
  FOR_CELLS(g)
  if (((i + j + k) & 1) == 0 || is_gauge(g, c)) continue;
  const double v = nbr_sum<T>(g, z, ld(r, c), c, i, j, k);
  z[c] = (T)jdiv(v, diag_at(g, c, i, j, k));
  END_CELLS
}

// Red reverse half-sweep z_r = y_r + (sum_black w z)/D, then partial r.z
// over all cells (black z are final).  Gauge cell 0 (red) stays 0.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_sgs_red(DevGrid g, const T* __restrict__ r, T* __restrict__ z, const PcgControl* ctl, DD* pdd,
          double* pmx, int it, int mode) {
  if (!sgs_active(ctl, it, mode)) return;
  DD s[1] = {{0.0, 0.0}};
  double none[1] = {0.0};
  FOR_CELLS(g)
  double zv;
  if (((i + j + k) & 1) == 0 && !is_gauge(g, c)) {
    const double corr = nbr_sum<T>(g, z, 0.0, c, i, j, k);
    const T zn = (T)(ld(z, c) + jdiv(corr, diag_at(g, c, i, j, k)));
    z[c] = zn;
    zv = (double)zn;
  } else {
    zv = ld(z, c);
  }
  dd_acc(s[0], ld(r, c) * zv);
  END_CELLS
  block_publish<1, 1>(s, none, pdd, pmx);
}

// ------------------------------------------------ lexicographic SGS (exact)
// precondition_sgs (src/pressure.rs:467-532, the sgs-pcg-v1 preconditioner):
// forward z_i = (r_i + sum_{d, lower} w_d z_{i - s_d}) / D_i in row order,
// then reverse z_i += (sum_{d, upper} w_d z_{i + s_d}) / D_i in reverse row
// order, the gauge row skipped (z = 0).  A cell's forward value depends only
// on its lower neighbours' final values, so any order that finishes them
// first gives the same bits: tiles of 8^3 cells run in waves bx + by + bz =
// w (one launch per wave, src/sgs.rs:56-128 uses the same tile waves on CPU
// threads), and inside a tile one wavefront walks its local planes
// li + lj + lk = l with the tile's values in LDS.  Box operator only.
template <typename T, bool kReverse>
__global__ void __launch_bounds__(64)
k_sgs_lex_wave(DevGrid g, const T* __restrict__ r, T* __restrict__ z, const SgsTile* __restrict__ tiles,
               const PcgControl* ctl, int it, int mode) {
  if (!sgs_active(ctl, it, mode)) return;
  constexpr int E = kSgsTileEdge;
  __shared__ T tz[E * E * E];
  const SgsTile t = tiles[blockIdx.x];
  const int x0 = t.bx * E, y0 = t.by * E, z0 = t.bz * E;
  const int ex = min(E, g.nx - x0), ey = min(E, g.ny - y0), ez = min(E, g.nz - z0);
  const int cells = ex * ey * ez, exy = ex * ey;
  const int sy = g.nx, sz = g.nx * g.ny;
  const int lane = threadIdx.x;
  if (kReverse)
    for (int q = lane; q < cells; q += 64) {
      const int li = q % ex, lj = (q / ex) % ey, lk = q / exy;
      tz[q] = z[(x0 + li) + sy * (y0 + lj) + sz * (z0 + lk)];
    }
  __syncthreads();
  const int planes = ex + ey + ez - 2;
  for (int step = 0; step < planes; ++step) {
    const int l = kReverse ? planes - 1 - step : step;
    for (int q = lane; q < cells; q += 64) {
      const int li = q % ex, lj = (q / ex) % ey, lk = q / exy;
      if (li + lj + lk != l) continue;
      const int i = x0 + li, j = y0 + lj, k = z0 + lk;
      const int c = i + sy * j + sz * k;
      if (is_gauge(g, c)) {
        tz[q] = (T)0.0;
        continue;
      }
      const double diag = diag_at(g, c, i, j, k);
      if (!kReverse) {
        double v = (double)r[c];
        if (i > 0) v += g.w[0] * (double)(li > 0 ? tz[q - 1] : z[c - 1]);
        if (j > 0) v += g.w[1] * (double)(lj > 0 ? tz[q - ex] : z[c - sy]);
        if (k > 0) v += g.w[2] * (double)(lk > 0 ? tz[q - exy] : z[c - sz]);
        tz[q] = (T)(v / diag);
      } else {
        double corr = 0.0;
        if (i + 1 < g.nx) corr += g.w[0] * (double)(li + 1 < ex ? tz[q + 1] : z[c + 1]);
        if (j + 1 < g.ny) corr += g.w[1] * (double)(lj + 1 < ey ? tz[q + ex] : z[c + sy]);
        if (k + 1 < g.nz) corr += g.w[2] * (double)(lk + 1 < ez ? tz[q + exy] : z[c + sz]);
        tz[q] = (T)((double)tz[q] + corr / diag);
      }
    }
    __syncthreads();
  }
  for (int q = lane; q < cells; q += 64) {
    const int li = q % ex, lj = (q / ex) % ey, lk = q / exy;
    z[(x0 + li) + sy * (y0 + lj) + sz * (z0 + lk)] = tz[q];
  }
}

// r.z after the lexicographic sweeps (the reference's preconditioned.dot
// (residual)) and a non-finite z flag (the sweeps' Breakdown checks).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_sgs_lex_dot(DevGrid g, const T* __restrict__ r, const T* __restrict__ z, const PcgControl* ctl, DD* pdd,
              double* pmx, int it, int mode) {
  if (!sgs_active(ctl, it, mode)) return;
  DD s[1] = {{0.0, 0.0}};
  double bad[1] = {0.0};
  FOR_CELLS(g)
  const double zv = ld(z, c);
  if (!isfinite(zv)) bad[0] = 1.0;
  dd_acc(s[0], zv * ld(r, c));
  END_CELLS
  block_publish<1, 1>(s, bad, pdd, pmx);
}

// next rz after the lexicographic SGS; a non-finite value is the
// preconditioner's Breakdown { iteration: 0 } (src/pressure.rs:493-498).
struct FinSgsLex {
  PcgControl* ctl;
  int it;
  int mode;
  __device__ bool active() const { return sgs_active(ctl, it, mode); }
  __device__ void operator()(const double* sum, const double* mx) const {
    if (mx[0] != 0.0) return fail(ctl, RH_BREAKDOWN, 0);
    ctl->rz[(it + 1) & 1] = sum[0];
  }
};

template <typename T>
void sgs_lex_apply(const DevGrid& g, const T* r, T* z, const SgsTile* tiles, const int* wave_start, int waves,
                   PcgControl* ctl, DD* pdd, double* pmx, int it, int mode, hipStream_t s) {
  for (int w = 0; w < waves; ++w) {
    const int n = wave_start[w + 1] - wave_start[w];
    if (n > 0) k_sgs_lex_wave<T, false><<<n, 64, 0, s>>>(g, r, z, tiles + wave_start[w], ctl, it, mode);
  }
  for (int w = waves - 1; w >= 0; --w) {
    const int n = wave_start[w + 1] - wave_start[w];
    if (n > 0) k_sgs_lex_wave<T, true><<<n, 64, 0, s>>>(g, r, z, tiles + wave_start[w], ctl, it, mode);
  }
  const dim3 grid = cells_grid(g);
  k_sgs_lex_dot<T><<<grid, block_shape(), 0, s>>>(g, r, z, ctl, pdd, pmx, it, mode);
  launch_finish<1, 1>(pdd, pmx, grid.x, FinSgsLex{ctl, it, mode}, s);
}

void sgs_plan(const DevGrid& g, std::vector<SgsTile>& tiles, std::vector<int>& wave_start) {
  const int E = kSgsTileEdge;
  const int tx = (g.nx + E - 1) / E, ty = (g.ny + E - 1) / E, tz = (g.nz + E - 1) / E;
  const int waves = tx + ty + tz - 2;
  tiles.clear();
  wave_start.assign(1, 0);
  for (int w = 0; w < waves; ++w) {
    for (int bx = 0; bx < tx; ++bx)
      for (int by = 0; by < ty; ++by) {
        const int bz = w - bx - by;
        if (bz >= 0 && bz < tz) tiles.push_back(SgsTile{bx, by, bz, 0});
      }
    wave_start.push_back((int)tiles.size());
  }
}

// d = z + beta d (i >= 1), or d = z after a replacement / at it = 0;
// next_rz breakdown check (src/pressure.rs:716-719).
template <typename T, bool SGS>
__global__ void __launch_bounds__(kBlock)
k_direction(DevGrid g, const T* __restrict__ r, const T* __restrict__ z, T* __restrict__ d,
            PcgControl* ctl, int it) {
  if (ctl->done) return;
  const double next_rz = ctl->rz[(it + 1) & 1];
  if (it > 0 && (!isfinite(next_rz) || next_rz <= 0.0)) {
    if (blockIdx.x == 0 && flat_tid() == 0) fail(ctl, RH_BREAKDOWN, it);
    return;
  }
  const bool copy = it == 0 || ctl->replaced;
  const double beta = copy ? 0.0 : next_rz / ctl->rz[it & 1];
  FOR_CELLS(g)
  if (is_gauge(g, c)) {
    d[c] = (T)0.0;
    continue;
  }
  const double zv = SGS ? ld(z, c) : jdiv(ld(r, c), diag_at(g, c, i, j, k));
  d[c] = copy ? (T)zv : (T)(zv + beta * ld(d, c));
  END_CELLS
  if (blockIdx.x == 0 && flat_tid() == 0) ctl->iteration = it;
}


// Jacobi PCG apply for iterations >= 1.  Not replaced: k_apply_dir.  After
// a replacement (ctl->replaced, set by the previous iteration's true
// residual held in `prod`): r = prod (r[gauge] = 0) and d = z = r / D are
// formed on the fly (the values k_replace and k_direction stored), q = K d,
// and the partial r.z of the new rz is published beside d.q.
template <typename T>
__device__ __forceinline__ T jdir_from(const DevGrid& g, const T* __restrict__ prod, int c, int i, int j, int k) {
  return is_gauge(g, c) ? (T)0.0 : (T)jdiv(ld(prod, c), diag_at(g, c, i, j, k));
}
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_apply_dir_j(DevGrid g, T* __restrict__ r, T* __restrict__ d, T* __restrict__ q, const T* __restrict__ prod,
              const PcgControl* ctl, DD* pdd, double* pmx) {
  if (ctl->done) return;
  const bool rep = ctl->replaced;
  DD s[2] = {{0.0, 0.0}, {0.0, 0.0}};
  double bad[1] = {0.0};
  FOR_CELLS(g)
  double v, dc;
  if (!rep) {
    v = apply_at<T>(g, d, c, i, j, k);
    dc = ld(d, c);
  } else {
    // apply_at over the virtual direction d = z(prod): box-clamped
    // neighbours (offset 0 = the cell itself), closed faces select d_c
    const Nbr n = nbr_offsets(g.nx, g.ny, g.nz, i, j, k);
    const T dT = jdir_from(g, prod, c, i, j, k);
    const double xc = (double)dT;
    double a = (double)jdir_from(g, prod, c + n.xm, i + n.xm, j, k);
    double b = (double)jdir_from(g, prod, c + n.xp, i + n.xp, j, k);
    double cy = (double)jdir_from(g, prod, c + n.ym, i, n.ym ? j - 1 : j, k);
    double dd = (double)jdir_from(g, prod, c + n.yp, i, n.yp ? j + 1 : j, k);
    double e = (double)jdir_from(g, prod, c + n.zm, i, j, n.zm ? k - 1 : k);
    double f = (double)jdir_from(g, prod, c + n.zp, i, j, n.zp ? k + 1 : k);
    if (g.mask) {
      const int m = g.mask[c];
      a = (m & 1) ? a : xc;
      b = (m & 2) ? b : xc;
      cy = (m & 4) ? cy : xc;
      dd = (m & 8) ? dd : xc;
      e = (m & 16) ? e : xc;
      f = (m & 32) ? f : xc;
    }
    v = 0.0;
    v += g.w[0] * (xc - a);
    v += g.w[0] * (xc - b);
    v += g.w[1] * (xc - cy);
    v += g.w[1] * (xc - dd);
    v += g.w[2] * (xc - e);
    v += g.w[2] * (xc - f);
    dc = xc;
    d[c] = dT;
    if (is_gauge(g, c)) {
      r[c] = (T)0.0;
    } else {
      const T rn = prod[c];
      r[c] = rn;
      const double rv = (double)rn;
      dd_acc(s[1], rv * jdiv(rv, diag_at(g, c, i, j, k)));
    }
  }
  if (!isfinite(v)) bad[0] = 1.0;
  if (is_gauge(g, c)) v = 0.0;
  const T qs = (T)v;
  q[c] = qs;
  dd_acc(s[0], dc * (double)qs);
  END_CELLS
  block_publish<2, 1>(s, bad, pdd, pmx);
}

// Jacobi PCG direction for iterations >= 1 (not replaced: FinUpdate cleared
// the flag), with the true residual of a candidate iteration in the same
// pass: prod = b - K p into `prod` and its partials (k_true_residual).  A
// candidate's next-rz check waits for FinTrueJ / FinApplyDirJ.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_direction_j(DevGrid g, const double* __restrict__ p, const double* __restrict__ bvec, const T* __restrict__ r,
              T* __restrict__ d, T* __restrict__ prod, PcgControl* ctl, DD* pdd, double* pmx, int it) {
  if (ctl->done) return;
  const bool cand = ctl->candidate;
  const double next_rz = ctl->rz[(it + 1) & 1];
  if (!cand && (!isfinite(next_rz) || next_rz <= 0.0)) {
    if (blockIdx.x == 0 && flat_tid() == 0) fail(ctl, RH_BREAKDOWN, it);
    return;
  }
  const double beta = next_rz / ctl->rz[it & 1];
  DD s[2] = {{0.0, 0.0}, {0.0, 0.0}};
  double m[2] = {0.0, 0.0};
  FOR_CELLS(g)
  if (cand) {
    const double v = apply_at<double>(g, p, c, i, j, k);
    if (!isfinite(v)) m[1] = 1.0;
    const double pr = bvec[c] - v;
    prod[c] = (T)pr;
    dd_acc(s[0], pr * pr);
    if (!is_gauge(g, c)) dd_acc(s[1], pr * pr);  // without the gauge rows
    m[0] = fmax(m[0], fabs(pr));
  }
  if (is_gauge(g, c)) {
    d[c] = (T)0.0;
    continue;
  }
  const double zv = jdiv(ld(r, c), diag_at(g, c, i, j, k));
  d[c] = (T)(zv + beta * ld(d, c));
  END_CELLS
  if (cand) block_publish<2, 2>(s, m, pdd, pmx);
  if (!cand && blockIdx.x == 0 && flat_tid() == 0) ctl->iteration = it;
}


// Stateless operator kernels for the C ABI -----------------------------------

__global__ void __launch_bounds__(kBlock)
k_apply_full(DevGrid g, const double* __restrict__ x, double* __restrict__ y, unsigned* flags) {
  unsigned bad = 0;
  FOR_CELLS(g)
  const double v = apply_at<double>(g, x, c, i, j, k);
  if (!isfinite(v)) bad |= kFlagArithmetic;
  y[c] = v;
  END_CELLS
  if (bad) atomicOr(flags, bad);
}

// Residual report of a supplied pressure (ProviderWorkspace qualification,
// src/solver_provider.rs:246-254): sum (b - K p)^2, max|b - K p|, nonfinite.
__device__ __forceinline__ void b_residual_report(VBlock vb, const DevGrid& g,
                                                  const double* __restrict__ p,
                                                  const double* __restrict__ b, DD* pdd, double* pmx) {
  DD s[1] = {{0.0, 0.0}};
  double m[2] = {0.0, 0.0};
  FOR_CELLS_VB(g, vb)
  const double v = apply_at<double>(g, p, c, i, j, k);
  if (!isfinite(v)) m[1] = 1.0;
  const double prod = b[c] - v;
  dd_acc(s[0], prod * prod);
  m[0] = fmax(m[0], fabs(prod));
  END_CELLS
  block_publish<1, 2>(s, m, pdd, pmx, vb);
}
__global__ void __launch_bounds__(kBlock)
k_residual_report(DevGrid g, const double* __restrict__ p, const double* __restrict__ b, DD* pdd,
                  double* pmx) {
  b_residual_report(hw_block(), g, p, b, pdd, pmx);
}

// ---------------------------------------------------------------- batched

#define RH_BATCH_PROLOGUE(ARGS)                                  \
  const int s_ = batch_scene(prefix, n, blockIdx.x);            \
  const ARGS& a = A[s_];                                        \
  const VBlock vb = batch_block(prefix, s_);

__global__ void __launch_bounds__(kBlock)
kb_rhs_sums(const BCellArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BCellArgs)
  if (a.ctl->done) return;
  b_rhs_sums(vb, *a.g, a.b, a.pdd, a.pmx);
}
// residual of the returned pressure: scenes whose solve ended RH_OK
__global__ void __launch_bounds__(kBlock)
kb_residual_report(const BCellArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BCellArgs)
  if (a.ctl->status != RH_OK) return;
  b_residual_report(vb, *a.g, a.p, a.b, a.pdd, a.pmx);
}
__global__ void __launch_bounds__(kBlock)
kb_fill_pressure(const BCellArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BCellArgs)
  const DevGrid& g = *a.g;
  FOR_CELLS_VB(g, vb)
  a.p[c] = 0.0;
  END_CELLS
}
// p <- p - p[gauge] (captured in spare[0] when spare[1] != 0), solid cells 0
__global__ void __launch_bounds__(kBlock)
kb_regauge(const BCellArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BCellArgs)
  if (a.ctl->spare[1] == 0.0) return;
  const double p0 = a.ctl->spare[0];
  const DevGrid& g = *a.g;
  FOR_CELLS_VB(g, vb)
  a.p[c] = is_fluid(g, c) ? a.p[c] - (g.ncomp > 1 ? g.comp_shift[g.comp_label[c]] : p0) : 0.0;
  END_CELLS
}
// Per-region compatibility of split-fluid scenes (blocks = chunks; scenes
// with one region have none), then one checking block per scene.
__global__ void __launch_bounds__(kBlock)
kb_comp_sums(const BCellArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BCellArgs)
  if (a.ctl->done) return;
  b_comp_sums(*a.g, a.b, vb, a.pdd, a.pmx);
}
__global__ void __launch_bounds__(kBlock) kb_comp_check(const BCellArgs* __restrict__ A, int n) {
  const BCellArgs& a = A[blockIdx.x];
  if (a.g->ncomp <= 1 || a.ctl->done) return;
  b_comp_check(*a.g, a.pdd, a.ctl);
}

void launch_b_rhs_sums(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks) kb_rhs_sums<<<bl.blocks, block_shape(), 0, s>>>((const BCellArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_residual_report(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_residual_report<<<bl.blocks, block_shape(), 0, s>>>((const BCellArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_fill_pressure(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_fill_pressure<<<bl.blocks, block_shape(), 0, s>>>((const BCellArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_regauge(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks) kb_regauge<<<bl.blocks, block_shape(), 0, s>>>((const BCellArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_comp_sums(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks) kb_comp_sums<<<bl.blocks, block_shape(), 0, s>>>((const BCellArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_comp_check(const BatchLaunch& bl, hipStream_t s) {
  if (bl.n) kb_comp_check<<<bl.n, block_shape(), 0, s>>>((const BCellArgs*)bl.args, bl.n);
}

// ------------------------------------------------------------ host launchers

template <typename T>
void pcg_prologue(const DevGrid& g, const PcgArrays& a, PcgControl* ctl, DD* pdd, double* pmx,
                  hipStream_t s) {
  const int n = g.nx * g.ny * g.nz;
  const int fill_blocks = n / kBlock + 1 < kMaxBlocks ? n / kBlock + 1 : kMaxBlocks;
  const dim3 grid = cells_grid(g);
  k_rhs_sums<<<grid, block_shape(), 0, s>>>(g, a.b, pdd, pmx);
  launch_finish<3, 1>(pdd, pmx, grid.x, FinRhsSums{ctl}, s);
  launch_comp_check(g, a.b, ctl, pdd, pmx, s);
  if (a.seed)
    launch_seed(g, a.seed, a.p64, ctl, pdd, pmx, s);
  else
    k_fill_zero<double><<<fill_blocks, block_shape(), 0, s>>>(a.p64, n);
  launch(k_true_residual<T>, grid, block_shape(), 0, s, g, a.p64, a.b, (T*)a.q, ctl, pdd, pmx, 0);
  launch_finish<2, 2>(pdd, pmx, grid.x, FinTrue{ctl, 0}, s);
}

template <typename T>
void pcg_iteration(const DevGrid& g, int method, const PcgArrays& a, PcgControl* ctl, DD* pdd,
                   double* pmx, int it, hipStream_t s) {
  const dim3 grid = cells_grid(g), block = block_shape();
  const int nb = grid.x;
  double* p = a.p64;
  T* r = (T*)a.r;
  T* d = (T*)a.d;
  T* q = (T*)a.q;
  T* z = (T*)a.z;
  const bool lex = method == RH_SGS_PCG;
  const bool sgs = method == RH_RBSGS_PCG || lex;
  if (!sgs && it > 0) {
    // apply (+ the previous iteration's replacement) -> update -> direction
    // (+ this iteration's true residual when it is a candidate); the true
    // residual goes to z, which the Jacobi iteration does not store
    {
      prof::Scope ps(prof::kApplyDir, s, it);
      launch(k_apply_dir_j<T>, grid, block, 0, s, g, r, d, q, (const T*)z, ctl, pdd, pmx);
    }
    launch_finish<2, 1>(pdd, pmx, nb, FinApplyDirJ{ctl, it}, s);
    {
      prof::Scope ps(prof::kUpdate, s, it);
      launch(k_update<T, false>, grid, block, 0, s, g, p, r, d, q, z, ctl, pdd, pmx);
    }
    launch_finish<3, 1>(pdd, pmx, nb, FinUpdate{ctl, it, 1}, s);
    {
      prof::Scope ps(prof::kDirection, s, it);
      launch(k_direction_j<T>, grid, block, 0, s, g, p, a.b, r, d, z, ctl, pdd, pmx, it);
    }
    launch_finish<2, 2>(pdd, pmx, nb, FinTrueJ{ctl, it}, s);
    return;
  }
  if (it > 0) {
    {
      prof::Scope ps(prof::kApplyDir, s, it);
      launch(k_apply_dir<T>, grid, block, 0, s, g, d, q, ctl, pdd, pmx);
    }
    launch_finish<1, 1>(pdd, pmx, nb, FinApplyDir{ctl, it}, s);
    {
      prof::Scope ps(prof::kUpdate, s, it);
      if (sgs)
        launch(k_update<T, true>, grid, block, 0, s, g, p, r, d, q, z, ctl, pdd, pmx);
      else
        launch(k_update<T, false>, grid, block, 0, s, g, p, r, d, q, z, ctl, pdd, pmx);
    }
    launch_finish<3, 1>(pdd, pmx, nb, FinUpdate{ctl, it, sgs ? 0 : 1}, s);
    if (lex) {
      prof::Scope ps(prof::kSgsRed, s, it);
      sgs_lex_apply<T>(g, r, z, a.tiles, a.wave_start, a.waves, ctl, pdd, pmx, it, 0, s);
    } else if (sgs) {
      {
        prof::Scope ps(prof::kSgsBlack, s, it);
        launch(k_sgs_black<T>, grid, block, 0, s, g, r, z, ctl, it, 0);
      }
      {
        prof::Scope ps(prof::kSgsRed, s, it);
        launch(k_sgs_red<T>, grid, block, 0, s, g, r, z, ctl, pdd, pmx, it, 0);
      }
      launch_finish<1, 1>(pdd, pmx, nb, FinSgsRed{ctl, it, 0}, s);
    }
    {
      prof::Scope ps(prof::kTrueResidual, s, it);
      launch(k_true_residual<T>, grid, block, 0, s, g, p, a.b, q, ctl, pdd, pmx, it);
    }
    launch_finish<2, 2>(pdd, pmx, nb, FinTrue{ctl, it}, s);
  }
  if (sgs) {
    {
      prof::Scope ps(prof::kReplace, s, it);
      launch(k_replace<T, true>, grid, block, 0, s, g, r, q, z, ctl, pdd, pmx, it);
    }
    if (lex) {
      prof::Scope ps(prof::kSgsRed, s, it);
      sgs_lex_apply<T>(g, r, z, a.tiles, a.wave_start, a.waves, ctl, pdd, pmx, it, 1, s);
    } else {
      {
        prof::Scope ps(prof::kSgsBlack, s, it);
        launch(k_sgs_black<T>, grid, block, 0, s, g, r, z, ctl, it, 1);
      }
      {
        prof::Scope ps(prof::kSgsRed, s, it);
        launch(k_sgs_red<T>, grid, block, 0, s, g, r, z, ctl, pdd, pmx, it, 1);
      }
      launch_finish<1, 1>(pdd, pmx, nb, FinSgsRed{ctl, it, 1}, s);
    }
    {
      prof::Scope ps(prof::kDirection, s, it);
      launch(k_direction<T, true>, grid, block, 0, s, g, r, z, d, ctl, it);
    }
  } else {
    {
      prof::Scope ps(prof::kReplace, s, it);
      launch(k_replace<T, false>, grid, block, 0, s, g, r, q, z, ctl, pdd, pmx, it);
    }
    launch_finish<1, 1>(pdd, pmx, nb, FinReplace{ctl, it}, s);
    {
      prof::Scope ps(prof::kDirection, s, it);
      launch(k_direction<T, false>, grid, block, 0, s, g, r, z, d, ctl, it);
    }
  }
}

// The solution accumulator is the f64 output itself for every vector type.
template <typename T>
void pcg_export_pressure(const DevGrid& g, const PcgArrays& a, hipStream_t s) {
  (void)g;
  (void)a;
  (void)s;
}

template void pcg_prologue<double>(const DevGrid&, const PcgArrays&, PcgControl*, DD*, double*,
                                   hipStream_t);
template void pcg_prologue<float>(const DevGrid&, const PcgArrays&, PcgControl*, DD*, double*,
                                  hipStream_t);
template void pcg_iteration<double>(const DevGrid&, int, const PcgArrays&, PcgControl*, DD*,
                                    double*, int, hipStream_t);
template void pcg_iteration<float>(const DevGrid&, int, const PcgArrays&, PcgControl*, DD*,
                                   double*, int, hipStream_t);
template void pcg_export_pressure<double>(const DevGrid&, const PcgArrays&, hipStream_t);
template void pcg_export_pressure<float>(const DevGrid&, const PcgArrays&, hipStream_t);

void launch_apply_full(const DevGrid& g, const double* x, double* y, unsigned* flags,
                       hipStream_t s) {
  k_apply_full<<<cells_grid(g), block_shape(), 0, s>>>(g, x, y, flags);
}

// Stateless z = M^-1 r (gauge row zeroed); rz = r.z lands in ctl->rz[1].
// Reuses the it = 0 path of the PCG kernels with prod aliasing r; `ctl`
// must be zeroed by the caller.
void launch_precondition(const DevGrid& g, int method, double* r, double* z, PcgControl* ctl,
                         DD* pdd, double* pmx, hipStream_t s, const SgsTile* tiles, const int* wave_start,
                         int waves) {
  const dim3 grid = cells_grid(g), block = block_shape();
  if (method == RH_SGS_PCG) {
    k_replace<double, true><<<grid, block, 0, s>>>(g, r, r, z, ctl, pdd, pmx, 0);
    sgs_lex_apply<double>(g, r, z, tiles, wave_start, waves, ctl, pdd, pmx, 0, 1, s);
  } else if (method == RH_RBSGS_PCG) {
    k_replace<double, true><<<grid, block, 0, s>>>(g, r, r, z, ctl, pdd, pmx, 0);
    k_sgs_black<double><<<grid, block, 0, s>>>(g, r, z, ctl, 0, 1);
    k_sgs_red<double><<<grid, block, 0, s>>>(g, r, z, ctl, pdd, pmx, 0, 1);
    launch_finish<1, 1>(pdd, pmx, grid.x, FinSgsRed{ctl, 0, 1}, s);
  } else {
    k_replace<double, false><<<grid, block, 0, s>>>(g, r, r, z, ctl, pdd, pmx, 0);
    launch_finish<1, 1>(pdd, pmx, grid.x, FinReplace{ctl, 0}, s);
    // z = r/D through the direction kernel (d := z, copy branch at it = 0)
    k_direction<double, false><<<grid, block, 0, s>>>(g, r, z, z, ctl, 0);
  }
}

void launch_residual_report(const DevGrid& g, const double* p, const double* b, const Reduce& red,
                            hipStream_t s) {
  const dim3 grid = cells_grid(g);
  k_residual_report<<<grid, block_shape(), 0, s>>>(g, p, b, red.pdd, red.pmx);
  launch_finish<1, 2>(red.pdd, red.pmx, grid.x, FinStore{red.out, 2, 0.0, 0.0, 0.0}, s);
}

void launch_rhs_sums_finish(const DevGrid& g, const double* b, PcgControl* ctl, DD* pdd, double* pmx,
                            hipStream_t s) {
  launch_finish<3, 1>(pdd, pmx, cells_grid(g).x, FinRhsSums{ctl}, s);
  launch_comp_check(g, b, ctl, pdd, pmx, s);
}
void launch_rhs_sums(const DevGrid& g, const double* b, PcgControl* ctl, DD* pdd, double* pmx,
                     hipStream_t s) {
  const dim3 grid = cells_grid(g);
  k_rhs_sums<<<grid, block_shape(), 0, s>>>(g, b, pdd, pmx);
  launch_finish<3, 1>(pdd, pmx, grid.x, FinRhsSums{ctl}, s);
  launch_comp_check(g, b, ctl, pdd, pmx, s);
}

}  // namespace rh
