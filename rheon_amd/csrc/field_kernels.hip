// Face-field kernels of the projection step: RK2 semi-Lagrangian advection,
// RHS (integrated divergence), pressure-gradient correction, divergence
// maximum and velocity rate.  Each restates one reference routine with the
// same operation order (cited per kernel); built with -ffp-contract=off.
#include <cfloat>

#include "common.hpp"
#include "kernels.hpp"
#include "batch.hpp"
#include "profiler.hpp"

namespace rh {

// ---------------------------------------------------------------- sampling
// (x - origin) / h of a world position clamped to the box, per axis: the
// part of ScalarSampler::sample (src/sampling.rs:87-88) shared by the three
// velocity components at one position.  When h is a power of two the
// reciprocal is exact and the product equals the quotient bit for bit.
// kInBox: the position is already clamped to the box (the RK2 midpoint and
// departure are), so the clamp is the identity and is left out.  The clamp
// is written as max/min: positions here are finite, where both forms agree
// (up to the sign of a zero coordinate, which cannot change a sample).
template <bool kInBox = false>
__device__ __forceinline__ void base_coords(const DevGrid& g, const double pos[3], double out[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double x = pos[d];
    if (!kInBox) x = fmin(fmax(x, g.origin[d]), g.upper[d]);
    out[d] = g.pow2[d] ? (x - g.origin[d]) * g.inv_h[d] : (x - g.origin[d]) / g.h[d];
  }
}

// Element `off` bytes past `vals`.  kOff32: every byte offset into the array
// fits in 32 bits (the host checks the array size), so the load takes the
// uniform base plus a 32-bit lane offset instead of a 64-bit address per lane.
template <typename T, bool kOff32>
__device__ __forceinline__ T load_elem(const T* __restrict__ vals, long long idx) {
  if (kOff32)
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(vals) + (unsigned)idx * (unsigned)sizeof(T));
  return vals[idx];
}

// x / h[d]: the product with the exact reciprocal where h is a power of two
// (the same value bit for bit), the quotient otherwise.
__device__ __forceinline__ double over_h(const DevGrid& g, int d, double x) {
  return g.pow2[d] ? x * g.inv_h[d] : x / g.h[d];
}

// ScalarSampler::sample (src/sampling.rs:74-122) from precomputed base
// coordinates: comp 0..2 = face lattice of that axis, 3 = cell centres; f64
// arithmetic on T samples, in the reference's weight and summation order.  The clamp of q to
// [0, top] is written as max/min after a NaN test: the reference flags only a
// NaN q (an infinite one clamps to a finite end), and the two forms agree
// elsewhere up to the sign of a zero q, which only changes zero-weight terms.
// t = q - floor(q) equals q - (double)lo exactly.  For T = float the corner
// minimum and maximum are taken in float (exact: the widening is monotone;
// they matter only when every corner is finite, else the sum is not).
// Branch-free: a failing sample still completes (max/min maps a NaN q to 0,
// so every corner index stays inside the array) and returns the flag; the
// caller discards the value of a flagged face.
template <typename T, bool kOff32 = false>
__device__ __forceinline__ unsigned sample_at(const DevGrid& g, int comp, const T* __restrict__ vals,
                                              const double base[3], double& out) {
  const int n[3] = {g.nx, g.ny, g.nz};
  int lo[3], hi[3], dims[3];
  double t[3];
  bool bad = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    dims[d] = n[d] + (comp == d ? 1 : 0);
    const double off = (comp == d) ? 0.0 : 0.5;
    double q = base[d] - off;
    const double top = (double)(dims[d] - 1);
    bad = bad || isnan(q);
    q = fmin(fmax(q, 0.0), top);
    const double fl = floor(q);
    lo[d] = (int)fl;
    hi[d] = lo[d] + 1 < dims[d] - 1 ? lo[d] + 1 : dims[d] - 1;
    t[d] = q - fl;
  }
  // corner (x, y, z) sits at c0 + {0, ox} + {0, oy} + {0, oz}: two
  // integer multiplies per sample instead of four (hi[d] - lo[d] is 0 or 1)
  const int sy = dims[0], sz = dims[0] * dims[1];
  // kOff32 also guarantees 24-bit operands (face_offsets_32): full-rate
  // 24-bit multiplies give the same index
  const int c0 = kOff32 ? (int)__umul24(__umul24(lo[2], dims[1]) + lo[1], dims[0]) + lo[0]
                        : (lo[2] * dims[1] + lo[1]) * dims[0] + lo[0];
  const int ox = hi[0] - lo[0];
  const int oy = hi[1] != lo[1] ? sy : 0;
  const int oz = hi[2] != lo[2] ? sz : 0;
  T raw[8];
#pragma unroll
  for (int z = 0; z < 2; ++z)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int x = 0; x < 2; ++x)
        raw[x + 2 * y + 4 * z] = load_elem<T, kOff32>(vals, c0 + (x ? ox : 0) + (y ? oy : 0) + (z ? oz : 0));
  double result = 0.0, minimum, maximum;
  if (sizeof(T) == 4) {
    T lo4 = raw[0], hi4 = raw[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) {
      lo4 = fminf(lo4, raw[q]);
      hi4 = fmaxf(hi4, raw[q]);
    }
    minimum = (double)lo4;
    maximum = (double)hi4;
  } else {
    minimum = INFINITY;
    maximum = -INFINITY;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      minimum = fmin(minimum, (double)raw[q]);
      maximum = fmax(maximum, (double)raw[q]);
    }
  }
#pragma unroll
  for (int z = 0; z < 2; ++z)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        double weight = (x ? t[0] : 1.0 - t[0]);
        weight = weight * (y ? t[1] : 1.0 - t[1]);
        weight = weight * (z ? t[2] : 1.0 - t[2]);
        result += weight * (double)raw[x + 2 * y + 4 * z];
      }
  bad = bad || !isfinite(result);
  if (result < minimum) result = minimum;
  if (result > maximum) result = maximum;
  out = result;
  return bad ? kFlagArithmetic : 0u;
}

// VelocitySampler::backtrace / departure (src/sampling.rs:151-175).
// The clamps to the box follow a finite test, so max/min is the same clamp
// for every accepted position.  Branch-free like sample_at: a NaN or
// infinite position is flagged and max/min puts it on the box (so later
// samples stay in bounds); the departure of a flagged face is discarded.
template <typename T, bool kOff32 = false>
__device__ __forceinline__ unsigned backtrace(const DevGrid& g, const T* __restrict__ u,
                                              const T* __restrict__ v, const T* __restrict__ w,
                                              const double pos[3], double dt, double out[3]) {
  double vel[3], mid[3], base[3];
  base_coords(g, pos, base);
  unsigned f = sample_at<T, kOff32>(g, 0, u, base, vel[0]) | sample_at<T, kOff32>(g, 1, v, base, vel[1]) |
               sample_at<T, kOff32>(g, 2, w, base, vel[2]);
  const double half = 0.5 * dt;
  bool bad = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double value = pos[d] - half * vel[d];
    bad = bad || !isfinite(value);
    mid[d] = fmin(fmax(value, g.origin[d]), g.upper[d]);
  }
  base_coords<true>(g, mid, base);
  f |= sample_at<T, kOff32>(g, 0, u, base, vel[0]) | sample_at<T, kOff32>(g, 1, v, base, vel[1]) |
       sample_at<T, kOff32>(g, 2, w, base, vel[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double value = pos[d] - dt * vel[d];
    bad = bad || !isfinite(value);
    out[d] = fmin(fmax(value, g.origin[d]), g.upper[d]);
  }
  return f | (bad ? kFlagArithmetic : 0u);
}

// First RK2 stage at a face position of a lattice grid (DevGrid::lattice,
// f32 faces).  For the face (i, j, k) of axis a every base coordinate is an
// exact lattice coordinate, so ScalarSampler::sample's weights are products
// of 0, 1/2 and 1 (src/sampling.rs:89-113):
//  * component a samples its own lattice point: weight 1 on the face itself
//    and exactly +0 on the other seven corners, so the sum is 0.0 + u;
//  * component c != a has t = 1/2 on axes a and c (lower corners p_a - 1 and
//    p_c, both upper corners inside the support) and t = 0 on the third axis:
//    four corners of weight exactly 1/4, taken in the reference's z, y, x
//    order, and four of weight +0.
// The terms of weight +0 are +-0 and leave the running sum unchanged: it
// starts at +0.0 and a round-to-nearest sum is -0 only when both operands
// are, so it never holds -0.  The clamp to the corners' [min, max] is the
// identity: the k-th partial sum is at most round(k/4 * max) = k/4 * max
// (representable for an f32 max) and, likewise, at least k/4 * min.  With
// finite faces (validated before sampling) every sum is finite.  So the
// value is the one the general sample returns, bit for bit.
template <bool kOff32>
__device__ __forceinline__ double lattice_half_sample(const DevGrid& g, int a, int c, const float* __restrict__ vals,
                                                      int i, int j, int k) {
  const int dx = g.nx + (c == 0), dy = g.ny + (c == 1);
  const int p[3] = {i - (a == 0), j - (a == 1), k - (a == 2)};
  const int c0 = kOff32 ? p[0] + (int)__umul24(dx, p[1] + (int)__umul24(dy, p[2])) : p[0] + dx * (p[1] + dy * p[2]);
  const int lo_ax = a < c ? a : c, hi_ax = a < c ? c : a;
  const int s_in = lo_ax == 0 ? 1 : dx;       // inner (lower-index) active axis
  const int s_out = hi_ax == 1 ? dx : dx * dy;  // outer active axis
  const double v0 = (double)load_elem<float, kOff32>(vals, c0);
  const double v1 = (double)load_elem<float, kOff32>(vals, c0 + s_in);
  const double v2 = (double)load_elem<float, kOff32>(vals, c0 + s_out);
  const double v3 = (double)load_elem<float, kOff32>(vals, c0 + s_in + s_out);
  double r = 0.0;
  r += 0.25 * v0;
  r += 0.25 * v1;
  r += 0.25 * v2;
  r += 0.25 * v3;
  return r;
}

// backtrace from the face (i, j, k) of axis a at `pos`, with the first
// stage's samples taken by lattice_half_sample (f32 faces, lattice grid);
// the rest as backtrace().
template <bool kOff32>
__device__ __forceinline__ unsigned backtrace_lattice_face(const DevGrid& g, int a, const float* __restrict__ u,
                                                           const float* __restrict__ v, const float* __restrict__ w,
                                                           int i, int j, int k, const double pos[3], double dt,
                                                           double out[3]) {
  const float* comp[3] = {u, v, w};
  double vel[3], mid[3], base[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c == a) {
      const int fx = g.nx + (a == 0), fy = g.ny + (a == 1);
      vel[c] = 0.0 + (double)load_elem<float, kOff32>(comp[c], i + fx * (j + fy * k));
    } else {
      vel[c] = lattice_half_sample<kOff32>(g, a, c, comp[c], i, j, k);
    }
  }
  const double half = 0.5 * dt;
  bool bad = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double value = pos[d] - half * vel[d];
    bad = bad || !isfinite(value);
    mid[d] = fmin(fmax(value, g.origin[d]), g.upper[d]);
  }
  base_coords<true>(g, mid, base);
  unsigned f = sample_at<float, kOff32>(g, 0, u, base, vel[0]) | sample_at<float, kOff32>(g, 1, v, base, vel[1]) |
               sample_at<float, kOff32>(g, 2, w, base, vel[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double value = pos[d] - dt * vel[d];
    bad = bad || !isfinite(value);
    out[d] = fmin(fmax(value, g.origin[d]), g.upper[d]);
  }
  return f | (bad ? kFlagArithmetic : 0u);
}

// Minimum waves per SIMD requested for the advection kernels.  The
// branch-free sampling keeps a face's 24 gathers of one RK2 stage in flight
// at once; the general form at 3 waves (up to 168 VGPRs) holds them without
// spilling, which measured faster than 4 waves with spills.  The lattice
// form fits 4 waves (128 VGPRs): 1.00 -> 0.88 ms per 256^3 launch.
#ifndef RH_ADVECT_WAVES
#define RH_ADVECT_WAVES 3
#endif
#ifndef RH_ADVECT_LATTICE_WAVES
#define RH_ADVECT_LATTICE_WAVES 4
#endif
constexpr int kAdvectWaves = RH_ADVECT_WAVES;
constexpr int kAdvectLatticeWaves = RH_ADVECT_LATTICE_WAVES;

// advect_velocity (src/advection.rs:28-68) of face (i, j, k) of component
// a, flat index idx: wall-normal faces are exactly zero; every other face
// samples its own component at the RK2 departure of its own face position
// (src/geometry.rs:240-248).  Returns the flags of a failed sample (the
// face is then left unwritten).
// kLattice (f32 faces): 1 = the grid is a lattice grid, 2 = decided per
// grid at run time; the first RK2 stage is then taken by
// backtrace_lattice_face, the same values with a fraction of the work.
template <typename T, bool kOff32, int kLattice>
__device__ __forceinline__ unsigned advect_face(const DevGrid& g, int a, int i, int j, int k, int idx, int top,
                                                const T* __restrict__ u, const T* __restrict__ v,
                                                const T* __restrict__ w, const T* __restrict__ src, double dt,
                                                T* __restrict__ dst) {
  const int p[3] = {i, j, k};
  if (p[a] == 0 || p[a] == top || !face_open(g, a, i, j, k)) {
    dst[idx] = (T)0.0;
    return 0u;
  }
  double pos[3], dep[3], value = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    pos[d] = g.origin[d] + ((double)p[d] + (d == a ? 0.0 : 0.5)) * g.h[d];
  unsigned f;
  if constexpr (kLattice != 0) {
    if (kLattice == 1 || g.lattice)
      f = backtrace_lattice_face<kOff32>(g, a, u, v, w, i, j, k, pos, dt, dep);
    else
      f = backtrace<T, kOff32>(g, u, v, w, pos, dt, dep);
  } else {
    f = backtrace<T, kOff32>(g, u, v, w, pos, dt, dep);
  }
  double base[3];
  base_coords<true>(g, dep, base);
  f |= sample_at<T, kOff32>(g, a, src, base, value);
  if (f) return f;
  dst[idx] = (T)value;
  return 0u;
}

// Component a (blockIdx.y) of block vb.  kFlat (batched launches): a
// contiguous chunk of the component's faces in memory order, a whole
// number of passes of the block's threads, so rows as short as 33 faces
// keep every lane busy.  Otherwise rows of faces, lanes along x (one-scene
// grids, whose rows are long; the flat form's index divisions cost 2% of a
// 256^3 launch there).
template <typename T, bool kOff32, int kLattice = 0, bool kFlat = false>
__device__ __forceinline__ void
b_advect_velocity(VBlock vb, int a, const DevGrid& g, const T* __restrict__ u, const T* __restrict__ v,
                  const T* __restrict__ w, double dt, T* __restrict__ ou, T* __restrict__ ov,
                  T* __restrict__ ow, unsigned* __restrict__ flags) {
  static_assert(kLattice == 0 || sizeof(T) == 4, "lattice first stage is stated for f32 faces");
  const T* src = a == 0 ? u : (a == 1 ? v : w);
  T* dst = a == 0 ? ou : (a == 1 ? ov : ow);
  const int fx = g.nx + (a == 0), fy = g.ny + (a == 1), fz = g.nz + (a == 2);
  const int top = (a == 0 ? fx : (a == 1 ? fy : fz)) - 1;
  unsigned bad = 0;
  if constexpr (kFlat) {
    const FastDiv dx = a == 0 ? g.div_x1 : g.div_x, dy = a == 1 ? g.div_y1 : g.div_y;
    const RowRange fr = chunk_span(fx * fy * fz, vb);
    for (int t = fr.begin + flat_tid(); t < fr.end; t += kBlock) {
      const int row = fast_div(t, dx), i = t - row * fx;
      const int k = fast_div(row, dy), j = row - k * fy;
      bad |= advect_face<T, kOff32, kLattice>(g, a, i, j, k, t, top, u, v, w, src, dt, dst);
    }
  } else {
    const RowRange rr = block_rows(fy * fz, vb);
    for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
      const int k = row / fy, j = row - k * fy;
      for (int i = threadIdx.x; i < fx; i += kLanesX)
        bad |= advect_face<T, kOff32, kLattice>(g, a, i, j, k, i + fx * row, top, u, v, w, src, dt, dst);
    }
  }
  if (bad) atomicOr(flags, bad);
}
template <typename T, bool kOff32, int kLattice = 0>
__global__ void __launch_bounds__(kBlock, kLattice == 1 ? kAdvectLatticeWaves : kAdvectWaves)
k_advect_velocity(DevGrid g, const T* __restrict__ u, const T* __restrict__ v,
                  const T* __restrict__ w, double dt, T* __restrict__ ou, T* __restrict__ ov,
                  T* __restrict__ ow, unsigned* __restrict__ flags) {
  b_advect_velocity<T, kOff32, kLattice>(hw_block(), (int)blockIdx.y, g, u, v, w, dt, ou, ov, ow, flags);
}

// First RK2 stage at a cell centre of a lattice grid (DevGrid::lattice, f32
// faces): component c has t = 1/2 on its own axis (corners p_c and p_c + 1)
// and t = 0 on the other two, so ScalarSampler::sample reduces to the two
// corners of weight exactly 1/2 in the reference's order; the other six
// weigh +0 and the clamp is the identity (the argument of
// lattice_half_sample with two terms).
template <bool kOff32>
__device__ __forceinline__ double lattice_cell_sample(const DevGrid& g, int c, const float* __restrict__ vals, int i,
                                                      int j, int k) {
  const int dx = g.nx + (c == 0), dy = g.ny + (c == 1);
  const int c0 = kOff32 ? i + (int)__umul24(dx, j + (int)__umul24(dy, k)) : i + dx * (j + dy * k);
  const int s = c == 0 ? 1 : (c == 1 ? dx : dx * dy);
  double r = 0.0;
  r += 0.5 * (double)load_elem<float, kOff32>(vals, c0);
  r += 0.5 * (double)load_elem<float, kOff32>(vals, c0 + s);
  return r;
}

// advect_tracer (src/advection.rs:72-162).  Cells outside the support box
// [lo, hi) are exactly +0.0, the value the reference writes there without
// sampling (TracerSupport, :104-153; the caller passes the whole box for
// the reference's dense fallback).  kLattice: the first RK2 stage by
// lattice_cell_sample (same values).
template <bool kOff32, bool kLattice>
__global__ void __launch_bounds__(kBlock)
k_advect_tracer(DevGrid g, const float* __restrict__ old, const float* __restrict__ u,
                const float* __restrict__ v, const float* __restrict__ w, double dt,
                float* __restrict__ out, unsigned* __restrict__ flags, TracerBox box) {
  const RowRange rr = block_rows(g.ny * g.nz);
  unsigned bad = 0;
  const float* comp[3] = {u, v, w};
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / g.ny, j = row - k * g.ny;
    const bool row_in = j >= box.lo[1] && j < box.hi[1] && k >= box.lo[2] && k < box.hi[2];
    for (int i = threadIdx.x; i < g.nx; i += kLanesX) {
      if (!row_in || i < box.lo[0] || i >= box.hi[0]) {
        out[i + g.nx * row] = 0.0f;
        continue;
      }
      const int p[3] = {i, j, k};
      double pos[3], dep[3], value = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) pos[d] = g.origin[d] + ((double)p[d] + 0.5) * g.h[d];
      unsigned f;
      if constexpr (kLattice) {
        double vel[3], mid[3], base[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) vel[c] = lattice_cell_sample<kOff32>(g, c, comp[c], i, j, k);
        const double half = 0.5 * dt;
        bool nf = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double x = pos[d] - half * vel[d];
          nf = nf || !isfinite(x);
          mid[d] = fmin(fmax(x, g.origin[d]), g.upper[d]);
        }
        base_coords<true>(g, mid, base);
        f = sample_at<float, kOff32>(g, 0, u, base, vel[0]) | sample_at<float, kOff32>(g, 1, v, base, vel[1]) |
            sample_at<float, kOff32>(g, 2, w, base, vel[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double x = pos[d] - dt * vel[d];
          nf = nf || !isfinite(x);
          dep[d] = fmin(fmax(x, g.origin[d]), g.upper[d]);
        }
        f |= nf ? kFlagArithmetic : 0u;
      } else {
        f = backtrace<float, kOff32>(g, u, v, w, pos, dt, dep);
      }
      double base[3];
      base_coords<true>(g, dep, base);
      f |= sample_at<float, kOff32>(g, 3, old, base, value);
      if (f) {
        bad |= f;
        continue;
      }
      out[i + g.nx * row] = (float)value;
    }
  }
  if (bad) atomicOr(flags, bad);
}

// Bounding box of the nonzero cells of a tracer (TracerSupport's exact
// support, src/advection.rs:131-145: a cell counts when its value != 0,
// so -0.0 does not) into bounds[0..5] = min i, j, k and max i, j, k (set to
// n, n, n, 0, 0, 0 beforehand); non-finite values raise the input flag
// (ScalarSampler::cells, src/sampling.rs:62-64).
__global__ void __launch_bounds__(kBlock)
k_tracer_bounds(DevGrid g, const float* __restrict__ t, unsigned* __restrict__ bounds,
                unsigned* __restrict__ flags) {
  const RowRange rr = block_rows(g.ny * g.nz);
  unsigned lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  bool any = false, bad = false;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / g.ny, j = row - k * g.ny;
    for (int i = threadIdx.x; i < g.nx; i += kLanesX) {
      const float x = t[i + g.nx * row];
      if (!isfinite(x)) bad = true;
      if (x != 0.0f) {
        any = true;
        lo[0] = min(lo[0], (unsigned)i);
        lo[1] = min(lo[1], (unsigned)j);
        lo[2] = min(lo[2], (unsigned)k);
        hi[0] = max(hi[0], (unsigned)i + 1);
        hi[1] = max(hi[1], (unsigned)j + 1);
        hi[2] = max(hi[2], (unsigned)k + 1);
      }
    }
  }
  if (bad) atomicOr(flags, kFlagNonfiniteInput);
  if (any) {
    for (int d = 0; d < 3; ++d) {
      atomicMin(&bounds[d], lo[d]);
      atomicMax(&bounds[3 + d], hi[d]);
    }
  }
}

__global__ void k_tracer_bounds_init(DevGrid g, unsigned* bounds) {
  if (threadIdx.x == 0) {
    bounds[0] = (unsigned)g.nx;
    bounds[1] = (unsigned)g.ny;
    bounds[2] = (unsigned)g.nz;
    bounds[3] = bounds[4] = bounds[5] = 0u;
  }
}

// Candidate tracer check and sum (src/simulation.rs:1322-1331): every value
// finite and in [0, 1] (else the arithmetic flag), and the double-double
// sum of the values (the reference sums them serially; this is the sum
// rounded once).
__global__ void __launch_bounds__(kBlock)
k_tracer_sum(DevGrid g, const float* __restrict__ t, Reduce red) {
  const RowRange rr = block_rows(g.ny * g.nz);
  DD dd[1] = {DD{0.0, 0.0}};
  double mx[1] = {0.0};
  bool bad = false;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock)
    for (int i = threadIdx.x; i < g.nx; i += kLanesX) {
      const float x = t[i + g.nx * row];
      if (!isfinite(x) || !(x >= 0.0f && x <= 1.0f)) bad = true;
      dd_acc(dd[0], (double)x);
    }
  if (bad) atomicOr(red.flags, kFlagArithmetic);
  block_publish<1, 1>(dd, mx, red.pdd, red.pmx);
}

// Finite check of every face and (optionally) exact zero on wall-normal faces
// (validate_velocity, src/operator.rs:850-875; SamplingError::NonFiniteField,
// src/sampling.rs:62-64).  blockIdx.y = component.
template <typename T>
__device__ __forceinline__ void
b_validate_faces(VBlock vb, const DevGrid& g, const T* __restrict__ u, const T* __restrict__ v,
                 const T* __restrict__ w, int check_walls, unsigned* __restrict__ flags) {
  const int a = blockIdx.y;
  const T* src = a == 0 ? u : (a == 1 ? v : w);
  const int fx = g.nx + (a == 0), fy = g.ny + (a == 1), fz = g.nz + (a == 2);
  const RowRange rr = block_rows(fy * fz, vb);
  unsigned bad = 0;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / fy, j = row - k * fy;
    for (int i = threadIdx.x; i < fx; i += kLanesX) {
      const double value = (double)src[i + fx * row];
      if (!isfinite(value)) bad |= kFlagNonfiniteInput;
      if (check_walls) {
        const int pa = a == 0 ? i : (a == 1 ? j : k);
        const int top = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
        if ((pa == 0 || pa == top) && value != 0.0) bad |= kFlagNonzeroWall;
      }
    }
  }
  if (bad) atomicOr(flags, bad);
}
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_validate_faces(DevGrid g, const T* __restrict__ u, const T* __restrict__ v,
                 const T* __restrict__ w, int check_walls, unsigned* __restrict__ flags) {
  b_validate_faces<T>(hw_block(), g, u, v, w, check_walls, flags);
}

// build_rhs (src/operator.rs:460-509): b = (sum_d A_d (u_lo - u_hi)) / dt,
// axis order X, Y, Z, f64 from T faces.  kValidate: also b_validate_faces
// with the wall check, each face checked by the cell that reads it as its
// lower face (or, for the last cell along an axis, its upper wall face);
// same flags, one pass fewer.  kSums: also the solver's compatibility
// partials of the values written (b_rhs_sums's sums in its cell order and
// block slots, so its finisher gives the same bits), published to pdd/pmx.
template <typename T, bool kValidate = false, bool kSums = false>
__device__ __forceinline__ void
b_build_rhs(VBlock vb, const DevGrid& g, const T* __restrict__ u, const T* __restrict__ v,
            const T* __restrict__ w, double dt, double* __restrict__ out,
            unsigned* __restrict__ flags, DD* pdd = nullptr, double* pmx = nullptr) {
  const int nx = g.nx, ny = g.ny;
  const RowRange rr = block_rows(g.ny * g.nz, vb);
  unsigned bad = 0;
  DD sums[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  double nonfinite[1] = {0.0};
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / ny, j = row - k * ny;
    for (int i = threadIdx.x; i < nx; i += kLanesX) {
      const int xu = i + (nx + 1) * row;
      const int yv = i + nx * (j + (ny + 1) * k);
      const int zw = i + nx * row;
      const double ulo = (double)u[xu], uhi = (double)u[xu + 1];
      const double vlo = (double)v[yv], vhi = (double)v[yv + nx];
      const double wlo = (double)w[zw], whi = (double)w[zw + nx * ny];
      if (kValidate) {
        if (!isfinite(ulo) || !isfinite(vlo) || !isfinite(wlo)) bad |= kFlagNonfiniteInput;
        if ((i == 0 && ulo != 0.0) || (j == 0 && vlo != 0.0) || (k == 0 && wlo != 0.0)) bad |= kFlagNonzeroWall;
        if (i + 1 == nx) {
          if (!isfinite(uhi)) bad |= kFlagNonfiniteInput;
          if (uhi != 0.0) bad |= kFlagNonzeroWall;
        }
        if (j + 1 == ny) {
          if (!isfinite(vhi)) bad |= kFlagNonfiniteInput;
          if (vhi != 0.0) bad |= kFlagNonzeroWall;
        }
        if (k + 1 == g.nz) {
          if (!isfinite(whi)) bad |= kFlagNonfiniteInput;
          if (whi != 0.0) bad |= kFlagNonzeroWall;
        }
      }
      double flux = 0.0;
      flux += g.area[0] * (ulo - uhi);
      flux += g.area[1] * (vlo - vhi);
      flux += g.area[2] * (wlo - whi);
      const double value = flux / dt;
      if (!isfinite(value)) bad |= kFlagArithmetic;
      out[i + nx * row] = value;
      if (kSums) {
        if (!isfinite(value)) nonfinite[0] = 1.0;
        dd_acc(sums[0], value);
        dd_acc(sums[1], fabs(value));
        dd_acc(sums[2], value * value);
      }
    }
  }
  if (bad) atomicOr(flags, bad);
  if (kSums) block_publish<3, 1>(sums, nonfinite, pdd, pmx, vb);
}
template <typename T, bool kValidate, bool kSums = false>
__global__ void __launch_bounds__(kBlock)
k_build_rhs(DevGrid g, const T* __restrict__ u, const T* __restrict__ v,
            const T* __restrict__ w, double dt, double* __restrict__ out,
            unsigned* __restrict__ flags, DD* pdd, double* pmx) {
  b_build_rhs<T, kValidate, kSums>(hw_block(), g, u, v, w, dt, out, flags, pdd, pmx);
}

// correct_candidate_in_place (src/operator.rs:627-668) for f32 storage:
// u = f32(f64(u) - dt*acc_d*(p_R - p_L)), evaluated left to right;
// correct_candidate_f64 (:746-781) for f64: u - dt*(acc_d*(p_R - p_L)).
template <typename T>
__device__ __forceinline__ void
b_correct(VBlock vb, const DevGrid& g, const double* __restrict__ p, double dt, T* __restrict__ u,
          T* __restrict__ v, T* __restrict__ w, unsigned* __restrict__ flags) {
  const int a = blockIdx.y;
  T* dst = a == 0 ? u : (a == 1 ? v : w);
  const int fx = g.nx + (a == 0), fy = g.ny + (a == 1), fz = g.nz + (a == 2);
  const int stride = a == 0 ? 1 : (a == 1 ? g.nx : g.nx * g.ny);
  const RowRange rr = block_rows(fy * fz, vb);
  const double acc = g.acc[a];
  unsigned bad = 0;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / fy, j = row - k * fy;
    const int pa_row = a == 1 ? j : k;
    const int top_row = a == 1 ? g.ny : g.nz;
    if (a != 0 && (pa_row == 0 || pa_row == top_row)) continue;
    for (int i = threadIdx.x; i < fx; i += kLanesX) {
      if (a == 0 && (i == 0 || i == g.nx)) continue;
      const int index = i + fx * row;
      if (!face_open(g, a, i, j, k)) {  // closed obstacle face stays 0
        dst[index] = (T)0.0;
        continue;
      }
      const int right = i + g.nx * (j + g.ny * k);
      const double difference = p[right] - p[right - stride];
      double value;
      if (sizeof(T) == 4) {
        value = (double)dst[index] - dt * acc * difference;
        if (!isfinite(value) || fabs(value) > (double)FLT_MAX) {
          bad |= kFlagArithmetic;
          continue;
        }
      } else {
        const double gradient = acc * difference;
        value = (double)dst[index] - dt * gradient;
        if (!isfinite(value)) {
          bad |= kFlagArithmetic;
          continue;
        }
      }
      dst[index] = (T)value;
    }
  }
  if (bad) atomicOr(flags, bad);
}
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_correct(DevGrid g, const double* __restrict__ p, double dt, T* __restrict__ u,
          T* __restrict__ v, T* __restrict__ w, unsigned* __restrict__ flags) {
  b_correct<T>(hw_block(), g, p, dt, u, v, w, flags);
}

// correct_candidate_in_place with the three components in one pass over the
// cells: cell (i, j, k) takes its three lower faces (the face is interior
// when its own index along the axis is >= 1; the upper wall faces belong to
// no cell and stay untouched), so p and its lower neighbours are read once
// for all three components (k_correct reads p once per component).  Each
// face gets b_correct's expressions and checks.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_correct_cells(DevGrid g, const double* __restrict__ p, double dt, T* __restrict__ u, T* __restrict__ v,
                T* __restrict__ w, unsigned* __restrict__ flags) {
  const int nx = g.nx, ny = g.ny;
  const RowRange rr = block_rows(ny * g.nz);
  unsigned bad = 0;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / ny, j = row - k * ny;
    for (int i = threadIdx.x; i < nx; i += kLanesX) {
      const int c = i + nx * row;
      const double pc = p[c];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int pa = a == 0 ? i : (a == 1 ? j : k);
        if (pa == 0) continue;
        T* dst = a == 0 ? u : (a == 1 ? v : w);
        const int index = a == 0 ? i + (nx + 1) * row : (a == 1 ? i + nx * (j + (ny + 1) * k) : c);
        if (!face_open(g, a, i, j, k)) {  // closed obstacle face stays 0
          dst[index] = (T)0.0;
          continue;
        }
        const int stride = a == 0 ? 1 : (a == 1 ? nx : nx * ny);
        const double difference = pc - p[c - stride];
        const double acc = g.acc[a];
        double value;
        if (sizeof(T) == 4) {
          value = (double)dst[index] - dt * acc * difference;
          if (!isfinite(value) || fabs(value) > (double)FLT_MAX) {
            bad |= kFlagArithmetic;
            continue;
          }
        } else {
          const double gradient = acc * difference;
          value = (double)dst[index] - dt * gradient;
          if (!isfinite(value)) {
            bad |= kFlagArithmetic;
            continue;
          }
        }
        dst[index] = (T)value;
      }
    }
  }
  if (bad) atomicOr(flags, bad);
}

// divergence (src/operator.rs:513-568) and its all-cell maximum
// (actual_divergence_max, src/pressure.rs:330-342):
// div = (u_hi-u_lo)/h0 + (v_hi-v_lo)/h1 + (w_hi-w_lo)/h2.
template <typename T>
__device__ __forceinline__ void
b_divergence_max(VBlock vb, const DevGrid& g, const T* __restrict__ u, const T* __restrict__ v,
                 const T* __restrict__ w, Reduce red) {
  const int nx = g.nx, ny = g.ny;
  const RowRange rr = block_rows(g.ny * g.nz, vb);
  double m[1] = {0.0};
  unsigned bad = 0;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / ny, j = row - k * ny;
    for (int i = threadIdx.x; i < nx; i += kLanesX) {
      const int xu = i + (nx + 1) * row;
      const int yv = i + nx * (j + (ny + 1) * k);
      const int zw = i + nx * row;
      const double value = over_h(g, 0, (double)u[xu + 1] - (double)u[xu]) +
                           over_h(g, 1, (double)v[yv + nx] - (double)v[yv]) +
                           over_h(g, 2, (double)w[zw + nx * ny] - (double)w[zw]);
      if (!isfinite(value)) bad |= kFlagArithmetic;
      m[0] = fmax(m[0], fabs(value));
    }
  }
  if (bad) atomicOr(red.flags, bad);
  DD dd[1] = {DD{0.0, 0.0}};
  block_publish<1, 1>(dd, m, red.pdd, red.pmx, vb);
}
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_divergence_max(DevGrid g, const T* __restrict__ u, const T* __restrict__ v,
                 const T* __restrict__ w, Reduce red) {
  b_divergence_max<T>(hw_block(), g, u, v, w, red);
}

// velocity_rate (src/simulation.rs:1415-1429): sum_d max|u_d| / h_d.
template <typename T>
__device__ __forceinline__ void
b_velocity_rate(VBlock vb, const DevGrid& g, const T* __restrict__ u, const T* __restrict__ v,
                const T* __restrict__ w, Reduce red) {
  double m[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < 3; ++a) {
    const T* src = a == 0 ? u : (a == 1 ? v : w);
    const int fx = g.nx + (a == 0), fy = g.ny + (a == 1), fz = g.nz + (a == 2);
    const RowRange rr = block_rows(fy * fz, vb);
    for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock)
      for (int i = threadIdx.x; i < fx; i += kLanesX)
        m[a] = fmax(m[a], fabs((double)src[i + fx * row]));
  }
  DD dd[1] = {DD{0.0, 0.0}};
  block_publish<1, 3>(dd, m, red.pdd, red.pmx, vb);
}
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_velocity_rate(DevGrid g, const T* __restrict__ u, const T* __restrict__ v,
                const T* __restrict__ w, Reduce red) {
  b_velocity_rate<T>(hw_block(), g, u, v, w, red);
}

// divergence maximum and velocity rate in one pass over the corrected
// faces (the scene step needs both): per cell the divergence of
// b_divergence_max, and the maxima of |u|, |v|, |w| over every face (each
// cell takes its three lower faces, the last cell of a row / column /
// pillar also the upper wall face), as b_velocity_rate.  Maxima do not
// depend on the order they are taken in, so the values are those of the
// two kernels.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_divergence_rate(DevGrid g, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
                  Reduce red) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const RowRange rr = block_rows(ny * nz);
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  DD e{0.0, 0.0};
  unsigned bad = 0;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / ny, j = row - k * ny;
    for (int i = threadIdx.x; i < nx; i += kLanesX) {
      const int xu = i + (nx + 1) * row;
      const int yv = i + nx * (j + (ny + 1) * k);
      const int zw = i + nx * row;
      const double u0 = (double)u[xu], u1 = (double)u[xu + 1];
      const double v0 = (double)v[yv], v1 = (double)v[yv + nx];
      const double w0 = (double)w[zw], w1 = (double)w[zw + nx * ny];
      const double value = over_h(g, 0, u1 - u0) + over_h(g, 1, v1 - v0) + over_h(g, 2, w1 - w0);
      if (!isfinite(value)) bad |= kFlagArithmetic;
      m[0] = fmax(m[0], fabs(value));
      m[1] = fmax(m[1], fabs(u0));
      m[2] = fmax(m[2], fabs(v0));
      m[3] = fmax(m[3], fabs(w0));
      if (i + 1 == nx) m[1] = fmax(m[1], fabs(u1));
      if (j + 1 == ny) m[2] = fmax(m[2], fabs(v1));
      if (k + 1 == nz) m[3] = fmax(m[3], fabs(w1));
      // squared faces for the kinetic energy (each face once)
      dd_acc(e, u0 * u0);
      dd_acc(e, v0 * v0);
      dd_acc(e, w0 * w0);
      if (i + 1 == nx) dd_acc(e, u1 * u1);
      if (j + 1 == ny) dd_acc(e, v1 * v1);
      if (k + 1 == nz) dd_acc(e, w1 * w1);
    }
  }
  if (bad) atomicOr(red.flags, bad);
  DD dd[1] = {e};
  block_publish<1, 4>(dd, m, red.pdd, red.pmx);
}

// correct_candidate_in_place followed by the divergence gate, the velocity
// rate, the component maxima and the squared-face sum over the corrected
// faces, in one pass and out of place (in -> out): cell (i, j, k) corrects
// its three lower faces (k_correct_cells's expressions and checks; a wall
// face is copied, a closed obstacle face becomes 0) and writes them, and
// also evaluates its three upper faces -- by the same expressions from the
// same inputs, so the same bits the neighbouring cell writes -- to form its
// divergence without reading corrected faces back.  Upper wall faces belong
// to no cell: the cell below them copies them.  The loop, the per-thread
// accumulation order and the partials are k_divergence_rate's, so the
// results are its results on the corrected faces, bit for bit.
template <typename T>
__device__ __forceinline__ double corrected_face(T x, double p_right, double p_left, double dt, double acc,
                                                 unsigned& bad) {
  const double difference = p_right - p_left;
  double value;
  if (sizeof(T) == 4) {
    value = (double)x - dt * acc * difference;
    if (!isfinite(value) || fabs(value) > (double)FLT_MAX) bad |= kFlagArithmetic;
  } else {
    const double gradient = acc * difference;
    value = (double)x - dt * gradient;
    if (!isfinite(value)) bad |= kFlagArithmetic;
  }
  return (double)(T)value;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_correct_divergence_rate(DevGrid g, const double* __restrict__ p, double dt, const T* __restrict__ iu,
                          const T* __restrict__ iv, const T* __restrict__ iw, T* __restrict__ ou,
                          T* __restrict__ ov, T* __restrict__ ow, Reduce red) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int sy = nx, sz = nx * ny, last = nx * ny * nz - 1;
  const RowRange rr = block_rows(ny * nz);
  const double a0 = g.acc[0], a1 = g.acc[1], a2 = g.acc[2];
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  DD e{0.0, 0.0};
  unsigned bad = 0;
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / ny, j = row - k * ny;
    for (int i = threadIdx.x; i < nx; i += kLanesX) {
      const int xu = i + (nx + 1) * row;
      const int yv = i + nx * (j + (ny + 1) * k);
      const int c = i + nx * row;
      // every load first, at in-range addresses (a wall's neighbour index is
      // clamped; its value is not used), so they are all in flight at once
      const T fu0 = iu[xu], fu1 = iu[xu + 1], fv0 = iv[yv], fv1 = iv[yv + nx], fw0 = iw[c], fw1 = iw[c + sz];
      const double pc = p[c];
      const double pxl = p[max(c - 1, 0)], pyl = p[max(c - sy, 0)], pzl = p[max(c - sz, 0)];
      const double pxh = p[min(c + 1, last)], pyh = p[min(c + sy, last)], pzh = p[min(c + sz, last)];
      // open lower faces of this cell and of its +x/+y/+z neighbours
      bool ox0 = true, oy0 = true, oz0 = true, ox1 = true, oy1 = true, oz1 = true;
      if (g.mask) {
        const unsigned mc = g.mask[c];
        ox0 = mc & 1u;
        oy0 = (mc >> 2) & 1u;
        oz0 = (mc >> 4) & 1u;
        ox1 = g.mask[min(c + 1, last)] & 1u;
        oy1 = (g.mask[min(c + sy, last)] >> 2) & 1u;
        oz1 = (g.mask[min(c + sz, last)] >> 4) & 1u;
      }
      const bool wx0 = i == 0, wy0 = j == 0, wz0 = k == 0;
      const bool wx1 = i + 1 == nx, wy1 = j + 1 == ny, wz1 = k + 1 == nz;
      // each interior open lower face's own check (its owner's, as in
      // k_correct_cells); an upper face is checked by the cell above it
      unsigned bu = 0, bv = 0, bw = 0, b = 0;
      const double cu0 = corrected_face<T>(fu0, pc, pxl, dt, a0, bu);
      const double cv0 = corrected_face<T>(fv0, pc, pyl, dt, a1, bv);
      const double cw0 = corrected_face<T>(fw0, pc, pzl, dt, a2, bw);
      if (!wx0 && ox0) bad |= bu;
      if (!wy0 && oy0) bad |= bv;
      if (!wz0 && oz0) bad |= bw;
      const double u0 = wx0 ? (double)fu0 : (ox0 ? cu0 : 0.0);
      const double v0 = wy0 ? (double)fv0 : (oy0 ? cv0 : 0.0);
      const double w0 = wz0 ? (double)fw0 : (oz0 ? cw0 : 0.0);
      ou[xu] = (T)u0;
      ov[yv] = (T)v0;
      ow[c] = (T)w0;
      const double cu1 = corrected_face<T>(fu1, pxh, pc, dt, a0, b);
      const double cv1 = corrected_face<T>(fv1, pyh, pc, dt, a1, b);
      const double cw1 = corrected_face<T>(fw1, pzh, pc, dt, a2, b);
      (void)b;
      const double u1 = wx1 ? (double)fu1 : (ox1 ? cu1 : 0.0);
      const double v1 = wy1 ? (double)fv1 : (oy1 ? cv1 : 0.0);
      const double w1 = wz1 ? (double)fw1 : (oz1 ? cw1 : 0.0);
      if (wx1) ou[xu + 1] = fu1;
      if (wy1) ov[yv + nx] = fv1;
      if (wz1) ow[c + sz] = fw1;
      const double value = over_h(g, 0, u1 - u0) + over_h(g, 1, v1 - v0) + over_h(g, 2, w1 - w0);
      if (!isfinite(value)) bad |= kFlagArithmetic;
      m[0] = fmax(m[0], fabs(value));
      m[1] = fmax(m[1], fabs(u0));
      m[2] = fmax(m[2], fabs(v0));
      m[3] = fmax(m[3], fabs(w0));
      if (wx1) m[1] = fmax(m[1], fabs(u1));
      if (wy1) m[2] = fmax(m[2], fabs(v1));
      if (wz1) m[3] = fmax(m[3], fabs(w1));
      dd_acc(e, u0 * u0);
      dd_acc(e, v0 * v0);
      dd_acc(e, w0 * w0);
      if (wx1) dd_acc(e, u1 * u1);
      if (wy1) dd_acc(e, v1 * v1);
      if (wz1) dd_acc(e, w1 * w1);
    }
  }
  if (bad) atomicOr(red.flags, bad);
  DD dd[1] = {e};
  block_publish<1, 4>(dd, m, red.pdd, red.pmx);
}

// ------------------------------------------------------------ host launchers
template <typename T>
void launch_advect(const DevGrid& g, const void* u, const void* v, const void* w, double dt,
                   void* ou, void* ov, void* ow, unsigned* flags, hipStream_t s) {
  const int rows = (g.ny + 1) * (g.nz + 1);
  dim3 grid(grid_blocks(rows), 3), block(kLanesX, kRowsPerBlock);
  const bool off32 = face_offsets_32(g, sizeof(T));
  if constexpr (sizeof(T) == 4) {
    if (g.lattice) {
      if (off32)
        launch(k_advect_velocity<T, true, 1>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, dt,
               (T*)ou, (T*)ov, (T*)ow, flags);
      else
        launch(k_advect_velocity<T, false, 1>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, dt,
               (T*)ou, (T*)ov, (T*)ow, flags);
      return;
    }
  }
  if (off32)
    launch(k_advect_velocity<T, true>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, dt,
                                                  (T*)ou, (T*)ov, (T*)ow, flags);
  else
    launch(k_advect_velocity<T, false>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, dt,
                                                   (T*)ou, (T*)ov, (T*)ow, flags);
}

template <typename T>
void launch_validate(const DevGrid& g, const void* u, const void* v, const void* w,
                     int check_walls, unsigned* flags, hipStream_t s) {
  const int rows = (g.ny + 1) * (g.nz + 1);
  dim3 grid(grid_blocks(rows), 3), block(kLanesX, kRowsPerBlock);
  k_validate_faces<T><<<grid, block, 0, s>>>(g, (const T*)u, (const T*)v, (const T*)w,
                                             check_walls, flags);
}

template <typename T>
void launch_rhs(const DevGrid& g, const void* u, const void* v, const void* w, double dt,
                double* out, unsigned* flags, hipStream_t s, bool validate, const Reduce* sums) {
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  if (sums)
    launch(k_build_rhs<T, true, true>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, dt, out, flags,
           sums->pdd, sums->pmx);
  else if (validate)
    launch(k_build_rhs<T, true>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, dt, out, flags,
           nullptr, nullptr);
  else
    launch(k_build_rhs<T, false>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, dt, out, flags,
           nullptr, nullptr);
}

template <typename T>
void launch_correct(const DevGrid& g, const double* p, double dt, void* u, void* v, void* w,
                    unsigned* flags, hipStream_t s) {
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  launch(k_correct_cells<T>, grid, block, 0, s, g, p, dt, (T*)u, (T*)v, (T*)w, flags);
}

template <typename T>
void launch_divergence_max(const DevGrid& g, const void* u, const void* v, const void* w,
                           const Reduce& red, hipStream_t s) {
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  launch(k_divergence_max<T>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, red);
  launch_finish<1, 1>(red.pdd, red.pmx, grid.x, FinStore{red.out, 0, 0.0, 0.0, 0.0}, s);
}

template <typename T>
void launch_velocity_rate(const DevGrid& g, const void* u, const void* v, const void* w,
                          const Reduce& red, hipStream_t s) {
  const int rows = (g.ny + 1) * (g.nz + 1);
  dim3 grid(grid_blocks(rows)), block(kLanesX, kRowsPerBlock);
  k_velocity_rate<T><<<grid, block, 0, s>>>(g, (const T*)u, (const T*)v, (const T*)w, red);
  launch_finish<1, 3>(red.pdd, red.pmx, grid.x, FinStore{red.out, 1, g.h[0], g.h[1], g.h[2]}, s);
}

// out[0] = max|div u|, out[1] = velocity rate, out[2..4] = max|u|, |v|, |w|
// (the slots launch_divergence_max into out and launch_velocity_rate into
// out + 1 fill), out[5] = sum of the squared faces (double-double).
template <typename T>
void launch_divergence_rate(const DevGrid& g, const void* u, const void* v, const void* w, const Reduce& red,
                            hipStream_t s) {
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  launch(k_divergence_rate<T>, grid, block, 0, s, g, (const T*)u, (const T*)v, (const T*)w, red);
  launch_finish<1, 4>(red.pdd, red.pmx, grid.x, FinStore{red.out, 3, g.h[0], g.h[1], g.h[2]}, s);
}

template <typename T>
void launch_correct_divergence_rate(const DevGrid& g, const double* p, double dt, void* const in[3], void* out[3],
                                    const Reduce& red, hipStream_t s) {
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  launch(k_correct_divergence_rate<T>, grid, block, 0, s, g, p, dt, (const T*)in[0], (const T*)in[1],
         (const T*)in[2], (T*)out[0], (T*)out[1], (T*)out[2], red);
  launch_finish<1, 4>(red.pdd, red.pmx, grid.x, FinStore{red.out, 3, g.h[0], g.h[1], g.h[2]}, s);
}

void launch_advect_tracer(const DevGrid& g, const float* old, const float* u, const float* v,
                          const float* w, double dt, float* out, unsigned* flags, hipStream_t s,
                          const TracerBox* box) {
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  const TracerBox all{{0, 0, 0}, {g.nx, g.ny, g.nz}};
  const TracerBox b = box ? *box : all;
  const bool off32 = face_offsets_32(g, sizeof(float));
  if (g.lattice && off32)
    launch(k_advect_tracer<true, true>, grid, block, 0, s, g, old, u, v, w, dt, out, flags, b);
  else if (g.lattice)
    launch(k_advect_tracer<false, true>, grid, block, 0, s, g, old, u, v, w, dt, out, flags, b);
  else if (off32)
    launch(k_advect_tracer<true, false>, grid, block, 0, s, g, old, u, v, w, dt, out, flags, b);
  else
    launch(k_advect_tracer<false, false>, grid, block, 0, s, g, old, u, v, w, dt, out, flags, b);
}

void launch_tracer_bounds(const DevGrid& g, const float* t, unsigned* bounds, unsigned* flags, hipStream_t s) {
  k_tracer_bounds_init<<<1, 64, 0, s>>>(g, bounds);
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  k_tracer_bounds<<<grid, block, 0, s>>>(g, t, bounds, flags);
}

void launch_tracer_sum(const DevGrid& g, const float* t, const Reduce& red, hipStream_t s) {
  dim3 grid(grid_blocks(g.ny * g.nz)), block(kLanesX, kRowsPerBlock);
  launch(k_tracer_sum, grid, block, 0, s, g, t, red);
  launch_finish<1, 1>(red.pdd, red.pmx, grid.x, FinStore{red.out, 4, 0.0, 0.0, 0.0}, s);
}

// Obstacle faces (both cells not fluid on either side) set to 0; walls and
// open faces untouched (src/obstacle_pressure.rs:566-578 keeps them at 0).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_zero_blocked(DevGrid g, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w) {
  const int a = blockIdx.y;
  T* dst = a == 0 ? u : (a == 1 ? v : w);
  const int fx = g.nx + (a == 0), fy = g.ny + (a == 1), fz = g.nz + (a == 2);
  const RowRange rr = block_rows(fy * fz);
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock) {
    const int k = row / fy, j = row - k * fy;
    for (int i = threadIdx.x; i < fx; i += kLanesX) {
      const int p[3] = {i, j, k};
      const int top = (a == 0 ? fx : (a == 1 ? fy : fz)) - 1;
      if (p[a] == 0 || p[a] == top) continue;
      if (!face_open(g, a, i, j, k)) dst[i + fx * row] = (T)0.0;
    }
  }
}

template <typename T>
void launch_zero_blocked(const DevGrid& g, void* u, void* v, void* w, hipStream_t s) {
  if (!g.mask) return;
  dim3 grid(grid_blocks((g.ny + 1) * (g.nz + 1)), 3), block(kLanesX, kRowsPerBlock);
  k_zero_blocked<T><<<grid, block, 0, s>>>(g, (T*)u, (T*)v, (T*)w);
}

// ---------------------------------------------------- batched (f32 faces)
// Records: BFaceArgs; `state_ok` records skip scenes whose step has failed.

#define RH_BATCH_PROLOGUE(ARGS)                                  \
  const int s_ = batch_scene(prefix, n, blockIdx.x);            \
  const ARGS& a = A[s_];                                        \
  const VBlock vb = batch_block(prefix, s_);

__global__ void __launch_bounds__(kBlock)
kb_validate(const BFaceArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BFaceArgs)
  b_validate_faces<float>(vb, *a.g, a.u, a.v, a.w, a.mode, a.flags);
}
// Lattice and general scenes go to separate launches (kLattice 1 or 0), so
// each form fits the one-scene kernel's occupancy; the grid record is
// copied into scalar registers (load_record) rather than read per use.
template <bool kOff32, int kLattice>
__global__ void __launch_bounds__(kBlock, kLattice == 1 ? kAdvectLatticeWaves : kAdvectWaves)
kb_advect(const BFaceArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BFaceArgs)
  const DevGrid g = load_record(a.g);
  b_advect_velocity<float, kOff32, kLattice, true>(vb, (int)blockIdx.y, g, a.u, a.v, a.w, *a.dt, a.ou, a.ov, a.ow,
                                             a.flags);
}
__global__ void __launch_bounds__(kBlock)
kb_rhs(const BFaceArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BFaceArgs)
  b_build_rhs<float>(vb, *a.g, a.u, a.v, a.w, *a.dt, a.cells, a.flags);
}
// correct / divergence / rate run only for scenes whose solve was accepted
// (ctl->status == RH_OK after the quality gate).
__global__ void __launch_bounds__(kBlock)
kb_correct(const BFaceArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BFaceArgs)
  if (a.ctl->status != RH_OK) return;
  b_correct<float>(vb, *a.g, a.cells, *a.dt, a.ou, a.ov, a.ow, a.flags);
}
__global__ void __launch_bounds__(kBlock)
kb_divergence_max(const BFaceArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BFaceArgs)
  if (a.ctl->status != RH_OK) return;
  Reduce red;
  red.pdd = a.pdd;
  red.pmx = a.pmx;
  red.out = nullptr;
  red.flags = a.flags;
  b_divergence_max<float>(vb, *a.g, a.u, a.v, a.w, red);
}
__global__ void __launch_bounds__(kBlock)
kb_velocity_rate(const BFaceArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BFaceArgs)
  if (a.ctl->status != RH_OK) return;
  Reduce red;
  red.pdd = a.pdd;
  red.pmx = a.pmx;
  red.out = nullptr;
  red.flags = a.flags;
  b_velocity_rate<float>(vb, *a.g, a.u, a.v, a.w, red);
}

static dim3 face_batch_grid(const BatchLaunch& bl) { return dim3(bl.blocks, 3); }
static dim3 block_shape() { return dim3(kLanesX, kRowsPerBlock); }

void launch_b_validate(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_validate<<<face_batch_grid(bl), block_shape(), 0, s>>>((const BFaceArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_advect(const BatchLaunch& bl, hipStream_t s, bool off32, bool lattice) {
  if (!bl.blocks) return;
  const BFaceArgs* A = (const BFaceArgs*)bl.args;
  if (off32 && lattice)
    kb_advect<true, 1><<<face_batch_grid(bl), block_shape(), 0, s>>>(A, bl.prefix, bl.n);
  else if (off32)
    kb_advect<true, 0><<<face_batch_grid(bl), block_shape(), 0, s>>>(A, bl.prefix, bl.n);
  else if (lattice)
    kb_advect<false, 1><<<face_batch_grid(bl), block_shape(), 0, s>>>(A, bl.prefix, bl.n);
  else
    kb_advect<false, 0><<<face_batch_grid(bl), block_shape(), 0, s>>>(A, bl.prefix, bl.n);
}
void launch_b_rhs(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks) kb_rhs<<<bl.blocks, block_shape(), 0, s>>>((const BFaceArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_correct(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_correct<<<face_batch_grid(bl), block_shape(), 0, s>>>((const BFaceArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_divergence_max(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_divergence_max<<<bl.blocks, block_shape(), 0, s>>>((const BFaceArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_velocity_rate(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_velocity_rate<<<bl.blocks, block_shape(), 0, s>>>((const BFaceArgs*)bl.args, bl.prefix, bl.n);
}

#define RH_INSTANTIATE(T)                                                                     \
  template void launch_zero_blocked<T>(const DevGrid&, void*, void*, void*, hipStream_t);      \
  template void launch_advect<T>(const DevGrid&, const void*, const void*, const void*,        \
                                 double, void*, void*, void*, unsigned*, hipStream_t);         \
  template void launch_validate<T>(const DevGrid&, const void*, const void*, const void*, int, \
                                   unsigned*, hipStream_t);                                    \
  template void launch_rhs<T>(const DevGrid&, const void*, const void*, const void*, double,   \
                              double*, unsigned*, hipStream_t, bool, const Reduce*);           \
  template void launch_correct<T>(const DevGrid&, const double*, double, void*, void*, void*,  \
                                  unsigned*, hipStream_t);                                     \
  template void launch_divergence_max<T>(const DevGrid&, const void*, const void*, const void*, \
                                         const Reduce&, hipStream_t);                          \
  template void launch_velocity_rate<T>(const DevGrid&, const void*, const void*, const void*,  \
                                        const Reduce&, hipStream_t);                           \
  template void launch_divergence_rate<T>(const DevGrid&, const void*, const void*, const void*, \
                                          const Reduce&, hipStream_t);                         \
  template void launch_correct_divergence_rate<T>(const DevGrid&, const double*, double,      \
                                                  void* const[3], void* [3], const Reduce&,   \
                                                  hipStream_t);
RH_INSTANTIATE(float)
RH_INSTANTIATE(double)
#undef RH_INSTANTIATE

}  // namespace rh
