// Host runtime of rheon_amd: geometry validation, device scratch, the
// device PCG solver (PressureWorkspace on the GPU) and stream helpers.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"
#include "profiler.hpp"

namespace rh {

static thread_local std::string g_last_error;

void set_last_error(const char* what, const char* where) {
  g_last_error = std::string(what ? what : "") + (where ? std::string(" [") + where + "]" : "");
}
const char* last_error() { return g_last_error.c_str(); }

int grid_blocks(int rows) {
  const long long groups = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
  const long long blocks = (groups + kGroupsPerBlock - 1) / kGroupsPerBlock;
  return blocks < 1 ? 1 : (blocks > kMaxBlocks ? kMaxBlocks : (int)blocks);
}

// One cell per thread (coarse levels are latency-bound: the more waves in
// flight the better), up to the block cap.
int level_blocks(long long cells) {
  const long long blocks = (cells + kBlock - 1) / kBlock;
  return blocks < 1 ? 1 : (blocks > kMaxBlocks ? kMaxBlocks : (int)blocks);
}

int quad_blocks(int quads_per_row, int rows) {
  const long long quads = (long long)quads_per_row * rows;
  const long long blocks = (quads + kGroupsPerBlock * kBlock - 1) / (kGroupsPerBlock * kBlock);
  return blocks < 1 ? 1 : (blocks > kMaxBlocks ? kMaxBlocks : (int)blocks);
}

// GridGeometry::new (src/geometry.rs:114-191), plus the 32-bit indexing
// bound of these kernels (every cell and face index must fit in int32).
rh_status validate_grid(const rh_grid* g, size_t* cells_out) {
  if (!g) return RH_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a) {
    if (g->counts[a] == 0) return RH_GEOMETRY;
    if (!std::isfinite(g->spacing[a]) || g->spacing[a] <= 0.0) return RH_GEOMETRY;
    if (!std::isfinite(g->origin[a])) return RH_GEOMETRY;
  }
  unsigned __int128 cells = (unsigned __int128)g->counts[0] * g->counts[1];
  if (cells > (unsigned __int128)INT64_MAX) return RH_GEOMETRY;
  cells *= g->counts[2];
  if (cells > (unsigned __int128)INT64_MAX) return RH_GEOMETRY;
  const double volume = g->spacing[0] * g->spacing[1] * g->spacing[2];
  if (!std::isfinite(volume) || volume <= 0.0) return RH_GEOMETRY;
  for (int a = 0; a < 3; ++a) {
    if (g->counts[a] > (1ULL << 52)) return RH_GEOMETRY;
    const double n = (double)g->counts[a];
    const double h = g->spacing[a], o = g->origin[a];
    const double extent = n * h;
    const double upper = o + extent;
    const double scale = std::fmax(std::fmax(std::fabs(o), std::fabs(upper)), extent);
    if (!std::isfinite(extent) || !std::isfinite(scale) || !std::isnormal(0.5 * h) ||
        h / scale < 32.0 * DBL_EPSILON)
      return RH_GEOMETRY;
    const double first_center = o + 0.5 * h;
    const double last_center = o + (n - 0.5) * h;
    if (!std::isfinite(upper) || upper <= o || first_center <= o || last_center >= upper ||
        (g->counts[a] > 1 && o + h <= o))
      return RH_GEOMETRY;
  }
  if (cells_out) *cells_out = (size_t)cells;
  return RH_OK;
}

rh_status check_indexable(const rh_grid* g) {
  // Largest face array: (n_a + 1) * n_b * n_c must fit int32.
  for (int a = 0; a < 3; ++a) {
    unsigned __int128 f = 1;
    for (int d = 0; d < 3; ++d) f *= g->counts[d] + (d == a ? 1 : 0);
    if (f >= (unsigned __int128)INT32_MAX) return RH_UNSUPPORTED;
  }
  return RH_OK;
}

// PressureOperator::new coefficients (src/operator.rs:58-97).
DevGrid make_dev_grid(const rh_grid& g, double density) {
  DevGrid d;
  std::memset(&d, 0, sizeof(d));
  d.nx = (int)g.counts[0];
  d.ny = (int)g.counts[1];
  d.nz = (int)g.counts[2];
  d.div_q = make_fast_div(d.nx / 4);
  d.div_y = make_fast_div(d.ny);
  d.div_x = make_fast_div(d.nx);
  d.div_x1 = make_fast_div(d.nx + 1);
  d.div_y1 = make_fast_div(d.ny + 1);
  d.ncomp = 1;
  for (int a = 0; a < 3; ++a) {
    d.h[a] = g.spacing[a];
    d.origin[a] = g.origin[a];
    d.upper[a] = g.origin[a] + (double)g.counts[a] * g.spacing[a];
  }
  d.volume = g.spacing[0] * g.spacing[1] * g.spacing[2];
  for (int a = 0; a < 3; ++a) {
    int e = 0;
    const double m = std::frexp(g.spacing[a], &e);
    const double inv = 1.0 / g.spacing[a];
    d.inv_h[a] = inv;
    d.pow2[a] = (m == 0.5 && std::isnormal(inv) && g.spacing[a] * inv == 1.0) ? 1 : 0;
  }
  // lattice: the base coordinate of field kernels (field_kernels.hip,
  // base_coords) recovers every lattice coordinate p + s exactly.
  d.lattice = 1;
  for (int a = 0; a < 3 && d.lattice; ++a)
    for (uint64_t p = 0; p <= g.counts[a] && d.lattice; ++p)
      for (int half = 0; half < (p < g.counts[a] ? 2 : 1); ++half) {
        const double s = (double)p + (half ? 0.5 : 0.0);
        double x = d.origin[a] + s * d.h[a];
        x = std::fmin(std::fmax(x, d.origin[a]), d.upper[a]);
        const double q = d.pow2[a] ? (x - d.origin[a]) * d.inv_h[a] : (x - d.origin[a]) / d.h[a];
        if (q != s) {
          d.lattice = 0;
          break;
        }
      }
  const double* h = g.spacing;
  d.area[0] = h[1] * h[2];
  d.area[1] = h[0] * h[2];
  d.area[2] = h[0] * h[1];
  for (int a = 0; a < 3; ++a) {
    d.acc[a] = 1.0 / (density * h[a]);
    d.w[a] = d.area[a] * d.acc[a];
  }
  for (int mask = 0; mask < 64; ++mask) {
    double value = 0.0;
    for (int side = 0; side < 6; ++side)
      if (mask & (1 << side)) value += d.w[side / 2];
    d.diag[mask] = value;
  }
  return d;
}

rh_status check_density(const rh_grid& g, double density) {
  if (!std::isfinite(density) || density <= 0.0) return RH_INVALID_DENSITY;
  DevGrid d = make_dev_grid(g, density);
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(d.area[a]) || d.area[a] <= 0.0 || !std::isfinite(d.acc[a]) ||
        d.acc[a] <= 0.0 || !std::isfinite(d.w[a]) || d.w[a] <= 0.0)
      return RH_INVALID_ARGUMENT;  // OperatorError::InvalidCoefficient
  return RH_OK;
}

// Solid obstacles (src/obstacle_pressure.rs:224-323 with all-or-nothing
// apertures): open-face mask per cell from a solid flag per cell, the fluid
// regions and one gauge per region (the first at *gauge).
rh_status build_obstacle_mask(const rh_grid& g, const uint8_t* solid, std::vector<uint8_t>& mask,
                              int* gauge, ObstacleComponents* comps) {
  const int64_t nx = (int64_t)g.counts[0], ny = (int64_t)g.counts[1], nz = (int64_t)g.counts[2];
  const int64_t n = nx * ny * nz;
  mask.assign((size_t)n, 0);
  int64_t first = -1;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t c = i + nx * (j + ny * k);
        if (solid[c]) continue;
        if (first < 0) first = c;
        uint8_t m = kFluidBit;
        if (i > 0 && !solid[c - 1]) m |= 1;
        if (i + 1 < nx && !solid[c + 1]) m |= 2;
        if (j > 0 && !solid[c - nx]) m |= 4;
        if (j + 1 < ny && !solid[c + nx]) m |= 8;
        if (k > 0 && !solid[c - nx * ny]) m |= 16;
        if (k + 1 < nz && !solid[c + nx * ny]) m |= 32;
        mask[(size_t)c] = m;
      }
  if (first < 0) {
    set_last_error("every cell is solid", "build_obstacle_mask");
    return RH_INVALID_ARGUMENT;
  }
  // Fluid regions: seeds in index order, breadth-first over open faces
  // (src/static_obstacle.rs:165-210).  Each region's gauge is its largest-
  // volume cell, lowest index on ties (src/obstacle_pressure.rs:310-323): with
  // every fluid volume equal, its seed.  The gauge cells carry kGaugeBit.
  std::vector<uint8_t> label((size_t)n, kNoComponent);
  std::vector<int64_t> queue;
  queue.reserve(1024);
  const int64_t off[6] = {-1, 1, -nx, nx, -nx * ny, nx * ny};
  std::vector<int> gauges;
  for (int64_t seed = first; seed < n; ++seed) {
    if (!(mask[(size_t)seed] & kFluidBit) || label[(size_t)seed] != kNoComponent) continue;
    if ((int)gauges.size() == kMaxComponents) {
      set_last_error("the fluid cells form more than 254 separate regions", "build_obstacle_mask");
      return RH_UNSUPPORTED;
    }
    const uint8_t id = (uint8_t)gauges.size();
    gauges.push_back((int)seed);
    mask[(size_t)seed] |= kGaugeBit;
    label[(size_t)seed] = id;
    queue.clear();
    queue.push_back(seed);
    for (size_t head = 0; head < queue.size(); ++head) {
      const int64_t c = queue[head];
      const uint8_t m = mask[(size_t)c];
      for (int side = 0; side < 6; ++side)
        if ((m >> side) & 1) {
          const int64_t d = c + off[side];
          if (label[(size_t)d] == kNoComponent) {
            label[(size_t)d] = id;
            queue.push_back(d);
          }
        }
    }
  }
  *gauge = (int)first;
  if (comps) {
    comps->gauges = gauges;
    if (gauges.size() > 1)
      comps->labels.swap(label);
    else
      comps->labels.clear();
  }
  return RH_OK;
}

// ------------------------------------------------------------------ Scratch

rh_status Scratch::init() {
  if (stream) return RH_OK;
  RH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const size_t bytes = kDeviceBytes;
  RH_HIP(hipMalloc(&base, bytes));
  RH_HIP(hipMemsetAsync(base, 0, bytes, stream));
  char* p = (char*)base;
  pdd = (DD*)p;
  p += sizeof(DD) * 4 * kMaxBlocks;
  pmx = (double*)p;
  p += sizeof(double) * 4 * kMaxBlocks;
  counters = (unsigned*)p;
  p += sizeof(unsigned) * 64;
  scal = (double*)p;
  p += sizeof(double) * 64;
  p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
  ctl = (PcgControl*)p;
  flags = counters + 32;
  RH_HIP(hipHostMalloc(&host, 4096, hipHostMallocDefault));
  RH_HIP(hipHostMalloc(&progress, 64, hipHostMallocMapped));
  RH_HIP(hipHostGetDevicePointer((void**)&progress_dev, progress, 0));
  *(volatile int*)progress = 0;
  RH_HIP(hipEventCreate(&ev0));
  RH_HIP(hipEventCreate(&ev1));
  RH_HIP(hipEventCreateWithFlags(&poll_ev[0], hipEventDisableTiming));
  RH_HIP(hipEventCreateWithFlags(&poll_ev[1], hipEventDisableTiming));
  RH_HIP(hipStreamSynchronize(stream));
  bytes_ = bytes;
  return RH_OK;
}

void Scratch::release() {
  if (host) (void)hipHostFree(host);
  if (progress) (void)hipHostFree(progress);
  if (base) (void)hipFree(base);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  for (auto& e : poll_ev)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
  *this = Scratch();
}

Reduce Scratch::reduce(int out_slot) {
  Reduce r;
  r.pdd = pdd;
  r.pmx = pmx;
  r.out = scal + out_slot;
  r.flags = flags;
  return r;
}

rh_status Scratch::clear_flags() {
  RH_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(unsigned), stream));
  return RH_OK;
}

rh_status Scratch::read_flags(unsigned* out) {
  RH_HIP(hipMemcpyAsync(host, flags, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
  RH_HIP(hipStreamSynchronize(stream));
  *out = *(unsigned*)host;
  return RH_OK;
}

rh_status Scratch::read_flags_scalars(unsigned f[2], double* out, int n) {
  // the flag words lie below result slots 0..n-1 in one allocation
  char* lo = (char*)flags;
  const size_t span = (char*)(scal + n) - lo;
  RH_HIP(hipMemcpyAsync(host, lo, span, hipMemcpyDeviceToHost, stream));
  RH_HIP(hipStreamSynchronize(stream));
  std::memcpy(f, host, 2 * sizeof(unsigned));
  std::memcpy(out, (char*)host + ((char*)scal - lo), sizeof(double) * n);
  return RH_OK;
}

void Scratch::last_readback(unsigned f[2], double* out, int slot, int n) const {
  const char* h = (const char*)host + 1024;  // where Solver::finish_report copies [flags, ctl]
  std::memcpy(f, h, 2 * sizeof(unsigned));
  std::memcpy(out, h + ((const char*)(scal + slot) - (const char*)flags), sizeof(double) * n);
}

rh_status Scratch::read_scalars(double* out, int n, int slot) {
  RH_HIP(hipMemcpyAsync(host, scal + slot, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  RH_HIP(hipStreamSynchronize(stream));
  std::memcpy(out, host, sizeof(double) * n);
  return RH_OK;
}

rh_status flags_to_status(unsigned f) {
  if (f & kFlagNonfiniteInput) return RH_NONFINITE_INPUT;
  if (f & kFlagNonzeroWall) return RH_NONZERO_WALL_VELOCITY;
  if (f & kFlagArithmetic) return RH_ARITHMETIC_FAILURE;
  return RH_OK;
}

// The process-wide scratch for stateless operator entry points.
Scratch& global_scratch(std::unique_lock<std::mutex>& lock) {
  static std::mutex mu;
  static Scratch s;
  lock = std::unique_lock<std::mutex>(mu);
  return s;
}

}  // namespace rh
