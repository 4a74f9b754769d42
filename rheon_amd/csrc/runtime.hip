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
  int groups = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
  if (groups < 1) groups = 1;
  if (groups > kMaxBlocks) return kMaxBlocks;
  // Multiples of 8 let block_rows() give each XCD a contiguous slab.
  if (groups >= 64) groups &= ~7;
  return groups;
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
  for (int a = 0; a < 3; ++a) {
    d.h[a] = g.spacing[a];
    d.origin[a] = g.origin[a];
    d.upper[a] = g.origin[a] + (double)g.counts[a] * g.spacing[a];
  }
  d.volume = g.spacing[0] * g.spacing[1] * g.spacing[2];
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

// ------------------------------------------------------------------ Scratch

rh_status Scratch::init() {
  if (stream) return RH_OK;
  RH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const size_t bytes = sizeof(DD) * 4 * kMaxBlocks + sizeof(double) * 4 * kMaxBlocks +
                       sizeof(unsigned) * 64 + sizeof(double) * 64 + 2 * sizeof(PcgControl) + 1024;
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
  r.counter = counters + 8;
  r.out = scal + out_slot;
  r.flags = flags;
  return r;
}

rh_status Scratch::clear_flags() {
  RH_HIP(hipMemsetAsync(flags, 0, sizeof(unsigned), stream));
  return RH_OK;
}

rh_status Scratch::read_flags(unsigned* out) {
  RH_HIP(hipMemcpyAsync(host, flags, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
  RH_HIP(hipStreamSynchronize(stream));
  *out = *(unsigned*)host;
  return RH_OK;
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

// --------------------------------------------------------------- PcgSolver

rh_status PcgSolver::create(const rh_grid& grid, const rh_solver_config& config) {
  grid_ = grid;
  config_ = config;
  if (config.method != RH_JACOBI_PCG && config.method != RH_RBSGS_PCG) return RH_UNSUPPORTED;
  if (config.precision != RH_F64 && config.precision != RH_F32) return RH_INVALID_ARGUMENT;
  rh_status st = scratch_.init();
  if (st) return st;
  n_ = (size_t)grid.counts[0] * grid.counts[1] * grid.counts[2];
  const size_t s = config.precision == RH_F64 ? 8 : 4;
  // r, d, q (+ z for RB-SGS) (+ p when vectors are f32; f64 p is the output).
  int vectors = 3 + (config.method == RH_RBSGS_PCG ? 1 : 0) + (config.precision == RH_F32 ? 1 : 0);
  bytes_ = s * n_ * vectors;
  if (config.memory_limit && bytes_ > config.memory_limit) return RH_BUFFER_LIMIT;
  RH_HIP(hipMalloc(&arena_, bytes_ + 256 * vectors));
  char* p = (char*)arena_;
  auto take = [&]() {
    void* r = p;
    p += ((s * n_ + 255) / 256) * 256;
    return r;
  };
  r_ = take();
  d_ = take();
  q_ = take();
  z_ = config.method == RH_RBSGS_PCG ? take() : nullptr;
  p32_ = config.precision == RH_F32 ? take() : nullptr;
  return RH_OK;
}

PcgSolver::~PcgSolver() {
  if (arena_) (void)hipFree(arena_);
  scratch_.release();
}

uint64_t PcgSolver::allocated_bytes() const { return bytes_ + scratch_.bytes_; }

template <typename T>
rh_status PcgSolver::run(const DevGrid& g, const double* d_rhs, const rh_pressure_settings& st,
                         double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel,
                         void* ctx) {
  hipStream_t s = scratch_.stream;
  PcgArrays a;
  a.p = sizeof(T) == 8 ? (void*)d_pressure : p32_;
  a.r = r_;
  a.d = d_;
  a.q = q_;
  a.z = z_;
  a.b = d_rhs;
  a.p64 = d_pressure;
  PcgControl h;
  std::memset(&h, 0, sizeof(h));
  h.rel = st.relative_residual;
  h.abs_tol = st.absolute_residual;
  h.div_limit = st.divergence_limit;
  h.dt = dt_;
  h.volume = g.volume;
  h.max_iterations = (int)(st.max_iterations > INT32_MAX - 2 ? INT32_MAX - 2 : st.max_iterations);
  std::memcpy(scratch_.host, &h, sizeof(h));
  RH_HIP(hipMemcpyAsync(scratch_.ctl, scratch_.host, sizeof(h), hipMemcpyHostToDevice, s));
  RH_HIP(hipEventRecord(scratch_.ev0, s));
  pcg_prologue<T>(g, a, scratch_.ctl, scratch_.pdd, scratch_.pmx, s);
  pcg_iteration<T>(g, config_.method, a, scratch_.ctl, scratch_.pdd, scratch_.pmx, 0, s);
  RH_HIP(hipGetLastError());
  const int poll = st.poll_interval > 0 ? st.poll_interval : 16;
  PcgControl* slot[2] = {(PcgControl*)scratch_.host + 1, (PcgControl*)scratch_.host + 2};
  bool cancelled = false;
  int it = 1;
  int pending = -1;  // slot index with an outstanding readback
  // Enqueue batch k+1 before waiting for batch k's readback, so the device
  // never idles on the host's poll.
  for (int batch = 0;; ++batch) {
    const int end = std::min(h.max_iterations, it + poll - 1);
    for (; it <= end; ++it)
      pcg_iteration<T>(g, config_.method, a, scratch_.ctl, scratch_.pdd, scratch_.pmx, it, s);
    const int cur = batch & 1;
    RH_HIP(hipMemcpyAsync(slot[cur], scratch_.ctl, sizeof(PcgControl), hipMemcpyDeviceToHost, s));
    RH_HIP(hipEventRecord(scratch_.poll_ev[cur], s));
    bool finished = it > h.max_iterations;
    if (pending >= 0) {
      RH_HIP(hipEventSynchronize(scratch_.poll_ev[pending]));
      if (slot[pending]->done) finished = true;
      else if (cancel && cancel(ctx)) {
        cancelled = true;
        finished = true;
      }
    }
    pending = cur;
    if (finished) break;
  }
  RH_HIP(hipGetLastError());
  pcg_export_pressure<T>(g, a, s);
  RH_HIP(hipEventRecord(scratch_.ev1, s));
  RH_HIP(hipMemcpyAsync(slot[0], scratch_.ctl, sizeof(PcgControl), hipMemcpyDeviceToHost, s));
  RH_HIP(hipStreamSynchronize(s));
  const PcgControl& c = *slot[0];
  if (prof::enabled()) prof::settle(c.iteration);
  float ms = 0.0f;
  RH_HIP(hipEventElapsedTime(&ms, scratch_.ev0, scratch_.ev1));
  if (report) {
    report->iterations = c.iteration;
    report->true_residual_l2 = c.true_l2;
    report->true_residual_max = c.true_max;
    report->predicted_divergence_max = c.pred_div;
    report->error_iteration = c.error_iteration;
    report->error_value = c.error_value;
    report->threshold = c.threshold;
    report->seconds = ms * 1e-3;
  }
  if (c.done) return (rh_status)c.status;
  if (cancelled) {
    if (report) report->error_iteration = c.iteration;
    return RH_CANCELLED;
  }
  set_last_error("solver ended without a terminal status", "PcgSolver::run");
  return RH_EXECUTION;
}

rh_status PcgSolver::solve(double density, const double* d_rhs, double dt,
                           const rh_pressure_settings& st, double* d_pressure,
                           rh_pressure_report* report, rh_cancel_fn cancel, void* ctx) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!d_rhs || !d_pressure) return RH_INVALID_ARGUMENT;
  rh_status s = check_density(grid_, density);
  if (s) return s;
  // solve_loaded settings check (src/pressure.rs:588-599).
  if (!std::isfinite(dt) || dt <= 0.0 || !std::isfinite(st.relative_residual) ||
      st.relative_residual < 0.0 || !std::isfinite(st.absolute_residual) ||
      st.absolute_residual < 0.0 || !std::isfinite(st.divergence_limit) ||
      st.divergence_limit < 0.0 || st.max_iterations < 0)
    return RH_INVALID_SETTINGS;
  if (cancel && cancel(ctx)) return RH_CANCELLED;
  dt_ = dt;
  const DevGrid g = make_dev_grid(grid_, density);
  if (config_.precision == RH_F64) return run<double>(g, d_rhs, st, d_pressure, report, cancel, ctx);
  return run<float>(g, d_rhs, st, d_pressure, report, cancel, ctx);
}

}  // namespace rh
