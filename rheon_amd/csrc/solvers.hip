// Device pressure solvers (PressureWorkspace on the GPU, src/pressure.rs:243-740):
// the PCG family and the multigrid defect-correction solver.  Both keep every
// scalar decision on the device; the host enqueues iterations in batches and
// reads the control block back one batch behind, so the device never waits
// for the host.
#include <functional>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "mg.hpp"
#include "profiler.hpp"
#include "runtime.hpp"

namespace rh {

Solver* make_solver(rh_method method) {
  switch (method) {
    case RH_JACOBI_PCG:
    case RH_RBSGS_PCG:
    case RH_SGS_PCG: return new PcgSolver();
    case RH_MIXED_MG: return new MgSolver();
    default: return nullptr;
  }
}

// solve_loaded settings check (src/pressure.rs:588-599) and cancellation
// before any work (:600-602).
rh_status Solver::check_and_begin(double density, const double* d_rhs, double dt,
                                  const rh_pressure_settings& st, double* d_pressure,
                                  rh_pressure_report* report, rh_cancel_fn cancel, void* ctx) {
  if (report) std::memset(report, 0, sizeof(*report));
  if (!d_rhs || !d_pressure) return RH_INVALID_ARGUMENT;
  rh_status s = check_density(grid_, density);
  if (s) return s;
  if (!std::isfinite(dt) || dt <= 0.0 || !std::isfinite(st.relative_residual) ||
      st.relative_residual < 0.0 || !std::isfinite(st.absolute_residual) ||
      st.absolute_residual < 0.0 || !std::isfinite(st.divergence_limit) ||
      st.divergence_limit < 0.0 || st.max_iterations < 0)
    return RH_INVALID_SETTINGS;
  if (cancel && cancel(ctx)) return RH_CANCELLED;
  dt_ = dt;
  return RH_OK;
}

template <typename Step>
rh_status Solver::drive(int max_iterations, int poll, Step step, bool* cancelled,
                        rh_cancel_fn cancel, void* ctx) {
  hipStream_t s = scratch_.stream;
  PcgControl* slot[2] = {(PcgControl*)scratch_.host + 1, (PcgControl*)scratch_.host + 2};
  *cancelled = false;
  int it = 1;
  int pending = -1;
  for (int batch = 0;; ++batch) {
    const int end = std::min(max_iterations, it + poll - 1);
    for (; it <= end; ++it) step(it);
    const int cur = batch & 1;
    RH_HIP(hipMemcpyAsync(slot[cur], scratch_.ctl, sizeof(PcgControl), hipMemcpyDeviceToHost, s));
    RH_HIP(hipEventRecord(scratch_.poll_ev[cur], s));
    bool finished = it > max_iterations;
    if (pending >= 0) {
      RH_HIP(hipEventSynchronize(scratch_.poll_ev[pending]));
      if (slot[pending]->done) {
        finished = true;
      } else if (cancel && cancel(ctx)) {
        *cancelled = true;
        finished = true;
      }
    }
    pending = cur;
    if (finished) break;
  }
  RH_HIP(hipGetLastError());
  return RH_OK;
}

template <typename Step>
rh_status Solver::drive_paced(int max_iterations, int poll, int ahead, Step step, bool* cancelled,
                              rh_cancel_fn cancel, void* ctx, int* enqueued) {
  hipStream_t s = scratch_.stream;
  volatile const int* progress = scratch_.progress;
  *cancelled = false;
  int it = 0;           // iterations enqueued
  int polled = 0;       // decided iterations at the last cancellation poll
  unsigned spins = 0;
  bool drained = false;  // the stream had finished everything enqueued
  for (;;) {
    const int v = *progress;
    if (v & 1) break;                  // terminal status on the device
    const int seen = (v >> 1) - 1;     // last decided iteration (-1: none yet)
    if (cancel && seen + 1 - polled >= poll) {
      polled = seen + 1;
      if (cancel(ctx)) {
        *cancelled = true;
        break;
      }
    }
    if (it < max_iterations && it < std::max(ahead, seen + 2)) {
      step(++it);
      spins = 0;
      drained = false;
      continue;
    }
    // Waiting on the device.  Once the stream has drained (and the word,
    // read again above, allows no further iteration -- e.g. all are
    // enqueued and the last one's decision is not launched yet) the loop
    // ends; the caller reads the control block itself.
    if (drained) break;
    if ((++spins & 1023u) == 0) {
      const hipError_t q = hipStreamQuery(s);
      if (q == hipSuccess) {
        drained = true;
      } else if (q != hipErrorNotReady) {
        set_last_error(hipGetErrorString(q), "Solver::drive_paced");
        return RH_DEVICE_ERROR;
      }
    }
    __builtin_ia32_pause();
  }
  *enqueued = it;
  RH_HIP(hipGetLastError());
  return RH_OK;
}

rh_status Solver::finish_report(rh_pressure_report* report, bool cancelled, PcgControl* out, double* scal3,
                                 bool ev1_recorded) {
  hipStream_t s = scratch_.stream;
  PcgControl* c = (PcgControl*)scratch_.host + 1;
  if (!ev1_recorded) RH_HIP(hipEventRecord(scratch_.ev1, s));
  if (scal3) {
    // the flag words, every result slot (the report sums in 0..2) and the
    // control block, which follow one another in the scratch allocation,
    // in one copy (Scratch::last_readback reads the rest of it)
    char* lo = (char*)scratch_.flags;
    const size_t span = (char*)(scratch_.ctl + 1) - lo;
    char* h = (char*)scratch_.host + 1024;
    RH_HIP(hipMemcpyAsync(h, lo, span, hipMemcpyDeviceToHost, s));
    RH_HIP(hipStreamSynchronize(s));
    std::memcpy(scal3, h + ((char*)scratch_.scal - lo), 3 * sizeof(double));
    std::memcpy(c, h + ((char*)scratch_.ctl - lo), sizeof(PcgControl));
  } else {
    RH_HIP(hipMemcpyAsync(c, scratch_.ctl, sizeof(PcgControl), hipMemcpyDeviceToHost, s));
    RH_HIP(hipStreamSynchronize(s));
  }
  if (prof::enabled()) prof::settle(c->iteration);
  float ms = 0.0f;
  RH_HIP(hipEventElapsedTime(&ms, scratch_.ev0, scratch_.ev1));
  if (report) {
    report->iterations = c->iteration;
    report->true_residual_l2 = c->true_l2;
    report->true_residual_max = c->true_max;
    report->predicted_divergence_max = c->pred_div;
    report->error_iteration = c->error_iteration;
    report->error_value = c->error_value;
    report->threshold = c->threshold;
    report->seconds = ms * 1e-3;
  }
  if (out) *out = *c;
  if (c->done) return (rh_status)c->status;
  if (cancelled) {
    if (report) report->error_iteration = c->iteration;
    return RH_CANCELLED;
  }
  set_last_error("solver ended without a terminal status", "Solver::finish_report");
  return RH_EXECUTION;
}

// ------------------------------------------------------------ obstacles

void Solver::release_mask() {
  if (d_mask_) (void)hipFree(d_mask_);
  d_mask_ = nullptr;
  hmask_.clear();
  gauge_ = 0;
  release_components();
  ++generation_;
}

void Solver::release_components() {
  if (d_comp_) (void)hipFree(d_comp_);
  d_comp_ = nullptr;
  ncomp_ = 1;
  comp_nchunks_ = 0;
  comp_label_ = nullptr;
  comp_gauge_ = nullptr;
  comp_shift_ = nullptr;
  comp_cells_ = nullptr;
  comp_chunks_ = nullptr;
  comp_chunk_first_ = nullptr;
  gauges_.clear();
}

// Device tables of a fluid split into several regions: labels, gauges, one
// captured-pressure slot per region, and the fluid cells sorted by region
// (index order within a region) in chunks of at most `len` cells, none
// spanning two regions, so every region's compatibility sums are formed
// over its own cells.
rh_status Solver::upload_components(const ObstacleComponents& comps) {
  release_components();
  const int nc = (int)comps.gauges.size();
  if (nc <= 1) return RH_OK;
  const size_t n = (size_t)grid_.counts[0] * grid_.counts[1] * grid_.counts[2];
  std::vector<int> count(nc, 0);
  for (size_t c = 0; c < n; ++c)
    if (comps.labels[c] != kNoComponent) ++count[comps.labels[c]];
  std::vector<int> start(nc + 1, 0);
  for (int k = 0; k < nc; ++k) start[k + 1] = start[k] + count[k];
  const int fluid = start[nc];
  std::vector<int> cells(fluid), fill(start.begin(), start.end() - 1);
  for (size_t c = 0; c < n; ++c)
    if (comps.labels[c] != kNoComponent) cells[fill[comps.labels[c]]++] = (int)c;
  const int len = std::max(2048, (fluid + 8191) / 8192);
  std::vector<CompChunk> chunks;
  std::vector<int> first(nc + 1, 0);
  for (int k = 0; k < nc; ++k) {
    first[k] = (int)chunks.size();
    for (int b = start[k]; b < start[k + 1]; b += len) chunks.push_back(CompChunk{k, b, std::min(b + len, start[k + 1]), 0});
  }
  first[nc] = (int)chunks.size();
  if ((int)chunks.size() > kMaxBlocks) return RH_UNSUPPORTED;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t o_label = 0, o_gauge = up(n), o_shift = o_gauge + up(sizeof(int) * nc),
               o_cells = o_shift + up(sizeof(double) * nc), o_chunks = o_cells + up(sizeof(int) * (size_t)fluid),
               o_first = o_chunks + up(sizeof(CompChunk) * chunks.size()),
               total = o_first + up(sizeof(int) * (size_t)(nc + 1));
  char* d = nullptr;
  RH_HIP(hipMalloc(&d, total));
  d_comp_ = d;
  RH_HIP(hipMemset(d, 0, total));
  RH_HIP(hipMemcpy(d + o_label, comps.labels.data(), n, hipMemcpyHostToDevice));
  RH_HIP(hipMemcpy(d + o_gauge, comps.gauges.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
  RH_HIP(hipMemcpy(d + o_cells, cells.data(), sizeof(int) * (size_t)fluid, hipMemcpyHostToDevice));
  RH_HIP(hipMemcpy(d + o_chunks, chunks.data(), sizeof(CompChunk) * chunks.size(), hipMemcpyHostToDevice));
  RH_HIP(hipMemcpy(d + o_first, first.data(), sizeof(int) * (size_t)(nc + 1), hipMemcpyHostToDevice));
  ncomp_ = nc;
  comp_nchunks_ = (int)chunks.size();
  comp_label_ = (const unsigned char*)(d + o_label);
  comp_gauge_ = (const int*)(d + o_gauge);
  comp_shift_ = (double*)(d + o_shift);
  comp_cells_ = (const int*)(d + o_cells);
  comp_chunks_ = (const CompChunk*)(d + o_chunks);
  comp_chunk_first_ = (const int*)(d + o_first);
  gauges_ = comps.gauges;
  return RH_OK;
}

// Row flags stored on the device right after the level-0 mask (one byte per
// (j, k) row): bit 0 = every cell of the row is fluid with the plain box's
// open faces, bit 1 = bit 0 holds for the row and its four neighbour rows
// (clamped at walls).  The four-cell kernels skip the mask load on such rows
// (kRowPlain / kRowPlainNear in mg_quad.hpp).
static std::vector<uint8_t> mask_row_flags(const rh_grid& g, const std::vector<uint8_t>& mask) {
  const int nx = (int)g.counts[0], ny = (int)g.counts[1], nz = (int)g.counts[2];
  std::vector<uint8_t> plain((size_t)ny * nz), out((size_t)ny * nz);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const size_t row = (size_t)j + (size_t)ny * k;
      bool ok = true;
      for (int i = 0; i < nx && ok; ++i) {
        const int box = (i > 0) | ((i + 1 < nx) << 1) | ((j > 0) << 2) | ((j + 1 < ny) << 3) | ((k > 0) << 4) |
                        ((k + 1 < nz) << 5);
        ok = mask[(size_t)i + (size_t)nx * row] == (uint8_t)(box | kFluidBit);
      }
      plain[row] = ok ? 1 : 0;
    }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const size_t row = (size_t)j + (size_t)ny * k;
      const int jm = j > 0 ? j - 1 : j, jp = j + 1 < ny ? j + 1 : j;
      const int km = k > 0 ? k - 1 : k, kp = k + 1 < nz ? k + 1 : k;
      const bool near = plain[row] && plain[(size_t)jm + (size_t)ny * k] && plain[(size_t)jp + (size_t)ny * k] &&
                        plain[(size_t)j + (size_t)ny * km] && plain[(size_t)j + (size_t)ny * kp];
      out[row] = (uint8_t)(plain[row] | (near ? 2 : 0));
    }
  return out;
}

rh_status Solver::set_solid(const uint8_t* solid) {
  if (!solid) {
    release_mask();
    return RH_OK;
  }
  std::vector<uint8_t> mask;
  int gauge = 0;
  ObstacleComponents comps;
  rh_status st = build_obstacle_mask(grid_, solid, mask, &gauge, &comps);
  if (st) return st;
  if ((st = upload_components(comps))) return st;
  const std::vector<uint8_t> rows = mask_row_flags(grid_, mask);
  if (!d_mask_) RH_HIP(hipMalloc(&d_mask_, mask.size() + rows.size()));
  RH_HIP(hipMemcpy(d_mask_, mask.data(), mask.size(), hipMemcpyHostToDevice));
  RH_HIP(hipMemcpy(d_mask_ + mask.size(), rows.data(), rows.size(), hipMemcpyHostToDevice));
  hmask_.swap(mask);
  gauge_ = gauge;
  ++generation_;
  return RH_OK;
}

DevGrid Solver::dev_grid(double density) const {
  DevGrid g = make_dev_grid(grid_, density);
  g.mask = d_mask_;
  g.gauge = gauge_;
  g.ncomp = d_mask_ ? ncomp_ : 1;
  g.comp_nchunks = comp_nchunks_;
  g.comp_label = comp_label_;
  g.comp_gauge = comp_gauge_;
  g.comp_shift = comp_shift_;
  g.comp_cells = comp_cells_;
  g.comp_chunks = comp_chunks_;
  g.comp_chunk_first = comp_chunk_first_;
  return g;
}

static void init_control(PcgControl& h, const rh_pressure_settings& st, double dt, double volume) {
  std::memset(&h, 0, sizeof(h));
  h.rel = st.relative_residual;
  h.abs_tol = st.absolute_residual;
  h.div_limit = st.divergence_limit;
  h.dt = dt;
  h.volume = volume;
  h.max_iterations = (int)(st.max_iterations > INT32_MAX - 2 ? INT32_MAX - 2 : st.max_iterations);
}

// ------------------------------------------------------------------- PCG

rh_status PcgSolver::create(const rh_grid& grid, const rh_solver_config& config) {
  grid_ = grid;
  config_ = config;
  if (config.method != RH_JACOBI_PCG && config.method != RH_RBSGS_PCG && config.method != RH_SGS_PCG)
    return RH_UNSUPPORTED;
  if (config.precision != RH_F64 && config.precision != RH_F32) return RH_INVALID_ARGUMENT;
  rh_status st = scratch_.init();
  if (st) return st;
  n_ = (size_t)grid.counts[0] * grid.counts[1] * grid.counts[2];
  const size_t s = config.precision == RH_F64 ? 8 : 4;
  // r, d, q, z (RB-SGS: the preconditioned residual; Jacobi: the true
  // residual of a candidate iteration, held until the next apply folds a
  // replacement in); the solution accumulates in the caller's f64 pressure
  // for either vector precision.
  int vectors = 4;
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
  z_ = take();
  if (config.method == RH_SGS_PCG) {
    std::vector<SgsTile> tiles;
    sgs_plan(make_dev_grid(grid, 1.0), tiles, wave_start_);
    RH_HIP(hipMalloc(&d_tiles_, sizeof(SgsTile) * tiles.size()));
    RH_HIP(hipMemcpy(d_tiles_, tiles.data(), sizeof(SgsTile) * tiles.size(), hipMemcpyHostToDevice));
    bytes_ += sizeof(SgsTile) * tiles.size();
  }
  return RH_OK;
}

PcgSolver::~PcgSolver() {
  if (arena_) (void)hipFree(arena_);
  if (d_tiles_) (void)hipFree(d_tiles_);
  scratch_.release();
}

// The lexicographic SGS (sgs-pcg-v1) is stated for the sealed box.
rh_status PcgSolver::set_solid(const uint8_t* solid) {
  if (solid && config_.method == RH_SGS_PCG) {
    set_last_error("sgs-pcg-v1 runs on the sealed box only", "rh_solver_set_solid");
    return RH_UNSUPPORTED;
  }
  return Solver::set_solid(solid);
}

uint64_t PcgSolver::allocated_bytes() const { return bytes_ + scratch_.bytes_; }

template <typename T>
rh_status PcgSolver::run(const DevGrid& g, const double* d_rhs, const rh_pressure_settings& st,
                         double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel,
                         void* ctx) {
  hipStream_t s = scratch_.stream;
  PcgArrays a;
  a.p = d_pressure;
  a.r = r_;
  a.d = d_;
  a.q = q_;
  a.z = z_;
  a.b = d_rhs;
  a.p64 = d_pressure;
  a.tiles = d_tiles_;
  a.wave_start = wave_start_.data();
  a.waves = (int)wave_start_.size() - 1;
  a.seed = seed_;
  PcgControl h;
  init_control(h, st, dt_, g.volume);
  h.seeded = seed_ ? 1 : 0;
  std::memcpy(scratch_.host, &h, sizeof(h));
  RH_HIP(hipMemcpyAsync(scratch_.ctl, scratch_.host, sizeof(h), hipMemcpyHostToDevice, s));
  RH_HIP(hipEventRecord(scratch_.ev0, s));
  pcg_prologue<T>(g, a, scratch_.ctl, scratch_.pdd, scratch_.pmx, s);
  pcg_iteration<T>(g, config_.method, a, scratch_.ctl, scratch_.pdd, scratch_.pmx, 0, s);
  RH_HIP(hipGetLastError());
  bool cancelled = false;
  const int poll = st.poll_interval > 0 ? st.poll_interval : 16;
  rh_status rs = drive(
      h.max_iterations, poll,
      [&](int it) {
        pcg_iteration<T>(g, config_.method, a, scratch_.ctl, scratch_.pdd, scratch_.pmx, it, s);
      },
      &cancelled, cancel, ctx);
  if (rs) return rs;
  pcg_export_pressure<T>(g, a, s);
  return finish_report(report, cancelled, nullptr);
}

rh_status PcgSolver::solve(double density, const double* d_rhs, double dt,
                           const rh_pressure_settings& st, double* d_pressure,
                           rh_pressure_report* report, rh_cancel_fn cancel, void* ctx) {
  rh_status s = check_and_begin(density, d_rhs, dt, st, d_pressure, report, cancel, ctx);
  if (s) return s;
  const DevGrid g = dev_grid(density);
  if (config_.precision == RH_F64) return run<double>(g, d_rhs, st, d_pressure, report, cancel, ctx);
  return run<float>(g, d_rhs, st, d_pressure, report, cancel, ctx);
}

// -------------------------------------------------------------- multigrid

// Sweep counts of the V-cycle: explicit, or the smoother's default.
rh_status MgSolver::plan_sweeps(const rh_solver_config& config, int* pre, int* post, int* cpre, int* cpost) {
  if (config.mg_smoother != RH_SMOOTHER_CHEBYSHEV && config.mg_smoother != RH_SMOOTHER_JACOBI)
    return RH_INVALID_ARGUMENT;
  const bool jacobi = config.mg_smoother == RH_SMOOTHER_JACOBI;
  *pre = config.mg_pre_sweeps > 0 ? config.mg_pre_sweeps : (jacobi ? 2 : 3);
  *post = config.mg_post_sweeps > 0 ? config.mg_post_sweeps : (jacobi ? 2 : 1);
  *cpre = config.mg_coarse_pre_sweeps > 0 ? config.mg_coarse_pre_sweeps : *pre;
  *cpost = config.mg_coarse_post_sweeps > 0 ? config.mg_coarse_post_sweeps : *post;
  if (*pre > kMgMaxSweeps || *post > kMgMaxSweeps || *cpre > kMgMaxSweeps || *cpost > kMgMaxSweeps ||
      config.mg_pre_sweeps < 0 || config.mg_post_sweeps < 0 || config.mg_coarse_pre_sweeps < 0 ||
      config.mg_coarse_post_sweeps < 0 || config.mg_omega < 0.0)
    return RH_INVALID_ARGUMENT;
  return RH_OK;
}

// Weights tau_k of n sweeps: weighted Jacobi (all omega) or the reciprocal
// roots lambda_k = c + h cos((2k + 1) pi / 2n), k = 0..n-1, of the degree-n
// Chebyshev polynomial on [lo, 2] (c, h = centre and half-width).
static void sweep_weights(int n, bool jacobi, double omega, double* tau) {
  const double lo = n <= 2 ? 0.3 : 0.25, hi = 2.0;
  const double c = 0.5 * (hi + lo), h = 0.5 * (hi - lo);
  for (int k = 0; k < n; ++k)
    tau[k] = jacobi ? omega : 1.0 / (c + h * std::cos(M_PI * (2.0 * k + 1.0) / (2.0 * n)));
}

// Level shapes of the hierarchy for this grid (shared by create and the
// memory plan): returns the level count; *padded = first padded level.
int MgSolver::plan_levels(const rh_grid& grid, const rh_solver_config& config, int dims[][3],
                          int* padded) {
  int n[3] = {(int)grid.counts[0], (int)grid.counts[1], (int)grid.counts[2]};
  int levels = 1;
  dims[0][0] = n[0];
  dims[0][1] = n[1];
  dims[0][2] = n[2];
  const int want = config.mg_levels > 0 ? std::min(config.mg_levels, kMgMaxLevels) : kMgMaxLevels;
  // A level with an odd axis (or an axis below 4) above 64 cells is
  // coarsened with ceil(n/2) cells per axis: the last coarse cell of an odd
  // axis has one child, and the missing children act as solid cells (the
  // coarse levels from there on carry face fractions).  The coarsest level
  // then has at most 64 cells and a small dense inverse.
  *padded = 0;
  while (levels < want) {
    const int* m = &dims[levels - 1][0];
    const long cells = (long)m[0] * m[1] * m[2];
    const bool even = (m[0] % 2 == 0) && (m[1] % 2 == 0) && (m[2] % 2 == 0);
    const int mn = std::min(m[0], std::min(m[1], m[2]));
    const int mx = std::max(m[0], std::max(m[1], m[2]));
    if (even && mn >= 4) {
      if (config.mg_levels <= 0 && cells <= 64) break;
    } else {
      // odd axis (or an axis below 4): keep coarsening with padding down to
      // the small dense coarsest level
      if (config.mg_levels > 0 || cells <= 64 || mx < 3) break;
      if (!*padded) *padded = levels;
    }
    dims[levels][0] = (m[0] + 1) / 2;
    dims[levels][1] = (m[1] + 1) / 2;
    dims[levels][2] = (m[2] + 1) / 2;
    ++levels;
  }
  return levels;
}

// Device bytes a scene of this grid holds with this solver (admission plan).
uint64_t MgSolver::plan_bytes(const rh_grid& grid, const rh_solver_config& config) {
  int dims[kMgMaxLevels][3];
  int padded = 0;
  const int levels = plan_levels(grid, config, dims, &padded);
  const size_t s = config.precision == RH_F64 ? 8 : 4;
  uint64_t total = 0;
  for (int l = 0; l < levels; ++l) {
    const uint64_t cells = (uint64_t)dims[l][0] * dims[l][1] * dims[l][2];
    total += 3 * ((cells * s + 255) / 256 * 256);
    if (l > 0) total += (cells + 63) / 64 * 256 + (cells + 255) / 256 * 256;  // fractions, octants
    if (l == 0) total += cells + (uint64_t)dims[0][1] * dims[0][2];            // obstacle mask, row flags
  }
  const uint64_t coarse = (uint64_t)dims[levels - 1][0] * dims[levels - 1][1] * dims[levels - 1][2];
  // dense inverse, tables, the tail's per-cell 1/D (at most its LDS), scratch
  total += coarse * coarse * 8 + sizeof(MgHierarchy) + kMgTailLds + Scratch::kDeviceBytes;
  return total;
}

rh_status MgSolver::create(const rh_grid& grid, const rh_solver_config& config) {
  grid_ = grid;
  config_ = config;
  if (config.method != RH_MIXED_MG) return RH_UNSUPPORTED;
  if (config.precision != RH_F64 && config.precision != RH_F32) return RH_INVALID_ARGUMENT;
  rh_status st = scratch_.init();
  if (st) return st;
  st = plan_sweeps(config, &h_.pre, &h_.post, &h_.cpre, &h_.cpost);
  if (st) return st;
  h_.omega = config.mg_omega > 0.0 ? config.mg_omega : 6.0 / 7.0;
  {
    const bool jacobi = config.mg_smoother == RH_SMOOTHER_JACOBI;
    sweep_weights(h_.pre, jacobi, h_.omega, h_.tau_pre);
    sweep_weights(h_.post, jacobi, h_.omega, h_.tau_post);
    sweep_weights(h_.cpre, jacobi, h_.omega, h_.tau_cpre);
    sweep_weights(h_.cpost, jacobi, h_.omega, h_.tau_cpost);
  }
  h_.quad = (config.kernel_variant & RH_VARIANT_SCALAR_MG) == 0;
  // Coarsen while every axis is even and the level is above the one-
  // workgroup coarsest size (or up to the requested level count).
  int dims[kMgMaxLevels][3];
  int levels = plan_levels(grid, config, dims, &padded_level_);
  for (int l = 0; l < levels; ++l) {
    h_.lv[l].nx = dims[l][0];
    h_.lv[l].ny = dims[l][1];
    h_.lv[l].nz = dims[l][2];
    set_level_divisors(h_.lv[l]);
  }
  h_.levels = levels;
  const MgLevel& last = h_.lv[levels - 1];
  if ((long)last.nx * last.ny * last.nz > kMgCoarsestMax) return RH_UNSUPPORTED;
  const size_t s = config.precision == RH_F64 ? 8 : 4;
  size_t total = 0;
  for (int l = 0; l < levels; ++l) {
    const size_t cells = (size_t)h_.lv[l].nx * h_.lv[l].ny * h_.lv[l].nz;
    total += 3 * ((cells * s + 255) / 256 * 256);
  }
  bytes_ = total;
  if (config.memory_limit && bytes_ > config.memory_limit) return RH_BUFFER_LIMIT;
  RH_HIP(hipMalloc(&arena_, total));
  RH_HIP(hipMemset(arena_, 0, total));
  char* p = (char*)arena_;
  for (int l = 0; l < levels; ++l) {
    const size_t cells = (size_t)h_.lv[l].nx * h_.lv[l].ny * h_.lv[l].nz;
    const size_t sz = (cells * s + 255) / 256 * 256;
    h_.x[l] = p;
    p += sz;
    h_.y[l] = p;
    p += sz;
    h_.b[l] = p;
    p += sz;
  }
  if (padded_level_) return build_fractions();
  return RH_OK;
}

MgSolver::~MgSolver() {
  if (arena_) (void)hipFree(arena_);
  if (d_frac_) (void)hipFree(d_frac_);
  if (d_inv_) (void)hipFree(d_inv_);
  if (d_odc_) (void)hipFree(d_odc_);
  if (d_h_) (void)hipFree(d_h_);
  scratch_.release();
}

uint64_t MgSolver::allocated_bytes() const { return bytes_ + scratch_.bytes_; }

// Host view of one level's face weights: the box, the level-0 obstacle mask
// or the coarse face fractions (same rules as the device stencil).
struct LevelFaces {
  const MgLevel& L;
  const std::vector<uint8_t>* mask;   // level 0 with obstacles
  const std::vector<unsigned>* frac;  // levels >= 1 with obstacles
  double fscale;
  bool active(int c) const {
    if (frac) return ((*frac)[c] & kFracFluid) != 0;
    if (mask) return ((*mask)[c] & kFluidBit) != 0;
    return true;
  }
  // weight of face `side` (0..5 = x-, x+, y-, y+, z-, z+) of cell c; 0 if closed
  double weight(int c, int side) const {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    const int a = side / 2;
    const bool upper = side & 1;
    const int pos[3] = {i, j, k}, n[3] = {nx, ny, nz}, stride[3] = {1, nx, nx * ny};
    if (upper ? pos[a] + 1 >= n[a] : pos[a] == 0) return 0.0;
    if (frac) {
      const unsigned q = (*frac)[upper ? c + stride[a] : c];
      return L.w[a] * fscale * (double)((q >> (10 * a)) & 1023u);
    }
    if (mask) return ((*mask)[c] >> side) & 1 ? L.w[a] : 0.0;
    return L.w[a];
  }
};

// G = M^-1 P over the active cells, region by region: the active cells
// split into regions a_r connected through open faces (one region unless the
// obstacles cut the fluid apart at this level), M = K + sum_r c_r 1_r 1_r^T
// (SPD: K is block-diagonal over the regions and each rank-one term lifts
// its region's constant null vector; c_r = mean diagonal of a_r / n_r) by
// Cholesky on the host, P = I - sum_r 1_r 1_r^T / n_r.  For any b, x = G b
// solves K x = P b with every region's mean of x zero.  Inactive rows and
// columns of G are 0.
static bool dense_coarse_inverse(const LevelFaces& F, std::vector<double>& G) {
  const MgLevel& L = F.L;
  const int nx = L.nx, ny = L.ny, n = L.nx * L.ny * L.nz;
  std::vector<double> A((size_t)n * n, 0.0);
  const int off[6] = {-1, 1, -nx, nx, -nx * ny, nx * ny};
  for (int c = 0; c < n; ++c) {
    if (!F.active(c)) {
      A[(size_t)c * n + c] = 1.0;
      continue;
    }
    for (int s = 0; s < 6; ++s) {
      const double w = F.weight(c, s);
      if (w == 0.0) continue;
      A[(size_t)c * n + c] += w;
      A[(size_t)c * n + c + off[s]] -= w;
    }
  }
  // regions of the active cells (breadth-first through open faces)
  std::vector<int> region(n, -1), queue;
  int regions = 0;
  for (int seed = 0; seed < n; ++seed) {
    if (!F.active(seed) || region[seed] >= 0) continue;
    region[seed] = regions;
    queue.assign(1, seed);
    for (size_t h = 0; h < queue.size(); ++h) {
      const int c = queue[h];
      for (int s = 0; s < 6; ++s)
        if (F.weight(c, s) != 0.0 && region[c + off[s]] < 0) {
          region[c + off[s]] = regions;
          queue.push_back(c + off[s]);
        }
    }
    ++regions;
  }
  if (regions == 0) return false;
  std::vector<double> dsum(regions, 0.0);
  std::vector<int> na(regions, 0);
  for (int c = 0; c < n; ++c)
    if (region[c] >= 0) {
      dsum[region[c]] += A[(size_t)c * n + c];
      ++na[region[c]];
    }
  for (int r = 0; r < n; ++r)
    if (region[r] >= 0) {
      const int k = region[r];
      const double shift = dsum[k] > 0.0 ? dsum[k] / na[k] / na[k] : 1.0;
      for (int c = 0; c < n; ++c)
        if (region[c] == k) A[(size_t)r * n + c] += shift;
    }
  // Cholesky A = L L^T (lower, in place)
  for (int j = 0; j < n; ++j) {
    double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    A[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double v = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) v -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = v / d;
    }
  }
  // M^-1 column by column: solve L L^T x = e_c (column-major result)
  std::vector<double> inv((size_t)n * n, 0.0);
  std::vector<double> y(n);
  for (int col = 0; col < n; ++col) {
    for (int i = 0; i < n; ++i) {
      double v = i == col ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) v -= A[(size_t)i * n + k] * y[k];
      y[i] = v / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < n; ++k) v -= A[(size_t)k * n + i] * y[k];
      y[i] = v / A[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[(size_t)col * n + i] = y[i];
  }
  // G = M^-1 P: from each active column subtract the mean of its region's
  // columns
  std::vector<double> mean_col((size_t)regions * n, 0.0);
  for (int col = 0; col < n; ++col)
    if (region[col] >= 0)
      for (int i = 0; i < n; ++i) mean_col[(size_t)region[col] * n + i] += inv[(size_t)col * n + i];
  for (int k = 0; k < regions; ++k)
    for (int i = 0; i < n; ++i) mean_col[(size_t)k * n + i] /= na[k];
  G.assign((size_t)n * n, 0.0);
  for (int col = 0; col < n; ++col) {
    if (region[col] < 0) continue;
    for (int i = 0; i < n; ++i)
      if (F.active(i)) G[(size_t)col * n + i] = inv[(size_t)col * n + i] - mean_col[(size_t)region[col] * n + i];
  }
  return true;
}

// Obstacle hierarchy: level-0 mask and the coarse face fractions (open
// finest sub-faces / 4^l; stored exactly as counts while 4^l <= 256, else
// rounded to 1/1023 steps with an open face never rounded to 0).
rh_status MgSolver::set_solid(const uint8_t* solid) {
  rh_status st = Solver::set_solid(solid);
  if (st) return st;
  return build_fractions();
}

// Level-0 mask and coarse face fractions: from the obstacle mask, or from
// the plain box when only padded coarsening needs them (levels >= the first
// padded one then carry fractions; the finer ones keep the box rule).
rh_status MgSolver::build_fractions() {
  dirty_ = true;
  ++generation_;
  hfrac_.assign(h_.levels, {});
  for (int l = 0; l < h_.levels; ++l) {
    h_.lv[l].mask = nullptr;
    h_.lv[l].frac = nullptr;
    h_.lv[l].octant = nullptr;
    fscale_[l] = 0.0;
  }
  if (d_frac_) (void)hipFree(d_frac_);
  d_frac_ = nullptr;
  if (!d_mask_ && !padded_level_) return RH_OK;
  if (d_mask_) h_.lv[0].mask = d_mask_;
  if (h_.levels == 1) return RH_OK;
  const int first_frac = d_mask_ ? 1 : padded_level_;
  std::vector<uint8_t> box;
  if (!d_mask_) {  // open faces of the plain box, every cell fluid
    const MgLevel& L = h_.lv[0];
    box.assign((size_t)L.nx * L.ny * L.nz, 0);
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < L.ny; ++j)
        for (int i = 0; i < L.nx; ++i)
          box[i + (size_t)L.nx * (j + (size_t)L.ny * k)] =
              (uint8_t)(kFluidBit | (i > 0) | ((i + 1 < L.nx) << 1) | ((j > 0) << 2) | ((j + 1 < L.ny) << 3) |
                        ((k > 0) << 4) | ((k + 1 < L.nz) << 5));
  }
  const std::vector<uint8_t>& mask0 = d_mask_ ? hmask_ : box;
  // open sub-face counts of each level's three lower faces, fluid flags
  std::vector<uint32_t> cnt[3];
  std::vector<uint8_t> fluid;
  {
    const size_t n = mask0.size();
    for (int a = 0; a < 3; ++a) {
      cnt[a].resize(n);
      for (size_t c = 0; c < n; ++c) cnt[a][c] = (mask0[c] >> (2 * a)) & 1u;
    }
    fluid.resize(n);
    for (size_t c = 0; c < n; ++c) fluid[c] = (mask0[c] & kFluidBit) ? 1 : 0;
  }
  size_t total = 0;
  for (int l = 1; l < h_.levels; ++l) {
    const size_t nc = (size_t)h_.lv[l].nx * h_.lv[l].ny * h_.lv[l].nz;
    total += (nc + 63) / 64 * 64 + (nc + 255) / 256 * 64;  // words + octant bytes
  }
  RH_HIP(hipMalloc(&d_frac_, sizeof(unsigned) * total));
  size_t offset = 0;
  for (int l = 1; l < h_.levels; ++l) {
    const MgLevel& F = h_.lv[l - 1];
    const MgLevel& C = h_.lv[l];
    const size_t nc = (size_t)C.nx * C.ny * C.nz;
    std::vector<uint32_t> cc[3];
    std::vector<uint8_t> cf(nc, 0);
    for (int a = 0; a < 3; ++a) cc[a].assign(nc, 0);
    for (int K = 0; K < C.nz; ++K)
      for (int J = 0; J < C.ny; ++J)
        for (int I = 0; I < C.nx; ++I) {
          const size_t c = I + (size_t)C.nx * (J + (size_t)C.ny * K);
          for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
              for (int dx = 0; dx < 2; ++dx) {
                if (2 * I + dx >= F.nx || 2 * J + dy >= F.ny || 2 * K + dz >= F.nz) continue;  // padding
                const size_t f =
                    (2 * I + dx) + (size_t)F.nx * ((2 * J + dy) + (size_t)F.ny * (2 * K + dz));
                cf[c] |= fluid[f];
                // the child's lower face lies on the coarse lower face of axis a
                if (dx == 0) cc[0][c] += cnt[0][f];
                if (dy == 0) cc[1][c] += cnt[1][f];
                if (dz == 0) cc[2][c] += cnt[2][f];
              }
        }
    const double sub = std::ldexp(1.0, 2 * l);  // finest sub-faces per coarse face
    const bool exact = sub <= 256.0;
    fscale_[l] = exact ? 1.0 / sub : 1.0 / 1023.0;
    std::vector<unsigned> q(nc);
    for (size_t c = 0; c < nc; ++c) {
      unsigned word = cf[c] ? kFracFluid : 0u;
      for (int a = 0; a < 3; ++a) {
        unsigned v = cc[a][c];
        if (!exact && v) v = std::max(1u, (unsigned)std::lround(v * 1023.0 / sub));
        word |= (v & 1023u) << (10 * a);
      }
      q[c] = word;
    }
    // octant bits: fine child parities (x, y, z) -> its trilinear stencil
    // {P, P + s} per axis (s = -1 for even, +1 for odd, mirrored at walls)
    std::vector<uint8_t> oct(nc, 0);
    for (int K = 0; K < C.nz; ++K)
      for (int J = 0; J < C.ny; ++J)
        for (int I = 0; I < C.nx; ++I) {
          uint8_t bits = 0;
          for (int o = 0; o < 8; ++o) {
            const int P[3] = {I, J, K}, n[3] = {C.nx, C.ny, C.nz};
            int lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
              const int nb = P[a] + (((o >> a) & 1) ? 1 : -1);
              lo[a] = std::min(P[a], nb < 0 || nb >= n[a] ? P[a] : nb);
              hi[a] = std::max(P[a], nb < 0 || nb >= n[a] ? P[a] : nb);
            }
            bool all = true;
            for (int e = lo[2]; e <= hi[2] && all; ++e)
              for (int b = lo[1]; b <= hi[1] && all; ++b)
                for (int a = lo[0]; a <= hi[0] && all; ++a)
                  all = cf[a + (size_t)C.nx * (b + (size_t)C.ny * e)] != 0;
            if (all) bits |= (uint8_t)(1u << o);
          }
          oct[I + (size_t)C.nx * (J + (size_t)C.ny * K)] = bits;
        }
    RH_HIP(hipMemcpy(d_frac_ + offset, q.data(), sizeof(unsigned) * nc, hipMemcpyHostToDevice));
    unsigned char* d_oct = (unsigned char*)(d_frac_ + offset + (nc + 63) / 64 * 64);
    RH_HIP(hipMemcpy(d_oct, oct.data(), nc, hipMemcpyHostToDevice));
    if (l >= first_frac) {
      h_.lv[l].frac = d_frac_ + offset;
      h_.lv[l].octant = d_oct;
    } else {
      q.clear();  // a box level: the host view keeps the box rule too
    }
    offset += (nc + 63) / 64 * 64 + (nc + 255) / 256 * 64;
    hfrac_[l].swap(q);
    for (int a = 0; a < 3; ++a) cnt[a].swap(cc[a]);
    fluid.swap(cf);
  }
  return RH_OK;
}

// Rediscretised coefficients per level (spacing doubles each level) and the
// dense coarsest inverse for this density.
rh_status MgSolver::set_density(double density) {
  double h[3] = {grid_.spacing[0], grid_.spacing[1], grid_.spacing[2]};
  for (int l = 0; l < h_.levels; ++l) {
    MgLevel& L = h_.lv[l];
    const double area[3] = {h[1] * h[2], h[0] * h[2], h[0] * h[1]};
    for (int a = 0; a < 3; ++a) L.w[a] = area[a] * (1.0 / (density * h[a]));
    for (int a = 0; a < 3; ++a) L.wf[a] = L.w[a] * (l > 0 ? fscale_[l] : 1.0);
    for (int mask = 0; mask < 64; ++mask) {
      double v = 0.0;
      for (int side = 0; side < 6; ++side)
        if (mask & (1 << side)) v += L.w[side / 2];
      L.diag[mask] = v;
      L.odiag[mask] = v > 0.0 ? 1.0 / v : 0.0;
    }
    for (int a = 0; a < 3; ++a) h[a] *= 2.0;
  }
  if (!dirty_ && d_inv_ && d_h_ && inv_density_ == density) {
    h_.coarse_inv = d_inv_;
    h_.d_self = d_h_;
    return RH_OK;
  }
  const MgLevel& last = h_.lv[h_.levels - 1];
  const int n = last.nx * last.ny * last.nz;
  h_.coarse_inv = nullptr;
  std::vector<double> inv;
  const int cl = h_.levels - 1;
  const LevelFaces faces{last, (cl == 0 && !hmask_.empty()) ? &hmask_ : nullptr,
                         (cl > 0 && h_.lv[cl].frac) ? &hfrac_[cl] : nullptr, fscale_[cl]};
  if (n <= kMgDenseMax && dense_coarse_inverse(faces, inv)) {
    if (d_inv_ && inv_cap_ < inv.size()) {
      (void)hipFree(d_inv_);
      d_inv_ = nullptr;
    }
    if (!d_inv_) {
      RH_HIP(hipMalloc(&d_inv_, sizeof(double) * inv.size()));
      inv_cap_ = inv.size();
    }
    RH_HIP(hipMemcpy(d_inv_, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice));
    h_.coarse_inv = d_inv_;
  }
  // Tail: the first level below the finest whose x, y, b, fractions and
  // octants fit the tail's LDS (the tail needs the dense coarsest inverse);
  // 1/D is staged as well where the remaining room allows (box tables
  // first, then per-cell values from the largest level down).
  h_.tail_level = 0;
  h_.tail_lds = 0;
  for (int l = 0; l < kMgMaxLevels; ++l) h_.tail_odc[l] = nullptr;
  if (h_.coarse_inv) {
    const size_t s = config_.precision == RH_F64 ? 8 : 4;
    const size_t limit = (size_t)mg_tail_lds_limit();
    auto a16 = [](size_t b) { return (b + 15) / 16 * 16; };
    // per tail level: x, y, b; fractions if the level has them; octants
    // where the level is the coarse side of a tail prolongation
    auto sizes = [&](int l, int m, size_t q[6]) {
      const size_t n = (size_t)h_.lv[m].nx * h_.lv[m].ny * h_.lv[m].nz;
      q[0] = q[1] = q[2] = a16(s * n);
      q[3] = h_.lv[m].frac ? a16(4 * n) : 0;
      q[4] = (m > l && h_.lv[m].octant) ? a16(n) : 0;
      q[5] = 0;
    };
    for (int l = 1; l < h_.levels; ++l) {
      size_t q[kMgMaxLevels][6];
      size_t bytes = 0;
      for (int m = l; m < h_.levels; ++m) {
        sizes(l, m, q[m]);
        for (int k = 0; k < 5; ++k) bytes += q[m][k];
      }
      if (bytes > limit) continue;
      for (int m = l; m < h_.levels; ++m)
        if (!h_.lv[m].frac && bytes + 64 * 8 <= limit) {
          q[m][5] = 64 * 8;
          bytes += q[m][5];
        }
      for (int m = l; m < h_.levels; ++m) {
        const size_t n = (size_t)h_.lv[m].nx * h_.lv[m].ny * h_.lv[m].nz;
        if (h_.lv[m].frac && bytes + a16(8 * n) <= limit) {
          q[m][5] = a16(8 * n);
          bytes += q[m][5];
        }
      }
      size_t off = 0;
      for (int m = l; m < h_.levels; ++m)
        for (int k = 0; k < 6; ++k) {
          h_.tail_off[m][k] = q[m][k] ? (int)off : -1;
          off += q[m][k];
        }
      h_.tail_level = l;
      h_.tail_lds = (int)bytes;
      // per-cell 1/D of the staged fraction levels: lv_odiag's fraction
      // form (IEEE products, sum and quotient: the device's values)
      std::vector<double> od;
      size_t at[kMgMaxLevels] = {};
      for (int m = l; m < h_.levels; ++m) {
        if (!h_.lv[m].frac || !q[m][5]) continue;
        const MgLevel& L = h_.lv[m];
        const std::vector<unsigned>& fr = hfrac_[m];
        at[m] = od.size();
        for (int k = 0; k < L.nz; ++k)
          for (int j = 0; j < L.ny; ++j)
            for (int i = 0; i < L.nx; ++i) {
              const size_t c = i + (size_t)L.nx * (j + (size_t)L.ny * k);
              const unsigned w = fr[c];
              const unsigned xp = i + 1 < L.nx ? fr[c + 1] & 1023u : 0u;
              const unsigned yp = j + 1 < L.ny ? (fr[c + L.nx] >> 10) & 1023u : 0u;
              const unsigned zp = k + 1 < L.nz ? (fr[c + (size_t)L.nx * L.ny] >> 20) & 1023u : 0u;
              const int sx = (int)(w & 1023u) + (int)xp, sy = (int)((w >> 10) & 1023u) + (int)yp,
                        sz = (int)((w >> 20) & 1023u) + (int)zp;
              const double d = L.wf[0] * (double)sx + L.wf[1] * (double)sy + L.wf[2] * (double)sz;
              od.push_back(d > 0.0 ? 1.0 / d : 0.0);
            }
      }
      if (!od.empty()) {
        if (d_odc_ && odc_cap_ < od.size()) {
          (void)hipFree(d_odc_);
          d_odc_ = nullptr;
        }
        if (!d_odc_) {
          RH_HIP(hipMalloc(&d_odc_, sizeof(double) * od.size()));
          odc_cap_ = od.size();
        }
        RH_HIP(hipMemcpy(d_odc_, od.data(), sizeof(double) * od.size(), hipMemcpyHostToDevice));
        for (int m = l; m < h_.levels; ++m)
          if (h_.lv[m].frac && q[m][5]) h_.tail_odc[m] = d_odc_ + at[m];
      }
      break;
    }
  }
  if (!d_h_) RH_HIP(hipMalloc(&d_h_, sizeof(MgHierarchy)));
  h_.d_self = d_h_;
  RH_HIP(hipMemcpy(d_h_, &h_, sizeof(MgHierarchy), hipMemcpyHostToDevice));
  inv_density_ = density;
  dirty_ = false;
  ++generation_;
  return RH_OK;
}

template <typename S>
rh_status MgSolver::run(const DevGrid& g, const double* d_rhs, const rh_pressure_settings& st,
                        double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel,
                        void* ctx) {
  hipStream_t s = scratch_.stream;
  // the caller's work to enqueue after the solve's kernels (taken now, so
  // an early return drops it)
  std::function<void()> tail = std::move(tail_);
  tail_ = nullptr;
  PcgControl h;
  init_control(h, st, dt_, g.volume);
  std::memcpy(scratch_.host, &h, sizeof(h));
  RH_HIP(hipMemcpyAsync(scratch_.ctl, scratch_.host, sizeof(h), hipMemcpyHostToDevice, s));
  RH_HIP(hipEventRecord(scratch_.ev0, s));
  // compatibility + threshold, p = 0, initial true residual (it = 0)
  const bool sums_ready = rhs_sums_for_ == d_rhs;
  rhs_sums_for_ = nullptr;
  if (sums_ready)
    launch_rhs_sums_finish(g, d_rhs, scratch_.ctl, scratch_.pdd, scratch_.pmx, s);
  else
    launch_rhs_sums(g, d_rhs, scratch_.ctl, scratch_.pdd, scratch_.pmx, s);
  const bool x0_implicit = sizeof(S) == 4 && mg_first_from_b(g, h_, d_pressure, d_rhs);
  // Four-cell FP32 path: p = 0 is never read -- the first residual is b
  // and the first cycle's update writes p = 0 + x over every cell -- so p
  // is filled only if no cycle runs (below).
  // A seed starts the defect correction from it (the first residual reads p).
  const bool p_zero = !seed_ && x0_implicit && mg_quad_update(h_, d_pressure);
  if (seed_)
    launch_seed(g, seed_, d_pressure, scratch_.ctl, scratch_.pdd, scratch_.pmx, s);
  else if (!p_zero)
    RH_HIP(hipMemsetAsync(d_pressure, 0, sizeof(double) * (size_t)g.nx * g.ny * g.nz, s));
  // Four-cell FP32 path: each outer residual's acceptance runs inside the
  // next cycle's first sweep (MgFold), the last one after the loop.
  const bool fold = x0_implicit;
  int* progress = scratch_.progress_dev;
  *(volatile int*)scratch_.progress = 0;
  const int nb = mg_outer_blocks<S>(g, h_, d_pressure, d_rhs);
  mg_outer_residual<S>(g, h_, d_pressure, d_rhs, scratch_.ctl, scratch_.pdd, scratch_.pmx, 0, s, x0_implicit,
                       p_zero, !fold, progress);
  RH_HIP(hipGetLastError());
  bool cancelled = false;
  const int poll = st.poll_interval > 0 ? st.poll_interval : 2;
  int enqueued = 0;
  rh_status rs = drive_paced(
      h.max_iterations, poll, std::max(0, last_cycles_ - 1),
      [&](int it) {
        const MgFold f{scratch_.pdd, scratch_.pmx, nb, FinMgTrue{scratch_.ctl, it - 1, progress}};
        mg_cycle<S>(h_, d_pressure, scratch_.ctl, it, s, x0_implicit, p_zero && it == 1, fold ? &f : nullptr);
        mg_outer_residual<S>(g, h_, d_pressure, d_rhs, scratch_.ctl, scratch_.pdd, scratch_.pmx, it, s,
                             x0_implicit, false, !fold, progress);
      },
      &cancelled, cancel, ctx, &enqueued);
  if (rs) return rs;
  // the last enqueued outer residual's acceptance (a no-op once decided)
  if (fold) launch_finish<1, 2>(scratch_.pdd, scratch_.pmx, nb, FinMgTrue{scratch_.ctl, enqueued, progress}, s);
  // Fix the gauge p[0] = 0 (the reference's eliminated gauge cell), p = 0
  // when no cycle ran, and report the true residual of the returned
  // pressure -- all enqueued before the one read-back of the control block
  // and the report's sums (the gauge shift acts only on an OK or
  // iteration-limit outcome; the report is used only on OK).
  mg_regauge(g, d_pressure, scratch_.ctl, s, p_zero);
  if (h_.quad && mg_quad_outer_ok(g, h_.lv[0], d_pressure, d_rhs))
    launch_residual_report_q(g, d_pressure, d_rhs, scratch_.reduce(0), s);
  else
    launch_residual_report(g, d_pressure, d_rhs, scratch_.reduce(0), s);
  RH_HIP(hipEventRecord(scratch_.ev1, s));  // the solve's time ends here, before the caller's tail
  if (tail) tail();
  PcgControl c;
  double out[3];
  rs = finish_report(report, cancelled, &c, out, true);
  if (c.done) last_cycles_ = c.iteration;
  if (rs != RH_OK) return rs;
  if (report) {
    report->true_residual_l2 = std::sqrt(out[0]);
    report->true_residual_max = out[1];
    report->predicted_divergence_max = out[1] * dt_ / g.volume;
    report->seconds += 0.0;
  }
  if (!(std::sqrt(out[0]) <= c.threshold) || !(out[1] * dt_ / g.volume <= c.div_limit)) {
    set_last_error("gauge shift moved the residual past the threshold", "MgSolver::run");
    return RH_QUALITY_GATE;
  }
  return RH_OK;
}

rh_status MgSolver::solve(double density, const double* d_rhs, double dt,
                          const rh_pressure_settings& st, double* d_pressure,
                          rh_pressure_report* report, rh_cancel_fn cancel, void* ctx) {
  rh_status s = check_and_begin(density, d_rhs, dt, st, d_pressure, report, cancel, ctx);
  if (s) return s;
  const DevGrid g = dev_grid(density);
  if ((s = set_density(density))) return s;
  if (config_.precision == RH_F64) return run<double>(g, d_rhs, st, d_pressure, report, cancel, ctx);
  return run<float>(g, d_rhs, st, d_pressure, report, cancel, ctx);
}

}  // namespace rh
