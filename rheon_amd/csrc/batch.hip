// Batched projection steps over many device-resident scenes.
//
// Every stage of the step (validation, advection, RHS, the multigrid solve
// with its defect-correction loop, gauge fix, quality gate, correction,
// divergence and Courant gates, commit) is one launch over all scenes: the
// grid is the concatenation of each scene's own block decomposition and each
// block runs the one-scene kernel body for its scene (see batch.hpp).  All
// decisions of the one-scene step (src/simulation.rs:988-1277) are taken on
// the device per scene, so one step is a fixed launch sequence: the solve
// enqueues `max_cycles` V-cycles and a converged scene's blocks return at
// once.  The sequence can be captured once in a hipGraph and replayed per
// step; the host reads the per-scene reports after the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "handles.hpp"
#include "kernels.hpp"
#include "mg.hpp"
#include "runtime.hpp"

namespace rh {

// Per-scene record of the batched step (device memory).
struct BatchScene {
  PcgControl tmpl;   // solver control block at the start of a solve
  PcgControl* ctl;   // the scene solver's control block
  unsigned* flags;   // validation / arithmetic flags
  double* scal;      // [0..2] residual report, [3] max|div u|, [4..7] rate
  const double* p;
  const DevGrid* g;  // the scene's operator record (regions of a split fluid)
  double volume, actual_limit, max_courant, dt;
  int gauge;
  int status;
  int accepted;
  int pad;
  rh_pressure_report rep;
  double actual_div, courant;
};

namespace {

__device__ int status_of_flags(unsigned f) {
  if (f & kFlagNonfiniteInput) return RH_NONFINITE_INPUT;
  if (f & kFlagNonzeroWall) return RH_NONZERO_WALL_VELOCITY;
  if (f & kFlagArithmetic) return RH_ARITHMETIC_FAILURE;
  return RH_OK;
}

__global__ void kb_begin_step(BatchScene* S, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  *S[s].flags = 0u;
  S[s].status = RH_OK;
  S[s].accepted = 0;
  S[s].actual_div = 0.0;
  S[s].courant = 0.0;
  S[s].ctl->spare[2] = 0.0;
}

// Control block of the solve; a scene whose fields failed validation is
// marked done with that status, so every solver stage skips it.
__global__ void kb_begin_solve(BatchScene* S, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  PcgControl c = S[s].tmpl;
  const unsigned f = *S[s].flags;
  if (f) {
    c.done = 1;
    c.status = status_of_flags(f);
    S[s].status = c.status;
  }
  *S[s].ctl = c;
}

// After the last cycle: the report, and the gauge shift p - p[gauge] for a
// converged (or iteration-limited) solve (the one-scene MgSolver::run).
__global__ void kb_end_solve(BatchScene* S, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  PcgControl* c = S[s].ctl;
  rh_pressure_report& r = S[s].rep;
  r.iterations = c->iteration;
  r.true_residual_l2 = c->true_l2;
  r.true_residual_max = c->true_max;
  r.predicted_divergence_max = c->pred_div;
  r.error_iteration = c->error_iteration;
  r.error_value = c->error_value;
  r.threshold = c->threshold;
  r.seconds = 0.0;
  int st = c->done ? c->status : RH_EXECUTION;
  if (!c->done) {
    c->done = 1;
    c->status = st;
  }
  const bool shift = st == RH_OK || st == RH_ITERATION_LIMIT;
  c->spare[0] = shift ? S[s].p[S[s].gauge] : 0.0;
  c->spare[1] = shift ? 1.0 : 0.0;
  const DevGrid& g = *S[s].g;
  if (shift && g.ncomp > 1)
    for (int k = 0; k < g.ncomp; ++k) g.comp_shift[k] = S[s].p[g.comp_gauge[k]];
  if (S[s].status == RH_OK) S[s].status = st;
}

// Quality gate on the gauge-shifted pressure (MgSolver::run's final check).
__global__ void kb_quality(BatchScene* S, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (S[s].status != RH_OK) return;
  const double l2 = sqrt(S[s].scal[0]), mx = S[s].scal[1];
  const double div = mx * S[s].dt / S[s].volume;
  S[s].rep.true_residual_l2 = l2;
  S[s].rep.true_residual_max = mx;
  S[s].rep.predicted_divergence_max = div;
  if (S[s].scal[2] != 0.0 || !(l2 <= S[s].rep.threshold) || !(div <= S[s].tmpl.div_limit)) {
    S[s].status = RH_QUALITY_GATE;
    S[s].ctl->status = RH_QUALITY_GATE;
  }
}

// Actual divergence and Courant gates, then the commit decision
// (src/simulation.rs:1235-1277).
__global__ void kb_gate(BatchScene* S, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (S[s].status != RH_OK) return;
  const unsigned f = *S[s].flags;
  if (f) {
    S[s].status = status_of_flags(f);
    return;
  }
  const double actual = S[s].scal[3];
  const double courant = S[s].dt * S[s].scal[4];
  S[s].actual_div = actual;
  S[s].courant = courant;
  if (actual > S[s].actual_limit) {
    S[s].status = RH_DIVERGENCE_LIMIT;
    return;
  }
  if (S[s].max_courant > 0.0 && (!isfinite(courant) || courant > S[s].max_courant)) {
    S[s].status = RH_COURANT_LIMIT;
    return;
  }
  S[s].accepted = 1;
  S[s].ctl->spare[2] = 1.0;
}

// Commit: accepted faces <- candidate faces (blockIdx.y = component).
__global__ void __launch_bounds__(kBlock)
kb_commit(const BFaceArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  const int s = batch_scene(prefix, n, blockIdx.x);
  const BFaceArgs& a = A[s];
  if (a.ctl->spare[2] == 0.0) return;
  const VBlock vb = batch_block(prefix, s);
  const DevGrid& g = *a.g;
  const int ax = blockIdx.y;
  const float* src = ax == 0 ? a.u : (ax == 1 ? a.v : a.w);
  float* dst = ax == 0 ? a.ou : (ax == 1 ? a.ov : a.ow);
  const int fx = g.nx + (ax == 0), fy = g.ny + (ax == 1), fz = g.nz + (ax == 2);
  const RowRange rr = block_rows(fy * fz, vb);
  for (int row = rr.begin + threadIdx.y; row < rr.end; row += kRowsPerBlock)
    for (int i = threadIdx.x; i < fx; i += kLanesX) dst[i + fx * row] = src[i + fx * row];
}

// Finishers: one block per scene, skipped when the scene has no producer
// blocks in this stage or its gate says so.
template <int ND, int NM, typename Fin>
struct BFin {
  const DD* pdd;
  const double* pmx;
  int nb;
  int pad;
  Fin fin;
};
template <int ND, int NM, typename Fin>
__global__ void __launch_bounds__(kBlock) kb_finish(const BFin<ND, NM, Fin>* __restrict__ A) {
  const BFin<ND, NM, Fin>& a = A[blockIdx.x];
  if (a.nb == 0) return;
  finish_body<ND, NM, Fin>(a.pdd, a.pmx, a.nb, a.fin);
}

// Runs `f` only while the scene's solve is live.
template <typename F>
struct FinNotDone {
  const PcgControl* ctl;
  F f;
  __device__ bool active() const { return !ctl->done && f.active(); }
  __device__ void operator()(const double* s, const double* m) const { f(s, m); }
};
// Runs `f` only when the scene's solve (and gates so far) succeeded.
template <typename F>
struct FinStatusOk {
  const PcgControl* ctl;
  F f;
  __device__ bool active() const { return ctl->status == RH_OK && f.active(); }
  __device__ void operator()(const double* s, const double* m) const { f(s, m); }
};

}  // namespace

// ------------------------------------------------------------------ host

namespace {

enum class Op {
  kBeginStep, kValidateAcc, kAdvect, kAdvectLattice, kValidateCand, kRhs, kBeginSolve, kRhsSums, kFinRhs, kCompSums, kCompCheck, kFill,
  kOuterQ, kOuterS, kFinOuter, kSmoothQ, kSmoothQUpdate, kSmoothQFirst, kSmoothQFrac, kSmoothS, kRestrictQ, kRestrictQFrac, kRestrictS,
  kProlongQ, kProlongS, kTail, kCoarseDense, kEndSolve, kRegauge, kReport, kFinReport,
  kQuality, kCorrect, kDivMax, kFinDiv, kRate, kFinRate, kGate, kCommit
};

struct Launch {
  Op op;
  size_t prefix_off = 0, args_off = 0;  // offsets in the table arena
  int blocks = 0;
  int extra = 0;  // dynamic LDS bytes (tail)
};

// Host staging of the device tables: records, prefixes and finisher
// records, appended with alignment and copied once.
struct Tables {
  std::vector<unsigned char> bytes;
  size_t add(const void* data, size_t size, size_t align = 16) {
    size_t off = (bytes.size() + align - 1) / align * align;
    bytes.resize(off + size);
    std::memcpy(bytes.data() + off, data, size);
    return off;
  }
};

}  // namespace

class Batch {
 public:
  ~Batch();
  rh_status create(const rh_grid* grids, size_t n, const rh_batch_config& cfg);
  rh_status step(int steps, rh_step_report* reports, rh_status* statuses);
  // Enqueue `steps` steps (tables rebuilt if needed).  With a graph the
  // whole step (max_cycles V-cycles) is one replay and nothing waits.
  // Without one, the V-cycles go out in batches of kPollCycles and the host
  // stops enqueuing once every scene's solve is done (polled one batch
  // behind); `host_work`, if given, runs once while the device works.
  rh_status launch(int steps, const std::function<rh_status()>& host_work = nullptr);
  rh_status wait(rh_step_report* reports, rh_status* statuses);
  double last_seconds() const { return last_seconds_; }  // device time of the last launch
  double last_build_seconds() const { return last_build_seconds_; }  // host time of the last table build
  rh_scene* scene(size_t i) { return i < scenes_.size() ? scenes_[i] : nullptr; }
  // Non-owning membership (serving): the batch steps exactly these scenes.
  rh_status set_members(const std::vector<rh_scene*>& scenes, const std::vector<double>& dt);
  rh_status init_empty(const rh_batch_config& cfg);
  size_t size() const { return scenes_.size(); }
  uint64_t bytes() const;
  rh_status set_dt_all(const double* dt, size_t n);
  const rh_batch_config& config() const { return cfg_; }
  rh_status set_variant(uint32_t variant);
  rh_status set_tuning(const rh_batch_tuning& t);
  rh_status autotune(int steps, rh_batch_tuning* table, size_t cap, size_t* count, rh_batch_tuning* chosen);

 private:
  rh_status build();
  rh_status reserve(size_t n);
  void run_program(hipStream_t s);
  void run_range(size_t begin, size_t end, hipStream_t s);
  rh_status run_polled(const std::function<rh_status()>* host_work);
  void release_graph();
  MgSolver* mg(size_t i) { return static_cast<MgSolver*>(scenes_[i]->solver->impl.get()); }

  std::vector<rh_scene*> scenes_;                 // members (owned or not)
  std::vector<std::unique_ptr<rh_scene>> owned_;  // scenes this batch created
  std::vector<double> dt_;
  std::vector<uint64_t> gens_;
  rh_batch_config cfg_{};
  uint32_t variant_ = 0;
  int block_div_ = 1;  // each scene's blocks per stage = its one-scene count / block_div_
  hipStream_t stream_ = nullptr;
  DevGrid* d_grids_ = nullptr;
  BatchScene* d_scenes_ = nullptr;
  size_t cap_ = 0;  // scene records the device tables hold
  std::vector<BatchScene> h_scenes_;
  BatchScene* pinned_ = nullptr;  // pinned readback of the per-scene records
  size_t pinned_cap_ = 0;
  unsigned char* d_tables_ = nullptr;
  size_t tables_bytes_ = 0;
  std::vector<Launch> program_;
  std::vector<size_t> cycle_at_;  // program_ index of V-cycle it's first launch (it = 1..max_cycles_)
  size_t post_at_ = 0;            // program_ index of the first launch after the solve
  int* d_active_ = nullptr;       // [2] scenes still solving (device)
  int* h_active_ = nullptr;       // [2] their pinned host copies
  hipEvent_t poll_ev_[2] = {nullptr, nullptr};
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  // use_graph == 2: the step as segments (prologue, one cycle, one cycle
  // then the count of scenes still solving, the post stage) and the
  // cycles the last step needed
  hipGraphExec_t seg_[4] = {nullptr, nullptr, nullptr, nullptr};
  int* cycle_state_ = nullptr;      // [2] pinned: scenes still solving, largest cycle count
  int* cycle_state_dev_ = nullptr;  // its device address
  int next_cycles_ = 0;
  hipEvent_t seg_ev_ = nullptr;
  rh_status run_segments(const std::function<rh_status()>* host_work);
  rh_status capture(size_t begin, size_t end, bool state, hipGraphExec_t* out);
  bool built_ = false;
  bool adv_off32_ = false;  // every scene's face arrays take 32-bit offsets
  int max_cycles_ = 0;
  int launched_steps_ = 0;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  double last_seconds_ = 0.0;
  double last_build_seconds_ = 0.0;
};

Batch::~Batch() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : poll_ev_)
    if (e) (void)hipEventDestroy(e);
  if (d_active_) (void)hipFree(d_active_);
  if (h_active_) (void)hipHostFree(h_active_);
  if (pinned_) (void)hipHostFree(pinned_);
  if (cycle_state_) (void)hipHostFree(cycle_state_);
  if (seg_ev_) (void)hipEventDestroy(seg_ev_);
  release_graph();
  if (d_tables_) (void)hipFree(d_tables_);
  if (d_grids_) (void)hipFree(d_grids_);
  if (d_scenes_) (void)hipFree(d_scenes_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::release_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  for (auto& e : seg_) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
}

uint64_t Batch::bytes() const {
  uint64_t b = tables_bytes_ + scenes_.size() * (sizeof(DevGrid) + sizeof(BatchScene));
  for (const auto& sc : scenes_) b += sc->bytes;
  return b;
}

rh_status Batch::create(const rh_grid* grids, size_t n, const rh_batch_config& cfg) {
  cfg_ = cfg;
  const rh_scene_config& sc = cfg.scene;
  if (n == 0 || !grids || cfg.use_graph < 0 || cfg.use_graph > 2) return RH_INVALID_ARGUMENT;
  if (sc.precision != RH_F32 || sc.solver.method != RH_MIXED_MG || sc.solver.precision != RH_F32) {
    set_last_error("batched steps run f32 faces with the mixed-precision multigrid solve", "rh_batch_create");
    return RH_UNSUPPORTED;
  }
  variant_ = sc.solver.kernel_variant;
  for (size_t i = 0; i < n; ++i) {
    rh_scene* raw = nullptr;
    rh_status s = rh_scene_create(&grids[i], &sc, &raw);
    if (s) return s;
    owned_.emplace_back(raw);
    raw->batched = true;
    raw->p_state = 0;
    scenes_.push_back(raw);
    if (mg(i)->levels() < 2) {
      set_last_error("a batched scene needs at least two multigrid levels", "rh_batch_create");
      return RH_UNSUPPORTED;
    }
  }
  dt_.assign(n, 0.0);
  gens_.assign(n, ~0ull);
  RH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  return reserve(n);
}

rh_status Batch::reserve(size_t n) {
  if (n <= cap_) return RH_OK;
  if (d_grids_) (void)hipFree(d_grids_);
  if (d_scenes_) (void)hipFree(d_scenes_);
  d_grids_ = nullptr;
  d_scenes_ = nullptr;
  cap_ = std::max<size_t>(n, 2 * cap_);
  RH_HIP(hipMalloc(&d_grids_, sizeof(DevGrid) * cap_));
  RH_HIP(hipMalloc(&d_scenes_, sizeof(BatchScene) * cap_));
  return RH_OK;
}

rh_status Batch::init_empty(const rh_batch_config& cfg) {
  cfg_ = cfg;
  const rh_scene_config& sc = cfg.scene;
  if (sc.precision != RH_F32 || sc.solver.method != RH_MIXED_MG || sc.solver.precision != RH_F32)
    return RH_UNSUPPORTED;
  variant_ = sc.solver.kernel_variant;
  RH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  return RH_OK;
}

rh_status Batch::set_members(const std::vector<rh_scene*>& scenes, const std::vector<double>& dt) {
  if (scenes.size() != dt.size()) return RH_INVALID_ARGUMENT;
  for (rh_scene* s : scenes) {
    if (!s || s->config.precision != RH_F32 || s->config.solver.method != RH_MIXED_MG ||
        s->config.solver.precision != RH_F32)
      return RH_UNSUPPORTED;
    if (static_cast<MgSolver*>(s->solver->impl.get())->levels() < 2) return RH_UNSUPPORTED;
  }
  for (rh_scene* s : scenes) {
    if (s->tracer[0]) {
      set_last_error("scenes stepped by a batch carry no tracer", "rh_batch");
      return RH_UNSUPPORTED;
    }
    s->batched = true;
    s->p_state = 0;
    s->init_policy = RH_INIT_COLD;
  }
  if (scenes != scenes_ || dt != dt_) {
    built_ = false;
    scenes_ = scenes;
    dt_ = dt;
    gens_.assign(scenes_.size(), ~0ull);
  }
  return reserve(scenes_.size());
}

rh_status Batch::set_dt_all(const double* dt, size_t n) {
  if (!dt || n != scenes_.size()) return RH_INVALID_ARGUMENT;
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(dt[i]) || dt[i] <= 0.0) return RH_INVALID_TIME_STEP;
    if (dt_[i] != dt[i]) built_ = false;
    dt_[i] = dt[i];
  }
  return RH_OK;
}

rh_status Batch::set_tuning(const rh_batch_tuning& t) {
  if (t.block_divisor < 1 || t.block_divisor > 64) return RH_INVALID_ARGUMENT;
  if (t.kernel_variant != variant_ || t.block_divisor != block_div_) built_ = false;
  variant_ = t.kernel_variant;
  block_div_ = t.block_divisor;
  return RH_OK;
}

// Times every candidate launch shape on this batch's own scenes and keeps
// the fastest.  Candidates: MG kernel variant (four cells per lane where a
// level allows / one cell per lane) x blocks per scene (the one-scene count
// divided by 1, 2, 4, 8).  Each candidate starts from the same accepted
// faces (saved first, restored after every run), runs one untimed step
// (table build and graph capture) and `steps` timed steps.
rh_status Batch::autotune(int steps, rh_batch_tuning* table, size_t cap, size_t* count,
                          rh_batch_tuning* chosen) {
  const size_t n = scenes_.size();
  if (steps < 1) steps = 1;
  for (size_t i = 0; i < n; ++i)
    if (!(dt_[i] > 0.0)) return RH_INVALID_TIME_STEP;
  size_t total = 0;
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) total += scenes_[i]->flen[a] * 4;
  void* save = nullptr;
  RH_HIP(hipMalloc(&save, total));
  auto copy_faces = [&](bool restore) -> rh_status {
    size_t off = 0;
    for (size_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        const size_t b = scenes_[i]->flen[a] * 4;
        char* keep = (char*)save + off;
        RH_HIP(hipMemcpyAsync(restore ? scenes_[i]->acc[a] : keep, restore ? keep : scenes_[i]->acc[a], b,
                              hipMemcpyDeviceToDevice, stream_));
        off += b;
      }
    RH_HIP(hipStreamSynchronize(stream_));
    return RH_OK;
  };
  rh_status st = copy_faces(false);
  std::vector<rh_batch_tuning> tried;
  const uint32_t variants[2] = {0u, RH_VARIANT_SCALAR_MG};
  const int divisors[6] = {1, 2, 4, 8, 16, 32};
  for (uint32_t v : variants)
    for (int d : divisors) {
      if (st) break;
      rh_batch_tuning t{v, d, 0.0};
      if ((st = set_tuning(t))) break;
      if ((st = copy_faces(true))) break;
      if ((st = step(1, nullptr, nullptr))) break;  // build + capture
      if ((st = copy_faces(true))) break;
      const auto t0 = std::chrono::steady_clock::now();
      if ((st = step(steps, nullptr, nullptr))) break;
      t.seconds_per_step =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / steps;
      tried.push_back(t);
    }
  rh_status rs = copy_faces(true);
  (void)hipFree(save);
  if (st) return st;
  if (rs) return rs;
  if (tried.empty()) return RH_EXECUTION;
  const rh_batch_tuning best = *std::min_element(
      tried.begin(), tried.end(),
      [](const rh_batch_tuning& a, const rh_batch_tuning& b) { return a.seconds_per_step < b.seconds_per_step; });
  if (count) *count = tried.size();
  if (table)
    for (size_t i = 0; i < tried.size() && i < cap; ++i) table[i] = tried[i];
  if (chosen) *chosen = best;
  return set_tuning(best);
}

rh_status Batch::set_variant(uint32_t variant) {
  if (variant == variant_) return RH_OK;
  variant_ = variant;
  built_ = false;
  return RH_OK;
}

// Builds the launch tables and the step program from the scenes' current
// device state (hierarchies, masks, buffers).
rh_status Batch::build() {
  const size_t n = scenes_.size();
  h_scenes_.assign(n, BatchScene{});
  if (gens_.size() != n) gens_.assign(n, ~0ull);
  const rh_scene_config& sc = cfg_.scene;
  const double density = sc.density;
  release_graph();
  // device-side operator data of every scene, then the DevGrid records
  std::vector<DevGrid> grids(n);
  for (size_t i = 0; i < n; ++i) {
    rh_status s = mg(i)->prepare(density);
    if (s) return s;
    grids[i] = mg(i)->dev_grid(density);
  }
  RH_HIP(hipMemcpy(d_grids_, grids.data(), sizeof(DevGrid) * n, hipMemcpyHostToDevice));
  int pre = 0, post = 0, cpre = 0, cpost = 0;
  {
    rh_status s = MgSolver::plan_sweeps(sc.solver, &pre, &post, &cpre, &cpost);
    if (s) return s;
  }
  int64_t cycles = sc.pressure.max_iterations;
  if (cfg_.max_cycles > 0) cycles = std::min<int64_t>(cycles, cfg_.max_cycles);
  if (cycles < 1) cycles = 1;
  max_cycles_ = (int)cycles;
  // per-scene records
  for (size_t i = 0; i < n; ++i) {
    rh_scene* s = scenes_[i];
    Scratch& scr = s->solver->impl->scratch();
    BatchScene& b = h_scenes_[i];
    b = BatchScene{};
    PcgControl& c = b.tmpl;
    std::memset(&c, 0, sizeof(c));
    c.rel = sc.pressure.relative_residual;
    c.abs_tol = sc.pressure.absolute_residual;
    c.div_limit = sc.pressure.divergence_limit;
    c.dt = dt_[i];
    c.volume = grids[i].volume;
    c.max_iterations = max_cycles_;
    b.ctl = scr.ctl;
    b.flags = scr.flags;
    b.scal = scr.scal;
    b.p = s->p;
    b.volume = grids[i].volume;
    b.actual_limit = sc.actual_divergence_limit;
    b.max_courant = sc.max_courant;
    b.dt = dt_[i];
    b.gauge = grids[i].gauge;
    b.g = d_grids_ + i;
  }
  RH_HIP(hipMemcpy(d_scenes_, h_scenes_.data(), sizeof(BatchScene) * n, hipMemcpyHostToDevice));

  Tables T;
  program_.clear();
  const bool quad_ok = (variant_ & RH_VARIANT_SCALAR_MG) == 0;
  auto nblk = [&](int rows) { return std::max(1, grid_blocks(rows) / block_div_); };
  auto qblk = [&](int blocks) { return std::max(1, blocks / block_div_); };  // quad launches
  auto lblk = [&](const MgLevel& L) {  // flat level launches
    return std::max(1, level_blocks((long long)L.nx * L.ny * L.nz) / block_div_);
  };
  auto face_blocks = [&](size_t i) { return nblk((int)((grids[i].ny + 1) * (grids[i].nz + 1))); };
  auto cell_blocks = [&](size_t i) { return nblk(grids[i].ny * grids[i].nz); };
  // a stage: records for all scenes + blocks per scene (0 = not taking part)
  auto stage = [&](Op op, const void* recs, size_t rec_size, const std::vector<int>& blocks, int extra = 0) {
    std::vector<int> prefix(n + 1, 0);
    for (size_t i = 0; i < n; ++i) prefix[i + 1] = prefix[i] + blocks[i];
    Launch L;
    L.op = op;
    L.blocks = prefix[n];
    L.extra = extra;
    if (L.blocks == 0) return;
    L.prefix_off = T.add(prefix.data(), sizeof(int) * (n + 1));
    L.args_off = T.add(recs, rec_size * n);
    program_.push_back(L);
  };
  auto per_scene = [&](Op op) {
    Launch L;
    L.op = op;
    L.blocks = (int)((n + 255) / 256);
    program_.push_back(L);
  };
  // --- field records
  std::vector<BFaceArgs> acc(n), adv(n), cand(n), rhs(n), cor(n), div(n), rate(n), commit(n);
  std::vector<int> fb(n), cb(n), rb(n);
  for (size_t i = 0; i < n; ++i) {
    rh_scene* s = scenes_[i];
    Scratch& scr = s->solver->impl->scratch();
    BFaceArgs base{};
    base.g = d_grids_ + i;
    base.dt = &d_scenes_[i].dt;
    base.flags = scr.flags;
    base.pdd = scr.pdd;
    base.pmx = scr.pmx;
    base.ctl = scr.ctl;
    BFaceArgs a = base;
    a.u = (const float*)s->acc[0];
    a.v = (const float*)s->acc[1];
    a.w = (const float*)s->acc[2];
    a.mode = 0;
    acc[i] = a;
    a.ou = (float*)s->cand[0];
    a.ov = (float*)s->cand[1];
    a.ow = (float*)s->cand[2];
    adv[i] = a;
    BFaceArgs c = base;
    c.u = (const float*)s->cand[0];
    c.v = (const float*)s->cand[1];
    c.w = (const float*)s->cand[2];
    c.mode = 1;
    cand[i] = c;
    c.cells = s->rhs;
    rhs[i] = c;
    BFaceArgs k = base;
    k.ou = (float*)s->cand[0];
    k.ov = (float*)s->cand[1];
    k.ow = (float*)s->cand[2];
    k.cells = s->p;
    cor[i] = k;
    div[i] = cand[i];
    rate[i] = cand[i];
    BFaceArgs m = base;
    m.u = (const float*)s->cand[0];
    m.v = (const float*)s->cand[1];
    m.w = (const float*)s->cand[2];
    m.ou = (float*)s->acc[0];
    m.ov = (float*)s->acc[1];
    m.ow = (float*)s->acc[2];
    commit[i] = m;
    fb[i] = face_blocks(i);
    cb[i] = cell_blocks(i);
    rb[i] = face_blocks(i);
  }
  // --- cell records
  std::vector<BCellArgs> cells(n);
  for (size_t i = 0; i < n; ++i) {
    rh_scene* s = scenes_[i];
    Scratch& scr = s->solver->impl->scratch();
    cells[i] = BCellArgs{d_grids_ + i, s->rhs, s->p, scr.pdd, scr.pmx, scr.ctl};
  }
  // --- finisher records
  using FRhs = FinNotDone<FinRhsSums>;
  using FReport = FinStatusOk<FinStore>;
  auto fin_table = [&](auto make, int nd, int nm, const std::vector<int>& nb) -> size_t {
    (void)nd;
    (void)nm;
    using R = decltype(make(0, 0));
    std::vector<R> recs(n);
    for (size_t i = 0; i < n; ++i) recs[i] = make(i, nb[i]);
    return T.add(recs.data(), sizeof(R) * n);
  };
  auto fin_launch = [&](Op op, size_t off) {
    Launch L;
    L.op = op;
    L.args_off = off;
    L.blocks = (int)n;
    program_.push_back(L);
  };

  // ===== program
  per_scene(Op::kBeginStep);
  stage(Op::kValidateAcc, acc.data(), sizeof(BFaceArgs), fb);
  adv_off32_ = true;
  for (size_t i = 0; i < n; ++i) adv_off32_ = adv_off32_ && face_offsets_32(grids[i], sizeof(float));
  // lattice grids take the cheaper first RK2 stage: one launch for them and
  // one for the others, each kernel built for its own form
  {
    std::vector<int> fb_lat(n, 0), fb_gen(n, 0);
    for (size_t i = 0; i < n; ++i) (grids[i].lattice ? fb_lat : fb_gen)[i] = fb[i];
    stage(Op::kAdvectLattice, adv.data(), sizeof(BFaceArgs), fb_lat);
    stage(Op::kAdvect, adv.data(), sizeof(BFaceArgs), fb_gen);
  }
  stage(Op::kValidateCand, cand.data(), sizeof(BFaceArgs), fb);
  stage(Op::kRhs, rhs.data(), sizeof(BFaceArgs), cb);
  per_scene(Op::kBeginSolve);
  stage(Op::kRhsSums, cells.data(), sizeof(BCellArgs), cb);
  fin_launch(Op::kFinRhs, fin_table(
                              [&](size_t i, int nb) {
                                Scratch& scr = scenes_[i]->solver->impl->scratch();
                                return BFin<3, 1, FRhs>{scr.pdd, scr.pmx, nb, 0, FRhs{scr.ctl, FinRhsSums{scr.ctl}}};
                              },
                              3, 1, cb));
  // per-region compatibility of split-fluid scenes
  std::vector<int> compb(n, 0);
  bool split = false;
  for (size_t i = 0; i < n; ++i)
    if (grids[i].ncomp > 1) {
      compb[i] = grids[i].comp_nchunks;
      split = true;
    }
  if (split) {
    stage(Op::kCompSums, cells.data(), sizeof(BCellArgs), compb);
    Launch L;
    L.op = Op::kCompCheck;
    L.args_off = T.add(cells.data(), sizeof(BCellArgs) * n);
    L.blocks = (int)n;
    program_.push_back(L);
  }
  stage(Op::kFill, cells.data(), sizeof(BCellArgs), cb);

  // outer residual records (quad / scalar split) and the V-cycle stages
  std::vector<BOuterArgs> outer(n);
  std::vector<int> outer_q(n, 0), outer_s(n, 0);
  std::vector<bool> from_b(n, false);  // second finest pre-sweep from b0 alone (x0 not stored)
  for (size_t i = 0; i < n; ++i) {
    const MgHierarchy& h = mg(i)->hierarchy();
    const MgHierarchy* dh = mg(i)->device_hierarchy();
    Scratch& scr = scenes_[i]->solver->impl->scratch();
    from_b[i] = quad_ok && mg_first_from_b(grids[i], h, scenes_[i]->p, scenes_[i]->rhs);
    outer[i] = BOuterArgs{d_grids_ + i, &dh->lv[0], scenes_[i]->p, scenes_[i]->rhs, (float*)h.b[0],
                          from_b[i] ? nullptr : (float*)h.x[0], h.tau_pre[0], scr.ctl, scr.pdd, scr.pmx};
    const bool q = quad_ok && mg_quad_outer_ok(grids[i], h.lv[0], scenes_[i]->p, scenes_[i]->rhs);
    if (q)
      outer_q[i] = qblk(mg_quad_outer_blocks(grids[i]));
    else
      outer_s[i] = cb[i];
  }
  std::vector<int> outer_nb(n);
  for (size_t i = 0; i < n; ++i) outer_nb[i] = outer_q[i] + outer_s[i];
  auto outer_stage = [&](int it) {
    stage(Op::kOuterQ, outer.data(), sizeof(BOuterArgs), outer_q);
    stage(Op::kOuterS, outer.data(), sizeof(BOuterArgs), outer_s);
    fin_launch(Op::kFinOuter, fin_table(
                                  [&](size_t i, int nb) {
                                    Scratch& scr = scenes_[i]->solver->impl->scratch();
                                    // cycles take their iteration from the control block, so
                                    // every cycle's launches are the same (segments replay)
                                    return BFin<1, 2, FinMgTrue>{scr.pdd, scr.pmx, nb, 0,
                                                                 FinMgTrue{scr.ctl, it == 0 ? 0 : -1}};
                                  },
                                  1, 2, outer_nb));
  };

  // V-cycle stages, built once (same records every cycle)
  std::vector<Launch> cycle;
  {
    const size_t mark = program_.size();
    int maxL = 0;
    for (size_t i = 0; i < n; ++i) maxL = std::max(maxL, mg(i)->hierarchy().levels);
    // bottom level of each scene: its tail level, else its coarsest level
    std::vector<int> bottom(n);
    for (size_t i = 0; i < n; ++i) {
      const MgHierarchy& h = mg(i)->hierarchy();
      bottom[i] = h.tail_level > 0 ? h.tail_level : h.levels - 1;
    }
    // ping-pong state per scene and level: true = current iterate in x
    std::vector<std::vector<bool>> in_x(n, std::vector<bool>(kMgMaxLevels, true));
    // sweep `sw` of the pre (post = false) or post side of level l
    auto smooth_stage = [&](int l, bool to_p, bool post_side, int sw) {
      std::vector<BSmoothArgs> r(n);
      std::vector<int> bq(n, 0), bs(n, 0), bf(n, 0), bqf(n, 0);
      for (size_t i = 0; i < n; ++i) {
        if (l >= bottom[i]) continue;
        const MgHierarchy& h = mg(i)->hierarchy();
        const MgHierarchy* dh = mg(i)->device_hierarchy();
        const bool cx = in_x[i][l];
        const double* tau = l == 0 ? (post_side ? h.tau_post : h.tau_pre) : (post_side ? h.tau_cpost : h.tau_cpre);
        r[i] = BSmoothArgs{&dh->lv[l], (const float*)(cx ? h.x[l] : h.y[l]), (const float*)h.b[l],
                           (float*)(cx ? h.y[l] : h.x[l]), to_p ? scenes_[i]->p : nullptr, tau[sw],
                           scenes_[i]->solver->impl->scratch().ctl, h.tau_pre[0]};
        const bool q = quad_ok && mg_quad_size_ok(h.lv[l], l) && mg_quad_level(h.lv[l]);
        if (l == 0 && !post_side && sw == 1 && from_b[i])
          bf[i] = qblk(mg_quad_level_blocks(h.lv[l]));  // x0 recomputed from b0
        else if (q)
          bq[i] = qblk(mg_quad_level_blocks(h.lv[l]));
        else if (quad_ok && !to_p && mg_quad_size_ok(h.lv[l], l) && mg_quad_frac_level(h.lv[l]))
          bqf[i] = qblk(mg_quad_level_blocks(h.lv[l]));
        else
          bs[i] = lblk(h.lv[l]);
        if (!to_p) in_x[i][l] = !cx;
      }
      stage(to_p ? Op::kSmoothQUpdate : Op::kSmoothQ, r.data(), sizeof(BSmoothArgs), bq);
      stage(Op::kSmoothQFirst, r.data(), sizeof(BSmoothArgs), bf);
      stage(Op::kSmoothQFrac, r.data(), sizeof(BSmoothArgs), bqf);
      stage(Op::kSmoothS, r.data(), sizeof(BSmoothArgs), bs);  // the scalar body updates p itself
    };
    for (int l = 0; l < maxL; ++l) {
      for (int sw = 1; sw < (l == 0 ? pre : cpre); ++sw) smooth_stage(l, false, false, sw);
      // restriction l -> l+1
      std::vector<BRestrictArgs> r(n);
      std::vector<int> bq(n, 0), bs(n, 0), bqf(n, 0);
      for (size_t i = 0; i < n; ++i) {
        if (l >= bottom[i]) continue;
        const MgHierarchy& h = mg(i)->hierarchy();
        const MgHierarchy* dh = mg(i)->device_hierarchy();
        const bool first_in_y = mg_first_in_y(h, l + 1);
        r[i] = BRestrictArgs{&dh->lv[l], &dh->lv[l + 1], (const float*)(in_x[i][l] ? h.x[l] : h.y[l]),
                             (const float*)h.b[l], (float*)h.b[l + 1],
                             (float*)(first_in_y ? h.y[l + 1] : h.x[l + 1]), h.tau_cpre[0],
                             scenes_[i]->solver->impl->scratch().ctl};
        if (quad_ok && mg_quad_size_ok(h.lv[l], l) && mg_quad_pair(h.lv[l], h.lv[l + 1]))
          (h.lv[l].frac ? bqf : bq)[i] = qblk(mg_quad_restrict_blocks(h.lv[l + 1]));
        else
          bs[i] = lblk(h.lv[l + 1]);
        in_x[i][l + 1] = !first_in_y;
      }
      stage(Op::kRestrictQ, r.data(), sizeof(BRestrictArgs), bq);
      stage(Op::kRestrictQFrac, r.data(), sizeof(BRestrictArgs), bqf);
      stage(Op::kRestrictS, r.data(), sizeof(BRestrictArgs), bs);
      // bottoms at level l + 1
      std::vector<BTailArgs> t(n);
      std::vector<BCoarseArgs> cd(n);
      std::vector<int> bt(n, 0), bc(n, 0);
      int lds = 0;
      for (size_t i = 0; i < n; ++i) {
        if (bottom[i] != l + 1) continue;
        const MgHierarchy& h = mg(i)->hierarchy();
        const MgHierarchy* dh = mg(i)->device_hierarchy();
        const PcgControl* ctl = scenes_[i]->solver->impl->scratch().ctl;
        if (h.tail_level > 0) {
          t[i] = BTailArgs{dh, l + 1, ctl};
          bt[i] = 1;
          lds = std::max(lds, h.tail_lds);
        } else if (h.coarse_inv) {
          cd[i] = BCoarseArgs{&dh->lv[l + 1], h.coarse_inv, (const float*)h.b[l + 1], (float*)h.x[l + 1], ctl};
          bc[i] = 1;
        } else {
          set_last_error("a batched scene's coarsest level has no dense inverse", "rh_batch_step");
          return RH_UNSUPPORTED;
        }
      }
      stage(Op::kTail, t.data(), sizeof(BTailArgs), bt, lds);
      stage(Op::kCoarseDense, cd.data(), sizeof(BCoarseArgs), bc);
    }
    for (int l = maxL - 1; l >= 0; --l) {
      std::vector<BProlongArgs> r(n);
      std::vector<int> bq(n, 0), bs(n, 0);
      for (size_t i = 0; i < n; ++i) {
        if (l >= bottom[i]) continue;
        const MgHierarchy& h = mg(i)->hierarchy();
        const MgHierarchy* dh = mg(i)->device_hierarchy();
        const bool cx = in_x[i][l];
        r[i] = BProlongArgs{&dh->lv[l], &dh->lv[l + 1], (const float*)h.x[l + 1],
                            (const float*)(cx ? h.x[l] : h.y[l]), (float*)(cx ? h.y[l] : h.x[l]),
                            scenes_[i]->solver->impl->scratch().ctl};
        if (quad_ok && mg_quad_size_ok(h.lv[l], l) && mg_quad_prolong_pair(h.lv[l], h.lv[l + 1]))
          bq[i] = qblk(mg_quad_level_blocks(h.lv[l]));
        else
          bs[i] = lblk(h.lv[l]);
        in_x[i][l] = !cx;
      }
      stage(Op::kProlongQ, r.data(), sizeof(BProlongArgs), bq);
      stage(Op::kProlongS, r.data(), sizeof(BProlongArgs), bs);
      const int npost = l == 0 ? post : cpost;
      for (int sw = 0; sw < npost; ++sw) {
        const bool last = sw + 1 == npost;
        if (last && l == 0) {
          smooth_stage(0, true, true, sw);
        } else {
          smooth_stage(l, false, true, sw);
        }
      }
      // a coarse level's result must end in x[l] for the parent's prolongation
      for (size_t i = 0; i < n; ++i)
        if (l > 0 && l < bottom[i] && !in_x[i][l]) {
          set_last_error("sweep counts leave a level result outside x", "rh_batch_step");
          return RH_UNSUPPORTED;
        }
    }
    cycle.assign(program_.begin() + mark, program_.end());
    program_.resize(mark);
  }
  outer_stage(0);
  cycle_at_.assign(max_cycles_ + 1, 0);
  for (int it = 1; it <= max_cycles_; ++it) {
    cycle_at_[it] = program_.size();
    program_.insert(program_.end(), cycle.begin(), cycle.end());
    outer_stage(it);
  }
  post_at_ = program_.size();
  per_scene(Op::kEndSolve);
  stage(Op::kRegauge, cells.data(), sizeof(BCellArgs), cb);
  stage(Op::kReport, cells.data(), sizeof(BCellArgs), cb);
  fin_launch(Op::kFinReport, fin_table(
                                 [&](size_t i, int nb) {
                                   Scratch& scr = scenes_[i]->solver->impl->scratch();
                                   return BFin<1, 2, FReport>{scr.pdd, scr.pmx, nb, 0,
                                                              FReport{scr.ctl, FinStore{scr.scal, 2, 0.0, 0.0, 0.0}}};
                                 },
                                 1, 2, cb));
  per_scene(Op::kQuality);
  stage(Op::kCorrect, cor.data(), sizeof(BFaceArgs), fb);
  stage(Op::kDivMax, div.data(), sizeof(BFaceArgs), cb);
  fin_launch(Op::kFinDiv, fin_table(
                              [&](size_t i, int nb) {
                                Scratch& scr = scenes_[i]->solver->impl->scratch();
                                return BFin<1, 1, FReport>{scr.pdd, scr.pmx, nb, 0,
                                                           FReport{scr.ctl, FinStore{scr.scal + 3, 0, 0.0, 0.0, 0.0}}};
                              },
                              1, 1, cb));
  stage(Op::kRate, rate.data(), sizeof(BFaceArgs), rb);
  fin_launch(Op::kFinRate, fin_table(
                               [&](size_t i, int nb) {
                                 Scratch& scr = scenes_[i]->solver->impl->scratch();
                                 return BFin<1, 3, FReport>{
                                     scr.pdd, scr.pmx, nb, 0,
                                     FReport{scr.ctl, FinStore{scr.scal + 4, 1, grids[i].h[0], grids[i].h[1], grids[i].h[2]}}};
                               },
                               1, 3, rb));
  per_scene(Op::kGate);
  stage(Op::kCommit, commit.data(), sizeof(BFaceArgs), fb);

  // upload the tables
  if (d_tables_) (void)hipFree(d_tables_);
  d_tables_ = nullptr;
  tables_bytes_ = T.bytes.size();
  RH_HIP(hipMalloc(&d_tables_, std::max<size_t>(tables_bytes_, 16)));
  RH_HIP(hipMemcpy(d_tables_, T.bytes.data(), tables_bytes_, hipMemcpyHostToDevice));
  for (size_t i = 0; i < n; ++i) gens_[i] = scenes_[i]->solver->impl->generation();
  if (cfg_.use_graph == 2) {
    if (!cycle_state_) {
      RH_HIP(hipHostMalloc(&cycle_state_, 2 * sizeof(int), hipHostMallocMapped));
      RH_HIP(hipHostGetDevicePointer((void**)&cycle_state_dev_, cycle_state_, 0));
    }
    const size_t c1 = cycle_at_[1], c2 = max_cycles_ >= 2 ? cycle_at_[2] : post_at_;
    rh_status s = capture(0, c1, false, &seg_[0]);
    if (!s) s = capture(c1, c2, false, &seg_[1]);
    if (!s) s = capture(c1, c2, true, &seg_[2]);
    if (!s) s = capture(post_at_, program_.size(), false, &seg_[3]);
    if (s) return s;
    if (next_cycles_ < 1 || next_cycles_ > max_cycles_) next_cycles_ = std::max(1, max_cycles_ / 2);
  } else if (cfg_.use_graph) {
    RH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    run_program(stream_);
    hipError_t e = hipStreamEndCapture(stream_, &graph_);
    if (e != hipSuccess) {
      set_last_error(hipGetErrorString(e), "hipStreamEndCapture");
      return RH_DEVICE_ERROR;
    }
    RH_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
  built_ = true;
  return RH_OK;
}

void Batch::run_program(hipStream_t s) { run_range(0, program_.size(), s); }

void Batch::run_range(size_t begin, size_t end, hipStream_t s) {
  const int n = (int)scenes_.size();
  for (size_t li = begin; li < end; ++li) {
    const Launch& L = program_[li];
    BatchLaunch bl;
    bl.blocks = L.blocks;
    bl.n = n;
    bl.prefix = (const int*)(d_tables_ + L.prefix_off);
    bl.args = d_tables_ + L.args_off;
    const int pb = L.blocks;
    switch (L.op) {
      case Op::kBeginStep: kb_begin_step<<<pb, 256, 0, s>>>(d_scenes_, n); break;
      case Op::kBeginSolve: kb_begin_solve<<<pb, 256, 0, s>>>(d_scenes_, n); break;
      case Op::kEndSolve: kb_end_solve<<<pb, 256, 0, s>>>(d_scenes_, n); break;
      case Op::kQuality: kb_quality<<<pb, 256, 0, s>>>(d_scenes_, n); break;
      case Op::kGate: kb_gate<<<pb, 256, 0, s>>>(d_scenes_, n); break;
      case Op::kValidateAcc:
      case Op::kValidateCand: launch_b_validate(bl, s); break;
      case Op::kAdvect: launch_b_advect(bl, s, adv_off32_, false); break;
      case Op::kAdvectLattice: launch_b_advect(bl, s, adv_off32_, true); break;
      case Op::kRhs: launch_b_rhs(bl, s); break;
      case Op::kRhsSums: launch_b_rhs_sums(bl, s); break;
      case Op::kFinRhs:
        kb_finish<3, 1, FinNotDone<FinRhsSums>><<<n, dim3(kLanesX, kRowsPerBlock), 0, s>>>(
            (const BFin<3, 1, FinNotDone<FinRhsSums>>*)bl.args);
        break;
      case Op::kCompSums: launch_b_comp_sums(bl, s); break;
      case Op::kCompCheck: launch_b_comp_check(bl, s); break;
      case Op::kFill: launch_b_fill_pressure(bl, s); break;
      case Op::kOuterQ: launch_b_mg_outer_q(bl, s); break;
      case Op::kOuterS: launch_b_mg_outer(bl, s); break;
      case Op::kFinOuter:
        kb_finish<1, 2, FinMgTrue><<<n, dim3(kLanesX, kRowsPerBlock), 0, s>>>(
            (const BFin<1, 2, FinMgTrue>*)bl.args);
        break;
      case Op::kSmoothQ: launch_b_mg_smooth_q(bl, false, s); break;
      case Op::kSmoothQUpdate: launch_b_mg_smooth_q(bl, true, s); break;
      case Op::kSmoothQFirst: launch_b_mg_smooth_first_q(bl, s); break;
      case Op::kSmoothQFrac: launch_b_mg_smooth_qf(bl, s); break;
      case Op::kSmoothS: launch_b_mg_smooth(bl, s); break;
      case Op::kRestrictQ: launch_b_mg_restrict_q(bl, false, s); break;
      case Op::kRestrictQFrac: launch_b_mg_restrict_q(bl, true, s); break;
      case Op::kRestrictS: launch_b_mg_restrict(bl, s); break;
      case Op::kProlongQ: launch_b_mg_prolong_q(bl, s); break;
      case Op::kProlongS: launch_b_mg_prolong(bl, s); break;
      case Op::kTail: launch_b_mg_tail(bl, L.extra, s); break;
      case Op::kCoarseDense: launch_b_mg_coarsest_dense(bl, s); break;
      case Op::kRegauge: launch_b_regauge(bl, s); break;
      case Op::kReport: launch_b_residual_report(bl, s); break;
      case Op::kFinReport:
        kb_finish<1, 2, FinStatusOk<FinStore>><<<n, dim3(kLanesX, kRowsPerBlock), 0, s>>>(
            (const BFin<1, 2, FinStatusOk<FinStore>>*)bl.args);
        break;
      case Op::kCorrect: launch_b_correct(bl, s); break;
      case Op::kDivMax: launch_b_divergence_max(bl, s); break;
      case Op::kFinDiv:
        kb_finish<1, 1, FinStatusOk<FinStore>><<<n, dim3(kLanesX, kRowsPerBlock), 0, s>>>(
            (const BFin<1, 1, FinStatusOk<FinStore>>*)bl.args);
        break;
      case Op::kRate: launch_b_velocity_rate(bl, s); break;
      case Op::kFinRate:
        kb_finish<1, 3, FinStatusOk<FinStore>><<<n, dim3(kLanesX, kRowsPerBlock), 0, s>>>(
            (const BFin<1, 3, FinStatusOk<FinStore>>*)bl.args);
        break;
      case Op::kCommit:
        kb_commit<<<dim3(L.blocks, 3), dim3(kLanesX, kRowsPerBlock), 0, s>>>((const BFaceArgs*)bl.args,
                                                                            bl.prefix, n);
        break;
    }
  }
}

// Scenes whose solve is still running (control block not done).
__global__ void kb_count_active(const BatchScene* __restrict__ S, int n, int* out) {
  __shared__ int count;
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  int mine = 0;
  for (int s = threadIdx.x; s < n; s += blockDim.x) mine += S[s].ctl->done ? 0 : 1;
  if (mine) atomicAdd(&count, mine);
  __syncthreads();
  if (threadIdx.x == 0) *out = count;
}

constexpr int kPollCycles = 2;

// Scenes still solving and the largest outer iteration any scene reached,
// written straight into mapped pinned memory (the host reads it after the
// segment's event; no copy node in the graph).
__global__ void kb_cycle_state(const BatchScene* __restrict__ S, int n, int* out) {
  __shared__ int count, most;
  if (threadIdx.x == 0) {
    count = 0;
    most = 0;
  }
  __syncthreads();
  int mine = 0, hi = 0;
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    const PcgControl* c = S[s].ctl;
    mine += c->done ? 0 : 1;
    hi = max(hi, c->iteration);
  }
  if (mine) atomicAdd(&count, mine);
  if (hi) atomicMax(&most, hi);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = count;
    out[1] = most;
  }
}

// Launches [begin, end) of the program, then (with `state`) the cycle
// state, captured into an executable graph.
rh_status Batch::capture(size_t begin, size_t end, bool state, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  RH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  run_range(begin, end, stream_);
  if (state) kb_cycle_state<<<1, 256, 0, stream_>>>(d_scenes_, (int)scenes_.size(), cycle_state_dev_);
  hipError_t e = hipStreamEndCapture(stream_, &g);
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    set_last_error(hipGetErrorString(e), "Batch::capture");
    return RH_DEVICE_ERROR;
  }
  return RH_OK;
}

// One step as graph segments: the prologue, the cycles the last step
// needed (the last of them followed by the cycle state), then one more
// cycle at a time while some scene is still solving, then the post stage.
// Cycles after a scene's solve is done return at once for it, so the
// results equal the fixed-cycle program's; only the launches of cycles no
// scene needs are saved, at one host round trip per check.
rh_status Batch::run_segments(const std::function<rh_status()>* host_work) {
  if (!seg_ev_) RH_HIP(hipEventCreateWithFlags(&seg_ev_, hipEventDisableTiming));
  RH_HIP(hipGraphLaunch(seg_[0], stream_));
  int run = 0, want = std::min(std::max(next_cycles_, 1), max_cycles_);
  bool worked = false;
  volatile const int* st = cycle_state_;
  for (;;) {
    for (; run + 1 < want; ++run) RH_HIP(hipGraphLaunch(seg_[1], stream_));
    RH_HIP(hipGraphLaunch(seg_[2], stream_));
    ++run;
    RH_HIP(hipEventRecord(seg_ev_, stream_));
    if (host_work && *host_work && !worked) {
      worked = true;
      rh_status w = (*host_work)();
      if (w) return w;
    }
    RH_HIP(hipEventSynchronize(seg_ev_));
    if (st[0] == 0 || run >= max_cycles_) break;
    want = run + 1;
  }
  next_cycles_ = std::max(1, std::min((int)st[1], max_cycles_));
  RH_HIP(hipGraphLaunch(seg_[3], stream_));
  return RH_OK;
}

// One step without a graph: the solve's V-cycles are enqueued kPollCycles
// at a time; after each batch a one-block kernel counts the scenes still
// solving and the count is copied to pinned host memory.  The host waits
// for the previous batch's count (so the device always has one batch
// queued) and stops enqueuing cycles once it is 0.  Skipped cycles would
// only have launched kernels that return at once for every scene, so the
// results equal the full program's.
rh_status Batch::run_polled(const std::function<rh_status()>* host_work) {
  const int n = (int)scenes_.size();
  if (!d_active_) {
    RH_HIP(hipMalloc(&d_active_, 2 * sizeof(int)));
    RH_HIP(hipHostMalloc(&h_active_, 2 * sizeof(int), hipHostMallocDefault));
    RH_HIP(hipEventCreateWithFlags(&poll_ev_[0], hipEventDisableTiming));
    RH_HIP(hipEventCreateWithFlags(&poll_ev_[1], hipEventDisableTiming));
  }
  run_range(0, max_cycles_ >= 1 ? cycle_at_[1] : post_at_, stream_);
  int it = 1, pending = -1;
  bool worked = false;
  for (int b = 0; it <= max_cycles_; ++b) {
    const int last = std::min(max_cycles_, it + kPollCycles - 1);
    run_range(cycle_at_[it], last < max_cycles_ ? cycle_at_[last + 1] : post_at_, stream_);
    it = last + 1;
    const int cur = b & 1;
    kb_count_active<<<1, 256, 0, stream_>>>(d_scenes_, n, d_active_ + cur);
    RH_HIP(hipMemcpyAsync(h_active_ + cur, d_active_ + cur, sizeof(int), hipMemcpyDeviceToHost, stream_));
    RH_HIP(hipEventRecord(poll_ev_[cur], stream_));
    if (host_work && *host_work && !worked) {
      worked = true;
      rh_status w = (*host_work)();
      if (w) return w;
    }
    if (pending >= 0) {
      RH_HIP(hipEventSynchronize(poll_ev_[pending]));
      if (h_active_[pending] == 0) break;
    }
    pending = cur;
  }
  if (host_work && *host_work && !worked) {
    rh_status w = (*host_work)();
    if (w) return w;
  }
  run_range(post_at_, program_.size(), stream_);
  return RH_OK;
}

rh_status Batch::step(int steps, rh_step_report* reports, rh_status* statuses) {
  rh_status s = launch(steps);
  if (s) return s;
  return wait(reports, statuses);
}

rh_status Batch::launch(int steps, const std::function<rh_status()>& host_work) {
  const size_t n = scenes_.size();
  if (steps < 1) return RH_INVALID_ARGUMENT;
  launched_steps_ = 0;
  if (n == 0) return RH_OK;
  for (size_t i = 0; i < n; ++i)
    if (!(dt_[i] > 0.0)) return RH_INVALID_TIME_STEP;
  for (size_t i = 0; i < n; ++i)
    if (scenes_[i]->solver->impl->generation() != gens_[i]) built_ = false;
  last_build_seconds_ = 0.0;
  if (!built_) {
    const auto t0 = std::chrono::steady_clock::now();
    rh_status s = build();
    last_build_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (s) return s;
  }
  // every scene's own stream must be idle (uploads went through them)
  for (size_t i = 0; i < n; ++i) RH_HIP(hipStreamSynchronize(scenes_[i]->solver->impl->scratch().stream));
  // the batch commits faces without the one-scene step's bookkeeping
  for (size_t i = 0; i < n; ++i) scenes_[i]->rate_known = false;
  if (!ev_[0]) {
    RH_HIP(hipEventCreate(&ev_[0]));
    RH_HIP(hipEventCreate(&ev_[1]));
  }
  RH_HIP(hipEventRecord(ev_[0], stream_));
  for (int k = 0; k < steps; ++k) {
    if (cfg_.use_graph == 2) {
      rh_status r = run_segments(k == 0 ? &host_work : nullptr);
      if (r) return r;
    } else if (exec_) {
      RH_HIP(hipGraphLaunch(exec_, stream_));
    } else {
      rh_status r = run_polled(k == 0 ? &host_work : nullptr);
      if (r) return r;
    }
  }
  if (exec_ && host_work) {
    rh_status r = host_work();
    if (r) return r;
  }
  RH_HIP(hipEventRecord(ev_[1], stream_));
  RH_HIP(hipGetLastError());
  // the records come back through pinned memory: a pageable copy would queue
  // behind other threads' pageable transfers (a serving admission upload)
  if (pinned_cap_ < n) {
    if (pinned_) (void)hipHostFree(pinned_);
    pinned_ = nullptr;
    pinned_cap_ = 0;
    RH_HIP(hipHostMalloc(&pinned_, sizeof(BatchScene) * n, hipHostMallocDefault));
    pinned_cap_ = n;
  }
  RH_HIP(hipMemcpyAsync(pinned_, d_scenes_, sizeof(BatchScene) * n, hipMemcpyDeviceToHost,
                        stream_));
  launched_steps_ = steps;
  return RH_OK;
}

// Waits for the launched steps and unpacks the per-scene reports.
rh_status Batch::wait(rh_step_report* reports, rh_status* statuses) {
  const size_t n = scenes_.size();
  const int steps = launched_steps_;
  if (n == 0 || steps == 0) return RH_OK;
  launched_steps_ = 0;
  RH_HIP(hipStreamSynchronize(stream_));
  std::memcpy(h_scenes_.data(), pinned_, sizeof(BatchScene) * n);
  float ms = 0.0f;
  RH_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  last_seconds_ = ms * 1e-3;
  for (size_t i = 0; i < n; ++i) {
    const BatchScene& b = h_scenes_[i];
    if (statuses) statuses[i] = (rh_status)b.status;
    if (reports) {
      rh_step_report& r = reports[i];
      std::memset(&r, 0, sizeof(r));
      r.dt = b.dt;
      r.pressure = b.rep;
      r.actual_divergence_max = b.actual_div;
      r.courant = b.courant;
      r.seconds = ms * 1e-3 / steps;
    }
  }
  return RH_OK;
}

}  // namespace rh

struct rh_batch {
  rh::Batch impl;
};

extern "C" rh_status rh_batch_create(const rh_grid* grids, size_t n, const rh_batch_config* config,
                                     rh_batch** out) {
  if (!grids || !config || !out || n == 0) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  auto* b = new (std::nothrow) rh_batch();
  if (!b) return RH_DEVICE_ERROR;
  rh_status s = b->impl.create(grids, n, *config);
  if (s) {
    delete b;
    return s;
  }
  *out = b;
  return RH_OK;
}

extern "C" void rh_batch_destroy(rh_batch* batch) { delete batch; }

extern "C" uint64_t rh_batch_allocated_bytes(const rh_batch* batch) {
  return batch ? batch->impl.bytes() : 0;
}

extern "C" rh_status rh_batch_scene(rh_batch* batch, size_t i, rh_scene** out) {
  if (!batch || !out) return RH_INVALID_ARGUMENT;
  *out = batch->impl.scene(i);
  return *out ? RH_OK : RH_INVALID_ARGUMENT;
}

extern "C" rh_status rh_batch_set_dt(rh_batch* batch, const double* dt, size_t n) {
  if (!batch) return RH_INVALID_ARGUMENT;
  return batch->impl.set_dt_all(dt, n);
}

extern "C" rh_status rh_batch_set_variant(rh_batch* batch, uint32_t kernel_variant) {
  if (!batch) return RH_INVALID_ARGUMENT;
  return batch->impl.set_variant(kernel_variant);
}

extern "C" rh_status rh_batch_step(rh_batch* batch, int32_t steps, rh_step_report* reports,
                                   rh_status* statuses) {
  if (!batch) return RH_INVALID_ARGUMENT;
  return batch->impl.step(steps, reports, statuses);
}

extern "C" rh_status rh_batch_set_tuning(rh_batch* batch, const rh_batch_tuning* tuning) {
  if (!batch || !tuning) return RH_INVALID_ARGUMENT;
  return batch->impl.set_tuning(*tuning);
}

extern "C" rh_status rh_batch_autotune(rh_batch* batch, int32_t steps, rh_batch_tuning* table,
                                       size_t table_len, size_t* tried, rh_batch_tuning* chosen) {
  if (!batch) return RH_INVALID_ARGUMENT;
  return batch->impl.autotune(steps, table, table_len, tried, chosen);
}

// ================================================================ serving
//
// A serving loop over device-resident scenes with continuous batching:
// requests are admitted in arrival order while their planned device bytes
// fit the budget (and the active-scene cap); every tick runs one batched
// step of every running request; requests leave as they finish and admitted
// ones join at the next tick.  Admission (scene creation, obstacle set-up,
// the coarse levels' host preparation, uploads) runs on a worker thread of
// the server, so a tick's latency is its batched step and nothing else; a
// tick with no running request admits on its own thread (the queue never
// waits for the worker to wake).  The batched step of a scene does not
// depend on which other scenes share the launch, so a request's result is
// the same alone and in any mix (for a fixed launch shape).

namespace rh {

struct ServerRequest {
  int64_t id = 0;
  rh_grid grid{};
  std::vector<float> u, v, w;
  std::vector<uint8_t> solid;
  double dt = 0.0;
  int steps = 0;
  uint64_t plan = 0;
  rh_request_state st{};
  std::unique_ptr<rh_scene> scene;
  std::vector<float> ru, rv, rw;
  std::vector<double> rp;
};

class Server {
 public:
  ~Server();
  rh_status create(const rh_server_config& cfg);
  rh_status submit(const rh_request& r, int64_t* id);
  rh_status tick(rh_tick_report* rep);
  rh_status state(int64_t id, rh_request_state* out);
  rh_status result(int64_t id, float* u, float* v, float* w, double* p);

 private:
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
  }
  rh_status admit(ServerRequest* r);
  // Host -> device copy through the worker's pinned staging buffer.
  rh_status upload(void* dst, const void* src, size_t bytes, hipStream_t s);
  // Frees the request's scene (after copying its result when keep_result)
  // and returns its bytes to the budget.
  void retire(ServerRequest* r, bool keep_result);
  // The next queued request may be admitted now (caller holds mu_).
  bool admissible() const;
  // Admits the request at the queue head: reserve under mu_, build outside.
  void admit_front(std::unique_lock<std::mutex>& lk);
  void worker();
  rh_server_config cfg_{};
  Batch batch_;
  std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
  int device_ = 0;
  // guarded by mu_
  std::mutex mu_;
  std::condition_variable cv_;
  int64_t next_id_ = 1;
  uint64_t budget_ = 0, used_ = 0;
  std::vector<std::unique_ptr<ServerRequest>> all_;
  std::vector<ServerRequest*> queue_, ready_;
  // finished requests whose results the worker copies out before their
  // state turns final (final state, status, keep the result)
  struct Retiring {
    ServerRequest* r;
    int32_t state;
    rh_status status;
  };
  std::vector<Retiring> retiring_;
  int retiring_busy_ = 0;  // retirements the worker is carrying out
  size_t running_ = 0;    // requests in active_ (the tick thread's list)
  int admitting_ = 0;
  bool stop_ = false;
  // tick thread only
  std::vector<ServerRequest*> active_;
  std::thread thread_;
  // admission staging (pinned; used by one admission at a time)
  std::mutex stage_mu_;
  void* stage_ = nullptr;
  size_t stage_bytes_ = 0;
};

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  if (stage_) (void)hipHostFree(stage_);
}

rh_status Server::create(const rh_server_config& cfg) {
  cfg_ = cfg;
  rh_batch_config bc{};
  bc.scene = cfg.scene;
  bc.max_cycles = 0;
  bc.use_graph = 0;
  rh_status s = batch_.init_empty(bc);
  if (s) return s;
  rh_batch_tuning t{cfg.kernel_variant, cfg.block_divisor > 0 ? cfg.block_divisor : 1, 0.0};
  if ((s = batch_.set_tuning(t))) return s;
  budget_ = cfg.hbm_budget;
  if (!budget_) {
    size_t free_b = 0, total_b = 0;
    RH_HIP(hipMemGetInfo(&free_b, &total_b));
    budget_ = (uint64_t)(0.8 * (double)free_b);
  }
  RH_HIP(hipGetDevice(&device_));
  try {
    thread_ = std::thread([this] { worker(); });
  } catch (...) {
    set_last_error("could not start the admission thread", "rh_server_create");
    return RH_EXECUTION;
  }
  return RH_OK;
}

rh_status Server::submit(const rh_request& r, int64_t* id) {
  size_t cells = 0;
  rh_status s = validate_grid(&r.grid, &cells);
  if (s) return s;
  if ((s = check_indexable(&r.grid))) return s;
  if (!r.u || !r.v || !r.w || r.steps < 1 || !std::isfinite(r.dt) || r.dt <= 0.0) return RH_INVALID_ARGUMENT;
  auto q = std::make_unique<ServerRequest>();
  q->grid = r.grid;
  const size_t f[3] = {(size_t)(r.grid.counts[0] + 1) * r.grid.counts[1] * r.grid.counts[2],
                       (size_t)r.grid.counts[0] * (r.grid.counts[1] + 1) * r.grid.counts[2],
                       (size_t)r.grid.counts[0] * r.grid.counts[1] * (r.grid.counts[2] + 1)};
  q->u.assign(r.u, r.u + f[0]);
  q->v.assign(r.v, r.v + f[1]);
  q->w.assign(r.w, r.w + f[2]);
  if (r.solid) q->solid.assign(r.solid, r.solid + cells);
  q->dt = r.dt;
  q->steps = r.steps;
  uint64_t faces = 0;
  for (size_t a = 0; a < 3; ++a) faces += 2 * ((f[a] * 4 + 255) / 256 * 256);
  q->plan = faces + 2 * ((cells * 8 + 255) / 256 * 256) + MgSolver::plan_bytes(r.grid, cfg_.scene.solver);
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (q->plan > budget_) return RH_BUFFER_LIMIT;
    q->id = next_id_++;
    q->st.state = 0;
    q->st.status = RH_OK;
    q->st.submitted = now();
    q->st.bytes = q->plan;
    if (id) *id = q->id;
    queue_.push_back(q.get());
    all_.push_back(std::move(q));
  }
  cv_.notify_all();
  return RH_OK;
}

rh_status Server::upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> lk(stage_mu_);
  if (!stage_) {
    RH_HIP(hipHostMalloc(&stage_, 8u << 20, hipHostMallocDefault));
    stage_bytes_ = 8u << 20;
  }
  for (size_t off = 0; off < bytes; off += stage_bytes_) {
    const size_t n = std::min(stage_bytes_, bytes - off);
    std::memcpy(stage_, (const char*)src + off, n);
    RH_HIP(hipMemcpyAsync((char*)dst + off, stage_, n, hipMemcpyHostToDevice, s));
    RH_HIP(hipStreamSynchronize(s));
  }
  return RH_OK;
}

// Scene creation, obstacles and initial faces of a newly admitted request.
// The faces go straight into the scene's accepted arrays through the pinned
// staging buffer; the obstacle set-up that follows closes obstacle faces.
rh_status Server::admit(ServerRequest* r) {
  rh_scene* raw = nullptr;
  rh_status s = rh_scene_create(&r->grid, &cfg_.scene, &raw);
  if (s) return s;
  r->scene.reset(raw);
  const hipStream_t st = raw->solver->impl->scratch().stream;
  const std::vector<float>* src[3] = {&r->u, &r->v, &r->w};
  for (int a = 0; a < 3; ++a)
    if ((s = upload(raw->acc[a], src[a]->data(), raw->flen[a] * sizeof(float), st))) return s;
  raw->acc_checked = false;
  raw->rate_known = false;
  if (!r->solid.empty() && (s = rh_scene_set_solid(raw, r->solid.data(), r->solid.size()))) return s;
  MgSolver* mg = static_cast<MgSolver*>(raw->solver->impl.get());
  if (mg->levels() < 2) return RH_UNSUPPORTED;
  if ((s = mg->prepare(cfg_.scene.density))) return s;
  RH_HIP(hipStreamSynchronize(raw->solver->impl->scratch().stream));
  return RH_OK;
}

void Server::retire(ServerRequest* r, bool keep_result) {
  if (keep_result && r->scene) {
    const size_t cells = r->scene->n;
    r->ru.resize(r->scene->flen[0]);
    r->rv.resize(r->scene->flen[1]);
    r->rw.resize(r->scene->flen[2]);
    r->rp.resize(cells);
    (void)rh_scene_get_velocity(r->scene.get(), r->ru.data(), r->rv.data(), r->rw.data());
    (void)rh_scene_get_pressure(r->scene.get(), r->rp.data(), cells);
  }
  r->scene.reset();
  r->u.clear();
  r->v.clear();
  r->w.clear();
  r->solid.clear();
  {
    std::lock_guard<std::mutex> lk(mu_);
    used_ -= std::min(used_, r->st.bytes);
    r->st.finished = now();
  }
  cv_.notify_all();
}

bool Server::admissible() const {
  if (queue_.empty()) return false;
  const size_t running = running_ + ready_.size() + (size_t)admitting_;
  if (cfg_.max_active > 0 && running >= (size_t)cfg_.max_active) return false;
  return used_ + queue_.front()->plan <= budget_;
}

void Server::admit_front(std::unique_lock<std::mutex>& lk) {
  ServerRequest* r = queue_.front();
  queue_.erase(queue_.begin());
  used_ += r->plan;
  r->st.admitted = now();
  ++admitting_;
  lk.unlock();
  const rh_status a = admit(r);
  if (a) retire(r, false);
  lk.lock();
  --admitting_;
  if (a) {
    r->st.state = 3;
    r->st.status = a;
  } else {
    r->st.state = 1;
    ready_.push_back(r);
  }
}

void Server::worker() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lk(mu_);
  while (true) {
    cv_.wait(lk, [&] { return stop_ || !retiring_.empty() || admissible(); });
    if (stop_) return;
    if (!retiring_.empty()) {
      // retirements first: they return device memory to the budget
      const Retiring t = retiring_.front();
      retiring_.erase(retiring_.begin());
      ++retiring_busy_;
      lk.unlock();
      retire(t.r, true);
      lk.lock();
      --retiring_busy_;
      t.r->st.state = t.state;
      t.r->st.status = t.status;
    } else {
      admit_front(lk);
    }
    cv_.notify_all();
  }
}

rh_status Server::tick(rh_tick_report* rep) {
  rh_tick_report local{};
  rh_tick_report& T = rep ? *rep : local;
  std::memset(&T, 0, sizeof(T));
  const double t_start = now();
  std::vector<ServerRequest*> joined;
  {
    std::unique_lock<std::mutex> lk(mu_);
    if (active_.empty()) {
      // nothing runs: admit what fits here rather than wait for the worker
      const double a0 = now();
      // with nothing running, finish pending retirements (they free budget
      // and make results final) and admissions before going on
      while (!retiring_.empty() || retiring_busy_ > 0 || admitting_ > 0 || admissible()) {
        if (admissible() && retiring_.empty() && retiring_busy_ == 0)
          admit_front(lk);
        else
          cv_.wait(lk);
      }
      T.admit_seconds = now() - a0;
    }
    joined.swap(ready_);
    T.admitted = (int32_t)joined.size();
    running_ = active_.size() + joined.size();
    if (active_.empty()) {
      // the joined requests step from this tick on
      active_.swap(joined);
    }
  }
  // 1. one batched step of the current members
  std::vector<rh_scene*> members;
  std::vector<double> dts;
  for (ServerRequest* r : active_) {
    members.push_back(r->scene.get());
    dts.push_back(r->dt);
  }
  // Requests taken from ready_ join active_ also when the step fails (they
  // keep their scene and budget like any other running request).
  struct KeepJoined {
    std::vector<ServerRequest*>& joined;
    std::vector<ServerRequest*>& active;
    bool armed = true;
    ~KeepJoined() {
      if (armed) active.insert(active.end(), joined.begin(), joined.end());
    }
  } keep{joined, active_};
  rh_status s = batch_.set_members(members, dts);
  if (s) return s;
  if (!members.empty() && (s = batch_.launch(1))) return s;
  // 2. wait, account, retire finished requests
  std::vector<rh_step_report> reps(members.size());
  std::vector<rh_status> sts(members.size(), RH_OK);
  const double w0 = now();
  if (!members.empty() && (s = batch_.wait(reps.data(), sts.data()))) return s;
  T.wait_seconds = now() - w0;
  T.build_seconds = members.empty() ? 0.0 : batch_.last_build_seconds();
  keep.armed = false;
  std::vector<ServerRequest*> still;
  for (size_t i = 0; i < active_.size(); ++i) {
    ServerRequest* r = active_[i];
    {
      std::lock_guard<std::mutex> lk(mu_);
      r->st.last = reps[i];
    }
    T.cells += (int64_t)r->scene->n;
    // a finished request's results are copied out by the worker, off the
    // tick's path; its state turns final once they are on the host
    bool done;
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (sts[i] != RH_OK) {
        retiring_.push_back(Retiring{r, 3, sts[i]});
        done = true;
      } else {
        done = ++r->st.steps_done >= r->steps;
        if (done) retiring_.push_back(Retiring{r, 2, RH_OK});
      }
    }
    if (done) {
      ++T.finished;
      continue;
    }
    still.push_back(r);
  }
  for (ServerRequest* r : joined) still.push_back(r);
  active_.swap(still);
  {
    std::lock_guard<std::mutex> lk(mu_);
    running_ = active_.size();
    T.active = (int32_t)active_.size();
    T.queued = (int32_t)queue_.size();
    T.bytes_in_use = used_;
  }
  cv_.notify_all();
  T.stepped = (int32_t)members.size();
  T.device_seconds = members.empty() ? 0.0 : batch_.last_seconds();
  T.seconds = now() - t_start;
  return RH_OK;
}

rh_status Server::state(int64_t id, rh_request_state* out) {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& r : all_)
    if (r->id == id) {
      *out = r->st;
      return RH_OK;
    }
  return RH_INVALID_ARGUMENT;
}

rh_status Server::result(int64_t id, float* u, float* v, float* w, double* p) {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& r : all_)
    if (r->id == id) {
      if (r->st.state < 2 || r->ru.empty()) return RH_INVALID_ARGUMENT;
      if (u) std::memcpy(u, r->ru.data(), r->ru.size() * 4);
      if (v) std::memcpy(v, r->rv.data(), r->rv.size() * 4);
      if (w) std::memcpy(w, r->rw.data(), r->rw.size() * 4);
      if (p) std::memcpy(p, r->rp.data(), r->rp.size() * 8);
      return RH_OK;
    }
  return RH_INVALID_ARGUMENT;
}

}  // namespace rh

struct rh_server {
  rh::Server impl;
};

extern "C" rh_status rh_server_create(const rh_server_config* config, rh_server** out) {
  if (!config || !out) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  auto* s = new (std::nothrow) rh_server();
  if (!s) return RH_DEVICE_ERROR;
  rh_status st = s->impl.create(*config);
  if (st) {
    delete s;
    return st;
  }
  *out = s;
  return RH_OK;
}

extern "C" void rh_server_destroy(rh_server* server) { delete server; }

extern "C" rh_status rh_server_submit(rh_server* server, const rh_request* request, int64_t* id) {
  if (!server || !request) return RH_INVALID_ARGUMENT;
  return server->impl.submit(*request, id);
}

extern "C" rh_status rh_server_tick(rh_server* server, rh_tick_report* report) {
  if (!server) return RH_INVALID_ARGUMENT;
  return server->impl.tick(report);
}

extern "C" rh_status rh_server_state(rh_server* server, int64_t id, rh_request_state* out) {
  if (!server || !out) return RH_INVALID_ARGUMENT;
  return server->impl.state(id, out);
}

extern "C" rh_status rh_server_result(rh_server* server, int64_t id, float* u, float* v, float* w,
                                      double* pressure) {
  if (!server) return RH_INVALID_ARGUMENT;
  return server->impl.result(id, u, v, w, pressure);
}

extern "C" uint64_t rh_server_plan_bytes(const rh_grid* grid, const rh_scene_config* config) {
  if (!grid || !config || rh::validate_grid(grid, nullptr)) return 0;
  const uint64_t n = grid->counts[0] * grid->counts[1] * grid->counts[2];
  uint64_t faces = 0;
  for (int a = 0; a < 3; ++a) {
    uint64_t f = 1;
    for (int d = 0; d < 3; ++d) f *= grid->counts[d] + (d == a ? 1 : 0);
    faces += 2 * ((f * 4 + 255) / 256 * 256);
  }
  return faces + 2 * ((n * 8 + 255) / 256 * 256) + rh::MgSolver::plan_bytes(*grid, config->solver);
}
