// Host-side objects of rheon_amd.
#pragma once
#include <functional>
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "mg.hpp"

namespace rh {

const char* last_error();
rh_status check_indexable(const rh_grid* g);
rh_status check_density(const rh_grid& g, double density);
rh_status flags_to_status(unsigned f);
// Fluid regions of an obstacle mask (src/static_obstacle.rs:165-210): label
// per cell (kNoComponent for solid cells) and each region's gauge cell.
constexpr uint8_t kNoComponent = 255;
constexpr int kMaxComponents = 254;
struct ObstacleComponents {
  std::vector<int> gauges;       // lowest-index cell of each region, in label order
  std::vector<uint8_t> labels;   // per cell; filled only with more than one region
};
rh_status build_obstacle_mask(const rh_grid& g, const uint8_t* solid, std::vector<uint8_t>& mask,
                              int* gauge, ObstacleComponents* comps);

// Per-object device scratch: a stream, reduction partials, counters, error
// flags, result scalars, the solver control block and a pinned readback page.
struct Scratch {
  hipStream_t stream = nullptr;
  void* base = nullptr;
  DD* pdd = nullptr;
  double* pmx = nullptr;
  unsigned* counters = nullptr;
  unsigned* flags = nullptr;  // two words; clear_flags clears both
  double* scal = nullptr;
  PcgControl* ctl = nullptr;
  void* host = nullptr;
  // One mapped pinned word the device writes a solve's progress into
  // (FinMgTrue::progress): host view and device pointer.
  int* progress = nullptr;
  int* progress_dev = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  size_t bytes_ = 0;

  // device bytes of one Scratch (reduction partials for kMaxBlocks blocks)
  static constexpr size_t kDeviceBytes = sizeof(DD) * 4 * kMaxBlocks + sizeof(double) * 4 * kMaxBlocks +
                                         sizeof(unsigned) * 64 + sizeof(double) * 64 +
                                         2 * sizeof(PcgControl) + 1024;
  rh_status init();
  void release();
  Reduce reduce(int out_slot);
  rh_status clear_flags();
  rh_status read_flags(unsigned* out);
  rh_status read_scalars(double* out, int n, int slot);
  // Both flag words (flags[0], flags[1]) and result slots 0..n-1 in one copy.
  rh_status read_flags_scalars(unsigned f[2], double* out, int n);
  // The flag words and result slots slot..slot+n-1 as the last multigrid
  // read-back with a tail brought them back (no device access).
  void last_readback(unsigned f[2], double* out, int slot, int n) const;
};

Scratch& global_scratch(std::unique_lock<std::mutex>& lock);

// A device pressure solver behind PressureWorkspace::solve_rhs
// (src/pressure.rs:365-393): the caller's RHS and pressure are device arrays.
class Solver {
 public:
  virtual ~Solver() { release_mask(); }
  virtual rh_status create(const rh_grid& grid, const rh_solver_config& config) = 0;
  virtual rh_status solve(double density, const double* d_rhs, double dt,
                          const rh_pressure_settings& st, double* d_pressure,
                          rh_pressure_report* report, rh_cancel_fn cancel, void* ctx) = 0;
  virtual uint64_t allocated_bytes() const = 0;
  Scratch& scratch() { return scratch_; }
  // The RHS at d_rhs was written by launch_rhs with sums into this
  // solver's scratch partials; the next multigrid solve of that RHS runs
  // only the sums' finisher (consumed by that solve either way).
  void rhs_sums_published(const double* d_rhs) { rhs_sums_for_ = d_rhs; }
  // Work the next multigrid solve enqueues on its stream after its own
  // kernels and before its read-back, which then also brings back the flag
  // words and every result slot (Scratch::last_readback); dropped unrun when
  // the solve returns before that point.
  void set_solve_tail(std::function<void()> f) { tail_ = std::move(f); }
  const rh_grid& grid() const { return grid_; }
  // Solid obstacles: `solid` holds one flag per cell (nonzero = solid) on
  // the host, or is null for the plain box.  Builds the open-face mask,
  // labels the fluid regions (one gauge each) and keeps the mask on the device.
  virtual rh_status set_solid(const uint8_t* solid);
  const unsigned char* device_mask() const { return d_mask_; }
  int gauge() const { return gauge_; }
  int components() const { return ncomp_; }
  // The next solve starts from this device pressure (null: from p = 0).
  void set_seed(const double* seed) { seed_ = seed; }
  // Operator constants with this solver's obstacle mask and gauge.
  DevGrid dev_grid(double density) const;
  // Changes whenever device-side operator data is replaced (obstacles,
  // hierarchy uploads), so batched launch tables know to rebuild.
  uint64_t generation() const { return generation_; }
  const rh_solver_config& config() const { return config_; }

 protected:
  rh_status check_and_begin(double density, const double* d_rhs, double dt,
                            const rh_pressure_settings& st, double* d_pressure,
                            rh_pressure_report* report, rh_cancel_fn cancel, void* ctx);
  // Enqueues `step(it)` for it = 1.. in batches of `poll`, reading the
  // control block back one batch behind; returns when the device reports a
  // terminal status, the iteration budget is spent or `cancel` fires.
  template <typename Step>
  rh_status drive(int max_iterations, int poll, Step step, bool* cancelled, rh_cancel_fn cancel,
                  void* ctx);
  // Enqueues `step(it)` for it = 1.. paced on the device's progress word
  // (scratch_.progress, written by each iteration's acceptance finisher):
  // at least `ahead` iterations, then one beyond the iteration the device
  // is running, so the device always has the next one queued and at most
  // one iteration past convergence is enqueued (its launches return at
  // once).  No copies or events per iteration; cancellation is polled
  // every `poll` decided iterations.
  template <typename Step>
  rh_status drive_paced(int max_iterations, int poll, int ahead, Step step, bool* cancelled,
                        rh_cancel_fn cancel, void* ctx, int* enqueued);
  // Reads the control block back (and, with scal3, result slots 0..2 in
  // the same copy) and fills the report.
  rh_status finish_report(rh_pressure_report* report, bool cancelled, PcgControl* out, double* scal3 = nullptr,
                          bool ev1_recorded = false);
  void release_mask();
  void release_components();
  rh_status upload_components(const ObstacleComponents& comps);
  rh_grid grid_{};
  rh_solver_config config_{};
  Scratch scratch_;
  const double* rhs_sums_for_ = nullptr;
  std::function<void()> tail_;
  double dt_ = 0.0;
  std::vector<uint8_t> hmask_;      // host copy of the obstacle mask (empty = box)
  unsigned char* d_mask_ = nullptr;  // device obstacle mask
  int gauge_ = 0;
  // Several fluid regions: one device allocation holding the region labels,
  // gauge cells, captured gauge pressures, the region-sorted cell list and
  // its chunk table (DevGrid's comp_* fields).
  void* d_comp_ = nullptr;
  int ncomp_ = 1, comp_nchunks_ = 0;
  const unsigned char* comp_label_ = nullptr;
  const int* comp_gauge_ = nullptr;
  double* comp_shift_ = nullptr;
  const int* comp_cells_ = nullptr;
  const CompChunk* comp_chunks_ = nullptr;
  const int* comp_chunk_first_ = nullptr;
  std::vector<int> gauges_;  // host copy of the gauge cells
  const double* seed_ = nullptr;
  uint64_t generation_ = 0;
};

Solver* make_solver(rh_method method);

// jacobi-pcg-v1 and RB-SGS PCG (src/pressure.rs:577-740).
class PcgSolver : public Solver {
 public:
  rh_status create(const rh_grid& grid, const rh_solver_config& config) override;
  ~PcgSolver() override;
  rh_status solve(double density, const double* d_rhs, double dt, const rh_pressure_settings& st,
                  double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel,
                  void* ctx) override;
  uint64_t allocated_bytes() const override;
  rh_status set_solid(const uint8_t* solid) override;

 private:
  template <typename T>
  rh_status run(const DevGrid& g, const double* d_rhs, const rh_pressure_settings& st,
                double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel, void* ctx);
  void* arena_ = nullptr;
  SgsTile* d_tiles_ = nullptr;    // lexicographic SGS tile waves (device)
  std::vector<int> wave_start_;   // their offsets per wave (host)
  void *r_ = nullptr, *d_ = nullptr, *q_ = nullptr, *z_ = nullptr;
  size_t n_ = 0;
  uint64_t bytes_ = 0;
};

// Multigrid V-cycles inside an f64 defect-correction loop.
class MgSolver : public Solver {
 public:
  rh_status create(const rh_grid& grid, const rh_solver_config& config) override;
  ~MgSolver() override;
  rh_status solve(double density, const double* d_rhs, double dt, const rh_pressure_settings& st,
                  double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel,
                  void* ctx) override;
  uint64_t allocated_bytes() const override;
  int levels() const { return h_.levels; }
  static int plan_levels(const rh_grid& grid, const rh_solver_config& config, int dims[][3], int* padded);
  static uint64_t plan_bytes(const rh_grid& grid, const rh_solver_config& config);
  // Sweeps per side of the finest level (pre, post) and of the coarse
  // levels (cpre, cpost).
  static rh_status plan_sweeps(const rh_solver_config& config, int* pre, int* post, int* cpre, int* cpost);
  rh_status set_solid(const uint8_t* solid) override;
  // Level data and coarsest inverse for `density` on the device (done by
  // every solve; batched runners call it before building launch tables).
  rh_status prepare(double density) { return set_density(density); }
  const MgHierarchy& hierarchy() const { return h_; }
  const MgHierarchy* device_hierarchy() const { return d_h_; }

 private:
  template <typename S>
  rh_status run(const DevGrid& g, const double* d_rhs, const rh_pressure_settings& st,
                double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel, void* ctx);
  rh_status set_density(double density);
  rh_status build_fractions();
  MgHierarchy h_;
  int padded_level_ = 0;  // first level made by padded (ceil) coarsening; 0 = none
  unsigned* d_frac_ = nullptr;                // coarse-level face fractions (device)
  std::vector<std::vector<unsigned>> hfrac_;  // host copies per level (index = level)
  double fscale_[kMgMaxLevels] = {};
  bool dirty_ = true;                         // hierarchy must be re-uploaded
  void* arena_ = nullptr;
  double* d_inv_ = nullptr;   // dense coarsest inverse (device)
  size_t inv_cap_ = 0;        // doubles allocated at d_inv_
  double* d_odc_ = nullptr;   // per-cell 1/D of the tail's fraction levels (device)
  size_t odc_cap_ = 0;        // doubles allocated at d_odc_
  MgHierarchy* d_h_ = nullptr;  // device copy of h_ read by the tail kernel
  double inv_density_ = 0.0;  // density the inverse was built for
  int last_cycles_ = 0;       // cycles the previous solve took (enqueued without waiting next time)
  uint64_t bytes_ = 0;
};

}  // namespace rh
