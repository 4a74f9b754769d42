// Host-side objects of rheon_amd.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace rh {

const char* last_error();
rh_status check_indexable(const rh_grid* g);
rh_status check_density(const rh_grid& g, double density);
rh_status flags_to_status(unsigned f);

// Per-object device scratch: a stream, reduction partials, counters, error
// flags, result scalars, the PCG control block and a pinned readback page.
struct Scratch {
  hipStream_t stream = nullptr;
  void* base = nullptr;
  DD* pdd = nullptr;
  double* pmx = nullptr;
  unsigned* counters = nullptr;
  unsigned* flags = nullptr;
  double* scal = nullptr;
  PcgControl* ctl = nullptr;
  void* host = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  size_t bytes_ = 0;

  rh_status init();
  void release();
  Reduce reduce(int out_slot);
  rh_status clear_flags();
  rh_status read_flags(unsigned* out);
  rh_status read_scalars(double* out, int n, int slot);
};

Scratch& global_scratch(std::unique_lock<std::mutex>& lock);

// PressureWorkspace on the device (src/pressure.rs:243-740).
class PcgSolver {
 public:
  rh_status create(const rh_grid& grid, const rh_solver_config& config);
  ~PcgSolver();
  rh_status solve(double density, const double* d_rhs, double dt, const rh_pressure_settings& st,
                  double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel, void* ctx);
  uint64_t allocated_bytes() const;
  Scratch& scratch() { return scratch_; }
  const rh_grid& grid() const { return grid_; }

 private:
  template <typename T>
  rh_status run(const DevGrid& g, const double* d_rhs, const rh_pressure_settings& st,
                double* d_pressure, rh_pressure_report* report, rh_cancel_fn cancel, void* ctx);
  rh_grid grid_{};
  rh_solver_config config_{};
  Scratch scratch_;
  void* arena_ = nullptr;
  void *r_ = nullptr, *d_ = nullptr, *q_ = nullptr, *z_ = nullptr, *p32_ = nullptr;
  size_t n_ = 0;
  uint64_t bytes_ = 0;
  double dt_ = 0.0;
};

}  // namespace rh
