// Opaque C handles of the device solver and the device-resident scene
// (shared by the C ABI and the batched runner).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>

#include "runtime.hpp"

struct rh_solver {
  std::unique_ptr<rh::Solver> impl;
};

struct rh_scene {
  rh_grid grid{};
  rh_scene_config config{};
  size_t n = 0, flen[3] = {0, 0, 0};
  void* acc[3] = {nullptr, nullptr, nullptr};
  void* cand[3] = {nullptr, nullptr, nullptr};
  double* rhs = nullptr;
  double* p = nullptr;
  void* arena = nullptr;
  uint64_t bytes = 0;
  std::unique_ptr<rh_solver> solver;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // The accepted faces are known finite: they were committed by a step
  // (whose correction checked every face it wrote) and nothing has written
  // them since.  `external`: their device pointers were handed out, so any
  // step validates them again.
  bool acc_checked = false;
  bool external = false;
  // Passive tracer (accepted, candidate), f32 per cell; null without one.
  float* tracer[2] = {nullptr, nullptr};
  // Simulation time and committed steps (src/simulation.rs StepReport).
  double time = 0.0;
  uint64_t generation = 0;
  // Velocity rate of the accepted faces, kept from the Courant gate of the
  // step that committed them (valid while nothing else wrote them).
  double acc_rate = 0.0;
  bool rate_known = false;
  // Stepped by a batch (rh_batch / rh_server), which carries no tracer.
  bool batched = false;
  ~rh_scene() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (arena) (void)hipFree(arena);
    if (tracer[0]) (void)hipFree(tracer[0]);
  }
};

