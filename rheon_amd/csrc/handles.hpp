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
  ~rh_scene() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (arena) (void)hipFree(arena);
  }
};

