// Opaque C handles of the device solver and the device-resident scene
// (shared by the C ABI and the batched runner).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

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
  // Pressure initialization of the solve (rh_scene_set_initialization) and
  // where the accepted pressure of the last committed step is: 0 none, 1 in
  // p (the last solve's output), 2 in p_acc (kept aside by a warm step, so a
  // failed attempt does not lose it).  p_acc is allocated by the first warm
  // step.  A failed attempt keeps state 1 unless its solve wrote p
  // (p_written); a cold scene whose attempt fails after its solve has no
  // accepted pressure left to seed a later warm step with.
  bool p_written = false;
  int init_policy = RH_INIT_COLD;
  int p_state = 0;
  double* p_acc = nullptr;
  // Faces the one-scene step corrects into (out of place), then swaps with
  // cand, so the three buffers of each component circulate among acc, cand
  // and spare; spare_base (allocated by the first such step) is freed with
  // the arena.
  void* spare[3] = {nullptr, nullptr, nullptr};
  void* spare_base = nullptr;
  ~rh_scene() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (arena) (void)hipFree(arena);
    if (tracer[0]) (void)hipFree(tracer[0]);
    if (p_acc) (void)hipFree(p_acc);
    if (spare_base) (void)hipFree(spare_base);
  }
};

// impl PressureProvider for the GPU solver (src/solver_provider.rs:67-76),
// or, with `supplied` set, the supplied-array pressure witness of the
// projection qualification (SuppliedPressure, src/solver_provider/
// composition/projection.rs:554-593): predict() returns the supplied values.
struct rh_provider {
  rh_provider_config config{};
  std::unique_ptr<rh_solver> solver;
  rh_grid grid{};
  double* d_rhs = nullptr;
  double* d_p = nullptr;
  size_t cap = 0;
  const double* supplied = nullptr;
  size_t supplied_len = 0;
  ~rh_provider() {
    if (d_rhs) (void)hipFree(d_rhs);
    if (d_p) (void)hipFree(d_p);
  }
};

// ProviderWorkspace (src/solver_provider.rs:94-327) on the device.
struct rh_provider_workspace {
  rh_grid grid{};
  rh_method fallback = RH_JACOBI_PCG;
  std::unique_ptr<rh_solver> solver;  // native fallback on the device
  double* d_rhs = nullptr;
  double* d_p = nullptr;             // the last selected pressure (device)
  std::vector<double> pressure;      // the last selected pressure (host)
  ~rh_provider_workspace() {
    if (d_rhs) (void)hipFree(d_rhs);
    if (d_p) (void)hipFree(d_p);
  }
};

// SolverMetadata id of the supplied-array pressure witness.
constexpr const char* kSuppliedPressureId = "supplied-projection-pressure-witness-v1";
