// Event-based kernel timer (see profiler.hpp).
#include "profiler.hpp"

#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace rh {
namespace prof {

namespace {

struct Pair {
  hipEvent_t a = nullptr, b = nullptr;
  int slot = -1;
  int tag = -1;
  bool taken = false;  // attached to a dispatch
};

struct State {
  std::mutex mu;
  bool on = false;
  int focus = -1;            // slot timed alone, or -1 = every slot
  std::vector<Pair> pool;    // reusable event pairs
  size_t used = 0;
  int64_t launches[kNumSlots] = {};
  double seconds[kNumSlots] = {};
};

State& st() {
  static State s;
  return s;
}

thread_local int armed_pair = -1;  // pool index armed on this thread

const char* kNames[kNumSlots] = {
    "pcg_apply_dir", "pcg_update",      "pcg_true_residual", "pcg_replace",   "pcg_direction",
    "pcg_sgs_black", "pcg_sgs_red",     "advect_velocity",   "build_rhs",     "correct",
    "divergence_max", "mg_smooth",      "mg_residual",       "mg_restrict",   "mg_prolong",
    "mg_coarse_levels", "mg_outer_residual", "outer_update",   "batch_step",    "mg_smooth_update",
    "mg_smooth_first", "dense3d_count", "dense3d_write"};

}  // namespace

bool enabled() { return st().on; }

bool enabled_for(int slot) {
  const State& S = st();
  return S.on && (S.focus < 0 || S.focus == slot);
}

const char* name(int slot) { return slot >= 0 && slot < kNumSlots ? kNames[slot] : "?"; }

void arm(int slot, int tag) {
  State& S = st();
  std::lock_guard<std::mutex> lk(S.mu);
  armed_pair = -1;
  if (S.used == S.pool.size()) {
    Pair p;
    if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) return;
    S.pool.push_back(p);
  }
  Pair& p = S.pool[S.used];
  p.slot = slot;
  p.tag = tag;
  p.taken = false;
  armed_pair = (int)S.used;
  ++S.used;
}

void disarm() { armed_pair = -1; }

bool take(hipEvent_t* a, hipEvent_t* b) {
  if (armed_pair < 0) return false;
  State& S = st();
  std::lock_guard<std::mutex> lk(S.mu);
  Pair& p = S.pool[armed_pair];
  armed_pair = -1;
  p.taken = true;
  *a = p.a;
  *b = p.b;
  return true;
}

void settle(int keep_max_tag) {
  State& S = st();
  std::lock_guard<std::mutex> lk(S.mu);
  for (size_t i = 0; i < S.used; ++i) {
    Pair& p = S.pool[i];
    if (!p.taken || (p.tag >= 0 && p.tag > keep_max_tag)) continue;
    if (hipEventSynchronize(p.b) != hipSuccess) continue;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess) continue;
    S.launches[p.slot] += 1;
    S.seconds[p.slot] += ms * 1e-3;
  }
  S.used = 0;
}

void enable(bool on) {
  State& S = st();
  std::lock_guard<std::mutex> lk(S.mu);
  S.on = on;
  S.used = 0;
  std::memset(S.launches, 0, sizeof(S.launches));
  std::memset(S.seconds, 0, sizeof(S.seconds));
}

bool set_focus(const char* kernel) {
  State& S = st();
  std::lock_guard<std::mutex> lk(S.mu);
  if (!kernel) {
    S.focus = -1;
    return true;
  }
  for (int i = 0; i < kNumSlots; ++i)
    if (std::strcmp(kernel, kNames[i]) == 0) {
      S.focus = i;
      return true;
    }
  return false;
}

bool query(const char* kernel, int64_t* launches, double* seconds) {
  State& S = st();
  std::lock_guard<std::mutex> lk(S.mu);
  for (int i = 0; i < kNumSlots; ++i)
    if (std::strcmp(kernel, kNames[i]) == 0) {
      *launches = S.launches[i];
      *seconds = S.seconds[i];
      return true;
    }
  return false;
}

}  // namespace prof
}  // namespace rh

extern "C" rh_status rh_profile_enable(int32_t on) {
  rh::prof::enable(on != 0);
  return RH_OK;
}

extern "C" rh_status rh_profile_query(const char* kernel, int64_t* launches, double* seconds) {
  if (!kernel || !launches || !seconds) return RH_INVALID_ARGUMENT;
  return rh::prof::query(kernel, launches, seconds) ? RH_OK : RH_INVALID_ARGUMENT;
}

extern "C" rh_status rh_profile_focus(const char* kernel) {
  return rh::prof::set_focus(kernel) ? RH_OK : RH_INVALID_ARGUMENT;
}
