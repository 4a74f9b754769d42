// Live per-kernel timing with HIP events recorded on the launching stream.
// Disabled by default (one branch per launch).  Every event pair costs the
// stream a few microseconds, so a timed run can focus on one slot: only
// that slot's launches are bracketed.  Entries carry the solver
// iteration they belong to, so launches enqueued past convergence (which
// return at once) can be discarded before they are counted.
#pragma once
#include <hip/hip_runtime.h>

namespace rh {
namespace prof {

enum Slot {
  kApplyDir = 0,
  kUpdate,
  kTrueResidual,
  kReplace,
  kDirection,
  kSgsBlack,
  kSgsRed,
  kAdvect,
  kRhs,
  kCorrect,
  kDivMax,
  kMgSmooth,
  kMgResidual,
  kMgRestrict,
  kMgProlong,
  kMgCoarse,
  kOuterResidual,
  kOuterUpdate,
  kBatch,
  kMgSmoothUpdate,
  kNumSlots
};

bool enabled();
// Timing on for this slot (profiling on and no focus, or focus == slot).
bool enabled_for(int slot);
// Record a start event for `slot` on stream s; `tag` = solver iteration or -1.
void start(int slot, hipStream_t s, int tag);
void stop(int slot, hipStream_t s);
// Resolve recorded pairs (synchronises on their events): pairs with
// tag <= keep_max_tag (or tag < 0) are accumulated, others dropped.
void settle(int keep_max_tag);
const char* name(int slot);

struct Scope {
  int slot;
  hipStream_t s;
  bool on;
  Scope(int sl, hipStream_t st, int tag) : slot(sl), s(st), on(enabled_for(sl)) {
    if (on) start(sl, st, tag);
  }
  ~Scope() {
    if (on) stop(slot, s);
  }
};

}  // namespace prof
}  // namespace rh
