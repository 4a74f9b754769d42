// Live per-kernel timing with HIP events.  Disabled by default (one branch
// per launch).  A Scope arms an event pair that the next rh::launch on the
// same thread attaches to the kernel's own dispatch (hipExtLaunchKernelGGL:
// the events take that dispatch's start and end timestamps, so no marker
// packets sit between kernels and the figure is the kernel's duration).  A
// timed run can still focus on one slot.  Entries carry the solver
// iteration they belong to, so launches enqueued past convergence (which
// return at once) can be discarded before they are counted.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

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
  kMgSmoothFirst,
  kDense3dCount,
  kDense3dWrite,
  kNumSlots
};

bool enabled();
// Timing on for this slot (profiling on and no focus, or focus == slot).
bool enabled_for(int slot);
// Arm an event pair for `slot` (`tag` = solver iteration or -1) that the
// next launch on this thread takes; disarm drops it if no launch took it.
void arm(int slot, int tag);
void disarm();
// The armed pair, if any, marked as taken by a launch.
bool take(hipEvent_t* a, hipEvent_t* b);
// Resolve recorded pairs (synchronises on their events): pairs with
// tag <= keep_max_tag (or tag < 0) are accumulated, others dropped.
void settle(int keep_max_tag);
const char* name(int slot);

struct Scope {
  bool on;
  Scope(int sl, hipStream_t, int tag) : on(enabled_for(sl)) {
    if (on) arm(sl, tag);
  }
  ~Scope() {
    if (on) disarm();
  }
};

}  // namespace prof

// Kernel launch that carries the armed event pair of the enclosing Scope.
template <typename... KArgs, typename... Args>
inline void launch(void (*kernel)(KArgs...), dim3 grid, dim3 block, uint32_t shmem, hipStream_t s,
                   Args... args) {
  hipEvent_t a, b;
  if (prof::enabled() && prof::take(&a, &b))
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0, static_cast<KArgs>(args)...);
  else
    kernel<<<grid, block, shmem, s>>>(static_cast<KArgs>(args)...);
}

}  // namespace rh
