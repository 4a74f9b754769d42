// Batched launches: one kernel launch runs one stage of the projection step
// for many scenes.  The grid is the concatenation of every participating
// scene's own block decomposition (prefix[s] .. prefix[s + 1] - 1); each
// block finds its scene, reads that scene's argument record from device
// memory and runs the same kernel body as a one-scene launch with its
// scene-local block index.  Argument records and prefixes are built once
// per batch, so the whole step can be captured in a graph and replayed.
#pragma once
#include "common.hpp"
#include "kernels.hpp"
#include "mg.hpp"

namespace rh {

// A batched launch: total blocks, scene count, device prefix and records.
struct BatchLaunch {
  int blocks = 0;
  int n = 0;
  const int* prefix = nullptr;
  const void* args = nullptr;
};

// --------------------------------------------------------- field kernels
struct BFaceArgs {  // validate / advect / rhs / correct / divergence / rate
  const DevGrid* g;
  const float* u;
  const float* v;
  const float* w;
  float* ou;
  float* ov;
  float* ow;
  const double* dt;  // the scene's step dt (device)
  double* cells;     // rhs out (rhs) or pressure in (correct)
  unsigned* flags;
  DD* pdd;
  double* pmx;
  const PcgControl* ctl;  // correct: skipped unless the solve succeeded
  int mode;               // validate: check walls
};

void launch_b_validate(const BatchLaunch& bl, hipStream_t s);
// lattice: every scene of the launch has a lattice grid (DevGrid::lattice)
void launch_b_advect(const BatchLaunch& bl, hipStream_t s, bool off32, bool lattice);
void launch_b_rhs(const BatchLaunch& bl, hipStream_t s);
void launch_b_correct(const BatchLaunch& bl, hipStream_t s);
void launch_b_divergence_max(const BatchLaunch& bl, hipStream_t s);
void launch_b_velocity_rate(const BatchLaunch& bl, hipStream_t s);

// --------------------------------------------------------- cell kernels
struct BCellArgs {  // rhs sums / residual report / regauge / fill
  const DevGrid* g;
  const double* b;  // rhs
  double* p;        // pressure
  DD* pdd;
  double* pmx;
  PcgControl* ctl;
};
void launch_b_rhs_sums(const BatchLaunch& bl, hipStream_t s);
void launch_b_residual_report(const BatchLaunch& bl, hipStream_t s);
void launch_b_fill_pressure(const BatchLaunch& bl, hipStream_t s);
void launch_b_regauge(const BatchLaunch& bl, hipStream_t s);
void launch_b_comp_sums(const BatchLaunch& bl, hipStream_t s);
void launch_b_comp_check(const BatchLaunch& bl, hipStream_t s);  // bl.n blocks, one per scene

// --------------------------------------------------------- multigrid
struct BOuterArgs {
  const DevGrid* g;
  const MgLevel* L0;
  const double* p;
  const double* b;
  float* b0;
  float* x0;
  double tau;  // first pre-smoothing sweep weight
  const PcgControl* ctl;
  DD* pdd;
  double* pmx;
};
struct BSmoothArgs {
  const MgLevel* L;
  const float* xin;
  const float* b;
  float* xout;
  double* p;  // non-null: p += x (finest level, last sweep)
  double tau;  // sweep weight
  const PcgControl* ctl;
  double tau_first;  // first-sweep weight (the sweep from b0 alone)
};
struct BRestrictArgs {
  const MgLevel* F;
  const MgLevel* C;
  const float* x;
  const float* b;
  float* bc;
  float* xc;
  double tau;  // first coarse pre-smoothing sweep weight
  const PcgControl* ctl;
};
struct BProlongArgs {
  const MgLevel* F;
  const MgLevel* C;
  const float* xc;
  const float* xin;
  float* xout;
  const PcgControl* ctl;
};
struct BTailArgs {  // one block per scene
  const MgHierarchy* H;
  int l0;
  const PcgControl* ctl;
};
struct BCoarseArgs {  // one block per scene: dense coarsest solve
  const MgLevel* L;
  const double* minv;
  const float* b;
  float* x;
  const PcgControl* ctl;
};

void launch_b_mg_outer_q(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_smooth_q(const BatchLaunch& bl, bool update, hipStream_t s);
void launch_b_mg_smooth_first_q(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_smooth_qf(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_restrict_q(const BatchLaunch& bl, bool frac, hipStream_t s);
void launch_b_mg_prolong_q(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_outer(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_smooth(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_restrict(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_prolong(const BatchLaunch& bl, hipStream_t s);
void launch_b_mg_tail(const BatchLaunch& bl, int lds_bytes, hipStream_t s);
void launch_b_mg_coarsest_dense(const BatchLaunch& bl, hipStream_t s);

}  // namespace rh
