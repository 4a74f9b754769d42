// Multigrid hierarchy shared by the host solver and the MG kernels.
#pragma once
#include "common.hpp"
#include "kernels.hpp"

namespace rh {

constexpr int kMgMaxLevels = 12;
constexpr int kMgCoarsestMax = 4096;  // cells solved by one workgroup in LDS

// One level's shape and rediscretised box-operator constants.
//
// Solid obstacles: level 0 reads the per-cell open-face mask of the fine
// operator (bits 0..5 = faces x-, x+, y-, y+, z-, z+ open, bit 6 = fluid).
// Levels >= 1 read `frac`: per cell the open fraction of its three lower
// faces (10 bits each: x in bits 0..9, y in 10..19, z in 20..29; fraction =
// value * fscale) and bit 30 = the cell holds fluid.  A coarse face's open
// fraction is the mean of its four sub-faces' (the coarse weight is half the
// sum of the sub-face weights, the area-weighted rediscretisation), so the
// coarse operator keeps a constant null space per fluid region.  Upper faces
// are the lower faces of the upper neighbours; wall faces are 0.
struct MgLevel {
  int nx, ny, nz, pad;
  double w[3];
  double diag[64];
  double odiag[64];  // 1 / diag (Jacobi scale); 0 for diag 0
  const unsigned char* mask = nullptr;  // level 0 with obstacles
  const unsigned* frac = nullptr;       // levels >= 1 with obstacles
  // levels >= 1 with obstacles: bit o (o = x | y << 1 | z << 2 parities of a
  // fine child) is set when the trilinear stencil of that child holds only
  // active coarse cells, so prolongation needs no renormalisation there
  const unsigned char* octant = nullptr;
  double wf[3] = {0.0, 0.0, 0.0};       // w[d] * fscale
  // Divisors of the four-cell chunk loops (set_level_divisors): nx / 4 and
  // nx / 2 chunks per row (sweeps; the restriction onto this level), ny.
  FastDiv div_q{}, div_h{}, div_y{};
};

inline void set_level_divisors(MgLevel& L) {
  L.div_q = make_fast_div(L.nx / 4);
  L.div_h = make_fast_div(L.nx / 2);
  L.div_y = make_fast_div(L.ny);
}

constexpr unsigned kFracFluid = 1u << 30;

constexpr int kMgMaxSweeps = 8;   // smoothing sweeps per side of the V-cycle

constexpr int kMgDenseMax = 512;  // coarsest levels up to this size use a dense inverse
// Dynamic LDS bytes for the tail levels' arrays (x, y, b, face fractions,
// octants, 1/D); the tail kernel's static LDS (8 KB) comes on top, within
// the 160 KB a gfx950 workgroup may hold.
constexpr int kMgTailLds = 120 * 1024;

// Smoothing sweep k of a side is x <- x + tau[k] (b - K x) / D.  Jacobi:
// every tau = omega.  Chebyshev (default): the taus are the reciprocals of
// the roots of the degree-nu Chebyshev polynomial on [lo, 2] (D^-1 K of a
// diagonally dominant operator has its spectrum in [0, 2]), so nu sweeps
// damp every eigencomponent in that interval by 1 / T_nu((2 + lo) /
// (2 - lo)); the sweeps commute, so their order is free.
// The finest level takes (pre, post) sweeps, every coarser level (cpre,
// cpost); a coarse level's first pre-sweep (x = tau_cpre[0] b / D) is fused
// into the restriction that produces its right-hand side.
struct MgHierarchy {
  int levels = 0;
  int pre = 2, post = 2;
  int cpre = 2, cpost = 2;
  double omega = 6.0 / 7.0;
  double tau_pre[kMgMaxSweeps] = {};
  double tau_post[kMgMaxSweeps] = {};
  double tau_cpre[kMgMaxSweeps] = {};
  double tau_cpost[kMgMaxSweeps] = {};
  // Coarsest level <= kMgDenseMax cells: G = M^-1 P with M = K + c 1_a 1_a^T
  // over the active (fluid) cells a and P the projection removing the mean
  // over a; column-major n x n f64 (rows and columns of inactive cells are
  // 0).  x = G b is the solution of K x = b - mean_a(b) with mean_a(x) = 0.
  const double* coarse_inv = nullptr;
  // Levels >= tail_level (> 0) run in one single-workgroup kernel that reads
  // this struct from device memory at d_self; 0 = no tail.
  bool quad = true;  // use the four-cells-per-lane kernels where a level allows
  int tail_level = 0;
  int tail_lds = 0;                  // dynamic LDS bytes of the tail launch
  // LDS byte offsets per tail level of x, y, b, fractions, octants and 1/D
  // (per cell on fraction levels, the 64-entry table on box levels); -1 =
  // not staged (1/D is then computed per use)
  int tail_off[kMgMaxLevels][6] = {};
  // per-cell 1/D in device memory of the tail fraction levels that stage it
  // (lv_odiag's values, computed once per hierarchy)
  const double* tail_odc[kMgMaxLevels] = {};
  const MgHierarchy* d_self = nullptr;
  MgLevel lv[kMgMaxLevels];
  void* x[kMgMaxLevels] = {};  // V-cycle iterate (S)
  void* y[kMgMaxLevels] = {};  // Jacobi ping-pong partner (S)
  void* b[kMgMaxLevels] = {};  // level right-hand side (S)
};

// A coarse level l (above the bottom: the tail level, else the coarsest)
// must end its V-cycle with its result in x[l] for the parent's
// prolongation.  Its iterate moves between x and y once per sweep after the
// first and once in the prolongation, (cpre - 1) + 1 + cpost moves in all;
// when that count is odd the restriction stores the first sweep in y[l].
inline bool mg_first_in_y(const MgHierarchy& h, int l) {
  const int bottom = h.tail_level > 0 ? h.tail_level : h.levels - 1;
  return l >= 1 && l < bottom && ((h.cpre + h.cpost) & 1) != 0;
}

// Four-cells-per-lane FP32 kernels (levels with nx % 4 == 0 and no face
// fractions); identical values to the scalar kernels.
bool mg_quad_level(const MgLevel& L);
bool mg_quad_pair(const MgLevel& F, const MgLevel& C);  // restriction F -> C (exact halving, nx % 4 == 0)
bool mg_quad_prolong_pair(const MgLevel& F, const MgLevel& C);  // prolongation C -> F (F may have fractions)
bool mg_quad_frac_level(const MgLevel& L);  // smoothing sweeps of fraction levels, four cells per lane
// Four cells per lane pays on the finest level and on coarse levels of at
// least 2^20 cells; smaller coarse levels are latency-bound and run faster
// with one cell per thread (more waves in flight).
constexpr long long kMgQuadMinCells = 1ll << 20;
inline bool mg_quad_size_ok(const MgLevel& L, int l) {
  return l == 0 || (long long)L.nx * L.ny * L.nz >= kMgQuadMinCells;
}
void launch_mg_smooth_qf(const MgLevel& L, const float* xin, const float* b, float* xout, double tau,
                         const PcgControl* ctl, hipStream_t s);
bool mg_quad_outer_ok(const DevGrid& g, const MgLevel& L0, const double* p, const double* b);
// Block counts of the four-cells-per-lane launches (one-scene grids; a
// batched stage uses the same counts per scene).
int mg_quad_outer_blocks(const DevGrid& g);
inline int mg_quad_level_blocks(const MgLevel& L) { return quad_blocks(L.nx / 4, L.ny * L.nz); }
inline int mg_quad_restrict_blocks(const MgLevel& C) { return quad_blocks(C.nx / 2, C.ny * C.nz); }
// p_zero: p is known to be 0 and is not read (the solve's first residual,
// the first cycle's update p = 0 + x).
void launch_mg_outer_residual_q(const DevGrid& g, const MgLevel& L0, const double* p, const double* b,
                                float* b0, float* x0, double tau, const PcgControl* ctl, DD* pdd,
                                double* pmx, hipStream_t s, bool p_zero = false);
void launch_mg_smooth_q(const MgLevel& L, const float* xin, const float* b, float* xout, double* p,
                        double tau, const PcgControl* ctl, hipStream_t s, bool p_zero = false);
// The finest level's last sweep (p += x) takes the four-cell kernel.
bool mg_quad_update(const MgHierarchy& h, const double* p);
// launch_residual_report (sum of squares, max, nonfinite flag of b - K p)
// four cells per lane; needs mg_quad_outer_ok.
void launch_residual_report_q(const DevGrid& g, const double* p, const double* b, const Reduce& red,
                              hipStream_t s);
// Second finest pre-sweep from b0 alone (x0 = tau1 b0 / D recomputed per
// position; the outer residual then does not store x0).
struct MgFold;
void launch_mg_smooth_first_q(const MgLevel& L, const float* b, float* xout, double tau1, double tau2,
                              const PcgControl* ctl, hipStream_t s, const MgFold* fold = nullptr);
struct MgHierarchy;
// Fused legs of the small coarse levels (mg_fused_kernels.hip): FP32 levels
// 1 <= l < the tail level (or the coarsest) below the four-cell size, with
// up to kMgLegMaxSweeps coarse sweeps per side.  The down leg reads b[l],
// stores the pre-smoothed x in y[l] and restricts onto b[l + 1]; the up
// leg adds the prolonged x[l + 1] to y[l], post-smooths, and stores x[l].
constexpr int kMgLegMaxSweeps = 3;
struct MgLegTaus {
  double t[kMgLegMaxSweeps];
};
bool mg_leg_fusable(const MgHierarchy& h, int l);
void launch_mg_down_leg(const MgHierarchy& h, int l, const PcgControl* ctl, hipStream_t s);
void launch_mg_up_leg(const MgHierarchy& h, int l, const PcgControl* ctl, hipStream_t s);
bool mg_first_from_b(const DevGrid& g, const MgHierarchy& h, const double* p, const double* b);
void launch_mg_restrict_q(const MgLevel& F, const MgLevel& Cl, const float* x, const float* b, float* bc,
                          float* xc, double tau, const PcgControl* ctl, hipStream_t s);
void launch_mg_prolong_q(const MgLevel& F, const MgLevel& Cl, const float* xc, const float* xin,
                         float* xout, const PcgControl* ctl, hipStream_t s);

// Acceptance of the outer defect-correction residual; every cycle is a
// "candidate" (the true residual is computed each cycle).
struct FinMgTrue {
  PcgControl* ctl;
  int it;  // outer iteration; < 0: the one after ctl->iteration (a captured cycle replayed any number of times)
  // Optional mapped pinned word the host paces its enqueueing on:
  // ((iteration + 1) << 1) | done once this iteration is decided.
  int* progress = nullptr;
  __device__ bool active() const { return !ctl->done; }
  __device__ void operator()(const double* sum, const double* mx) const {
    decide(sum, mx);
    if (progress)
      __hip_atomic_store(progress, ((ctl->iteration + 1) << 1) | (ctl->done ? 1 : 0), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __device__ void decide(const double* sum, const double* mx) const {
    const int it = this->it >= 0 ? this->it : ctl->iteration + 1;
    if (mx[1] != 0.0) return fail(ctl, RH_ARITHMETIC_FAILURE, it);
    const double squared = sum[0], maximum = mx[0];
    const double divergence = maximum * ctl->dt / ctl->volume;
    if (!isfinite(squared) || !isfinite(divergence)) return fail(ctl, RH_BREAKDOWN, it);
    const double l2 = sqrt(squared);
    ctl->true_l2 = l2;
    ctl->true_max = maximum;
    ctl->pred_div = divergence;
    ctl->iteration = it;
    if (l2 <= ctl->threshold && divergence <= ctl->div_limit) {
      ctl->done = 1;
      ctl->status = RH_OK;
      return;
    }
    if (it >= ctl->max_iterations) {
      ctl->error_value = maximum;
      return fail(ctl, RH_ITERATION_LIMIT, it);
    }
  }
};

// The previous outer residual's acceptance (FinMgTrue over its partials),
// run by block 0 of the next cycle's first kernel instead of as a launch of
// its own.  Kernels that run in the cycle before the decision lands do
// scratch work only; the cycle's pressure update comes later and sees it.
struct MgFold {
  const DD* pdd;
  const double* pmx;
  int nb;
  FinMgTrue fin;
};
// Blocks of the outer residual's launch (the acceptance finisher's partial count).
template <typename S>
int mg_outer_blocks(const DevGrid& g, const MgHierarchy& h, const double* p, const double* b);

// x0_implicit (FP32 only, mg_first_from_b): the outer residual does not
// store the first pre-sweep and the cycle starts from b0 alone.
// p_zero (four-cell FP32 path only): p holds 0 and is not read (see
// launch_mg_outer_residual_q); the caller fills p when no cycle runs.
// finish: launch the acceptance finisher after it (false: the next cycle's
// first sweep carries it, MgFold); progress: FinMgTrue::progress.
template <typename S>
void mg_outer_residual(const DevGrid& g, const MgHierarchy& h, const double* p, const double* b,
                       PcgControl* ctl, DD* pdd, double* pmx, int it, hipStream_t s, bool x0_implicit,
                       bool p_zero = false, bool finish = true, int* progress = nullptr);
// fold: the previous outer residual's acceptance, run by the cycle's first
// sweep (four-cell FP32 path with x0_implicit only).
template <typename S>
void mg_cycle(const MgHierarchy& h, double* p, const PcgControl* ctl, int it, hipStream_t s, bool x0_implicit,
              bool p_zero = false, const MgFold* fold = nullptr);
// p[gauge] subtracted from the solve's pressure when the solve ended OK or
// at its iteration limit (solid cells 0); p = 0 when no cycle ran and
// p_unwritten; decided on the device, so no host round trip precedes it.
void mg_regauge(const DevGrid& g, double* p, PcgControl* ctl, hipStream_t s, bool p_unwritten);
int mg_tail_lds_limit();  // dynamic LDS bytes the tail kernels may use

}  // namespace rh
