// Kernel launch interface shared by the host code of rheon_amd.
#pragma once
#include <vector>

#include "common.hpp"

namespace rh {

// Error bits collected by kernels with atomicOr into a device flag word.
enum : unsigned {
  kFlagNonfiniteInput = 1u,
  kFlagNonzeroWall = 2u,
  kFlagArithmetic = 4u,
};

// Grid-reduction scratch: per-block partials, an arrival counter and the
// output scalars of one reduction.
struct Reduce {
  DD* pdd;            // 4 * kMaxBlocks
  double* pmx;        // 4 * kMaxBlocks
  double* out;        // result scalars
  unsigned* flags;    // error bits
};

// ------------------------------------------------------------ field kernels
template <typename T>
void launch_advect(const DevGrid& g, const void* u, const void* v, const void* w, double dt,
                   void* ou, void* ov, void* ow, unsigned* flags, hipStream_t s);
template <typename T>
void launch_validate(const DevGrid& g, const void* u, const void* v, const void* w,
                     int check_walls, unsigned* flags, hipStream_t s);
// validate: also launch_validate's checks with the wall test, in the same pass.
// sums (implies validate): also the partials of launch_rhs_sums's kernel,
// published to sums->pdd / sums->pmx for launch_rhs_sums_finish.
template <typename T>
void launch_rhs(const DevGrid& g, const void* u, const void* v, const void* w, double dt,
                double* out, unsigned* flags, hipStream_t s, bool validate = false,
                const Reduce* sums = nullptr);
template <typename T>
void launch_correct(const DevGrid& g, const double* p, double dt, void* u, void* v, void* w,
                    unsigned* flags, hipStream_t s);
template <typename T>
void launch_divergence_max(const DevGrid& g, const void* u, const void* v, const void* w,
                           const Reduce& red, hipStream_t s);
template <typename T>
void launch_velocity_rate(const DevGrid& g, const void* u, const void* v, const void* w,
                          const Reduce& red, hipStream_t s);
// Both in one pass: divergence maximum into red.out[0], the velocity rate
// (and the three component maxima) into red.out[1..4].
template <typename T>
void launch_divergence_rate(const DevGrid& g, const void* u, const void* v, const void* w,
                            const Reduce& red, hipStream_t s);
// launch_correct then launch_divergence_rate in one pass, out of place: the
// corrected faces of in[] are written to out[] (every face: walls copied),
// flags into red.flags, the sums as launch_divergence_rate's, bit for bit.
template <typename T>
void launch_correct_divergence_rate(const DevGrid& g, const double* p, double dt, void* const in[3], void* out[3],
                                    const Reduce& red, hipStream_t s);
template <typename T>
void launch_zero_blocked(const DevGrid& g, void* u, void* v, void* w, hipStream_t s);
// Cells [lo, hi) of a tracer advection; the rest are written +0.0.
struct TracerBox {
  int lo[3], hi[3];
};
// box: nullptr = every cell.
void launch_advect_tracer(const DevGrid& g, const float* old, const float* u, const float* v,
                          const float* w, double dt, float* out, unsigned* flags, hipStream_t s,
                          const TracerBox* box = nullptr);
// bounds[0..5]: min i, j, k and one past max i, j, k of the nonzero cells
// (n, n, n, 0, 0, 0 when none is nonzero).
void launch_tracer_bounds(const DevGrid& g, const float* t, unsigned* bounds, unsigned* flags, hipStream_t s);
// red.out[0] = sum of the tracer values; arithmetic flag on a value outside [0, 1].
void launch_tracer_sum(const DevGrid& g, const float* t, const Reduce& red, hipStream_t s);

// -------------------------------------------------------------- PCG kernels
// Device-side solver control block: all scalar decisions of solve_loaded
// (src/pressure.rs:577-740) are taken on the device, so the host only
// enqueues iterations and polls `done` every few iterations.
struct PcgControl {
  double rz[2];        // rz of iteration it is rz[it & 1]
  double alpha;
  double threshold;
  double rel, abs_tol, div_limit;
  double dt, volume;
  double rr;           // last recursive r.r
  double true_l2, true_max, pred_div;
  double sum, sum_abs; // compatibility sums
  double error_value;
  int max_iterations;
  int iteration;       // last completed iteration
  int done;            // nonzero once a terminal status is set
  int status;          // rh_status
  int error_iteration;
  int candidate;       // this iteration needs a true report
  int replaced;        // this iteration replaced r by the true residual
  int seeded;          // started from a caller-supplied pressure
  double sum_r;        // sum of r (multigrid gauge shift)
  double spare[7];
};

__device__ __forceinline__ void fail(PcgControl* ctl, int status, int it) {
  ctl->done = 1;
  ctl->status = status;
  ctl->error_iteration = it;
}

// Compatibility sums and threshold (src/pressure.rs:603-633).
struct FinRhsSums {
  PcgControl* ctl;
  __device__ bool active() const { return true; }
  __device__ void operator()(const double* sum, const double* mx) const {
    if (mx[0] != 0.0) return fail(ctl, RH_NONFINITE_INPUT, 0);
    const double total = sum[0], total_abs = sum[1];
    ctl->sum = total;
    ctl->sum_abs = total_abs;
    const double budget = (64.0 * 2.220446049250313e-16) * total_abs;
    if (!isfinite(total) || !isfinite(total_abs)) return fail(ctl, RH_BREAKDOWN, 0);
    if (fabs(total) > budget) {
      ctl->error_value = total;
      return fail(ctl, RH_INCOMPATIBLE_RHS, 0);
    }
    const double rhs_l2 = sqrt(sum[2]);
    const double threshold = fmax(ctl->abs_tol, ctl->rel * rhs_l2);
    if (!isfinite(rhs_l2) || !isfinite(threshold)) return fail(ctl, RH_BREAKDOWN, 0);
    ctl->threshold = threshold;
  }
};

// One tile of the lexicographic SGS sweeps (tile coordinates).
struct SgsTile {
  int bx, by, bz, pad;
};
constexpr int kSgsTileEdge = 8;

struct PcgArrays {
  const double* seed = nullptr;  // caller-supplied starting pressure (none: p = 0)
  void* p;       // f64 solution (= p64)
  void* r;       // T
  void* d;       // T
  void* q;       // T (also holds the true residual)
  void* z;       // T (SGS: the preconditioned residual)
  const double* b;
  double* p64;   // f64 output pressure when T = float (may alias p when T = double)
  // lexicographic SGS: the tiles of wave w are tiles[wave_start[w] ..
  // wave_start[w + 1]) (device array; host offsets)
  const SgsTile* tiles = nullptr;
  const int* wave_start = nullptr;
  int waves = 0;
};

template <typename T>
void pcg_prologue(const DevGrid& g, const PcgArrays& a, PcgControl* ctl, DD* pdd, double* pmx,
                  hipStream_t s);
template <typename T>
void pcg_iteration(const DevGrid& g, int method, const PcgArrays& a, PcgControl* ctl, DD* pdd,
                   double* pmx, int it, hipStream_t s);
template <typename T>
void pcg_export_pressure(const DevGrid& g, const PcgArrays& a, hipStream_t s);

void launch_apply_full(const DevGrid& g, const double* x, double* y, unsigned* flags,
                       hipStream_t s);
// tiles/wave_start/waves: the lexicographic SGS plan (RH_SGS_PCG only).
void launch_precondition(const DevGrid& g, int method, double* r, double* z, PcgControl* ctl,
                         DD* pdd, double* pmx, hipStream_t s, const SgsTile* tiles = nullptr,
                         const int* wave_start = nullptr, int waves = 0);
// The tile-wave plan of the lexicographic SGS sweeps: tiles of kSgsTileEdge
// cells per axis, grouped by wave bx + by + bz (host).
void sgs_plan(const DevGrid& g, std::vector<SgsTile>& tiles, std::vector<int>& wave_start);
void launch_residual_report(const DevGrid& g, const double* p, const double* b, const Reduce& red,
                            hipStream_t s);
void launch_comp_check(const DevGrid& g, const double* b, PcgControl* ctl, DD* pdd, double* pmx, hipStream_t s);
// p = seed with the seed's checks (non-finite, nonzero gauge) into ctl.
void launch_seed(const DevGrid& g, const double* seed, double* p, PcgControl* ctl, DD* pdd, double* pmx,
                 hipStream_t s);
// launch_rhs_sums's finisher alone, on partials already published (by
// launch_rhs with sums) over the same grid.
void launch_rhs_sums_finish(const DevGrid& g, const double* b, PcgControl* ctl, DD* pdd, double* pmx,
                            hipStream_t s);
void launch_rhs_sums(const DevGrid& g, const double* b, PcgControl* ctl, DD* pdd, double* pmx,
                     hipStream_t s);

}  // namespace rh
