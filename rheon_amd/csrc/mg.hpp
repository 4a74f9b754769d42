// Multigrid hierarchy shared by the host solver and the MG kernels.
#pragma once
#include "common.hpp"
#include "kernels.hpp"

namespace rh {

constexpr int kMgMaxLevels = 12;
constexpr int kMgCoarsestMax = 4096;  // cells solved by one workgroup in LDS

// One level's shape and rediscretised box-operator constants.
struct MgLevel {
  int nx, ny, nz, pad;
  double w[3];
  double diag[64];
  double odiag[64];  // omega / diag (weighted-Jacobi scale)
};

constexpr int kMgDenseMax = 512;  // coarsest levels up to this size use a dense inverse
constexpr int kMgTailLds = 56 * 1024;  // LDS bytes for the tail levels' x, y, b arrays

struct MgHierarchy {
  int levels = 0;
  int pre = 2, post = 2;
  double omega = 6.0 / 7.0;
  // Coarsest level <= kMgDenseMax cells: (K + c 11^T)^-1, column-major n x n
  // f64 on the device; x = M^-1 (b - mean b) is the mean-free solution.
  const double* coarse_inv = nullptr;
  // Levels >= tail_level (> 0) run in one single-workgroup kernel that reads
  // this struct from device memory at d_self; 0 = no tail.
  int tail_level = 0;
  int tail_lds = 0;                  // dynamic LDS bytes of the tail launch
  int tail_off[kMgMaxLevels][3] = {};  // LDS byte offsets of x, y, b per tail level
  const MgHierarchy* d_self = nullptr;
  MgLevel lv[kMgMaxLevels];
  void* x[kMgMaxLevels] = {};  // V-cycle iterate (S)
  void* y[kMgMaxLevels] = {};  // Jacobi ping-pong partner (S)
  void* b[kMgMaxLevels] = {};  // level right-hand side (S)
};

template <typename S>
void mg_outer_residual(const DevGrid& g, const MgHierarchy& h, const double* p, const double* b,
                       PcgControl* ctl, DD* pdd, double* pmx, int it, hipStream_t s);
template <typename S>
void mg_cycle(const MgHierarchy& h, double* p, const PcgControl* ctl, int it, hipStream_t s);
void mg_regauge(const DevGrid& g, double* p, PcgControl* ctl, hipStream_t s);

}  // namespace rh
