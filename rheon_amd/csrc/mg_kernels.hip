// Geometric multigrid for the fixed-box pressure Poisson problem.
//
// Cell-centred hierarchy: level l+1 halves every axis of level l (all axes
// even), spacing doubles, and the operator is the box operator rediscretised
// on the coarse spacing (weights A_d/(rho h_d), the integrated operator of
// src/operator.rs:58-97 on that grid).  One V(nu1, nu2) cycle:
//   pre-smooth  x = tau_1 b/D (first sweep from zero), then nu1-1 sweeps
//               x <- x + tau_k (b - K x)/D (Chebyshev or weighted-Jacobi
//               weights tau_k, see MgHierarchy)
//   restrict    b_c = sum over the 8 children of the fine residual
//               (integrated residuals add), x_c = tau_1 b_c/D_c
//   recurse / coarsest solve
//   prolong     x += trilinear interpolation of x_c (cell-centred weights
//               27/64, 9/64, 3/64, 1/64; Neumann reflection at walls)
//   post-smooth nu2 sweeps; at the finest level the last sweep also adds
//               the correction to the f64 pressure.
// The coarsest level (<= 4096 cells) is solved by one workgroup with f64 CG
// on the mean-free right-hand side.  The outer loop (FP64 defect correction)
// computes the true residual b - K p every cycle with the reference's
// acceptance rule (src/pressure.rs:432-456, :686-715).
// Storage type S of the V-cycle is float (mixed precision) or double.
#include "common.hpp"
#include "kernels.hpp"
#include "mg.hpp"
#include "profiler.hpp"
#include "stencil.hpp"
#include "mg_stencil.hpp"
#include "mg_quad.hpp"
#include "batch.hpp"

namespace rh {


// ------------------------------------------------------------------ kernels

// Outer FP64 residual r = b - K p (every cell), written as the finest V-cycle
// right-hand side b0 = S(r) with the first pre-smoothing sweep x0 = tau
// b0/D fused in; partial sum r^2, max|r|, nonfinite flag.
template <typename S>
__device__ __forceinline__ void b_mg_outer_residual(VBlock vb, const DevGrid& g, const MgLevel& L0,
                                                    const double* __restrict__ p,
                                                    const double* __restrict__ b, S* __restrict__ b0,
                                                    S* __restrict__ x0, double tau, const PcgControl* ctl,
                                                    DD* pdd, double* pmx) {
  if (ctl->done) return;
  DD s[1] = {{0.0, 0.0}};
  double m[2] = {0.0, 0.0};
  FOR_CELLS_VB(g, vb)
  const double v = apply_at<double>(g, p, c, i, j, k);
  if (!isfinite(v)) m[1] = 1.0;
  const double r = b[c] - v;
  dd_acc(s[0], r * r);
  m[0] = fmax(m[0], fabs(r));
  const S rs = (S)r;
  b0[c] = rs;
  x0[c] = lv_scale<S>(L0, tau, c, i, j, k) * rs;
  END_CELLS
  block_publish<1, 2>(s, m, pdd, pmx, vb);
}
template <typename S>
__global__ void __launch_bounds__(kBlock)
k_mg_outer_residual(DevGrid g, MgLevel L0, const double* __restrict__ p, const double* __restrict__ b,
                    S* __restrict__ b0, S* __restrict__ x0, double tau, const PcgControl* ctl,
                    DD* pdd, double* pmx) {
  b_mg_outer_residual<S>(hw_block(), g, L0, p, b, b0, x0, tau, ctl, pdd, pmx);
}

// One smoothing sweep x_out = x_in + tau (b - K x_in)/D.  With
// `p` non-null (finest level, last post-sweep) also p += x_out.
template <typename S>
__device__ __forceinline__ void b_mg_smooth(VBlock vb, const MgLevel& L, const S* __restrict__ xin,
                                            const S* __restrict__ b, S* __restrict__ xout,
                                            double* __restrict__ p, double tau, const PcgControl* ctl) {
  if (ctl->done) return;
  FOR_LEVEL_CELLS_VB(L, vb)
  const S ax = s_apply<S>(L, xin, c, i, j, k);
  const S xn = xin[c] + lv_scale<S>(L, tau, c, i, j, k) * (b[c] - ax);
  if (p) {
    p[c] = p[c] + (double)xn;
  } else {
    xout[c] = xn;
  }
  END_CELLS
}
template <typename S>
__global__ void __launch_bounds__(kBlock)
k_mg_smooth(MgLevel L, const S* __restrict__ xin, const S* __restrict__ b, S* __restrict__ xout,
            double tau, double* __restrict__ p, const PcgControl* ctl) {
  b_mg_smooth<S>(hw_block(), L, xin, b, xout, p, tau, ctl);
}

// Restriction: coarse cell (I,J,K) sums the residual b - K x of its 8
// children; writes b_c and the first coarse pre-smoothing sweep x_c.
template <typename S>
__device__ __forceinline__ void b_mg_restrict(VBlock vb, const MgLevel& F, const MgLevel& Cl,
                                              const S* __restrict__ x, const S* __restrict__ b,
                                              S* __restrict__ bc, S* __restrict__ xc, double tau,
                                              const PcgControl* ctl) {
  if (ctl->done) return;
  FOR_LEVEL_CELLS_VB(Cl, vb)
  S sum = (S)0;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int fi = 2 * i + dx, fj = 2 * j + dy, fk = 2 * k + dz;
        if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;  // padded coarse cell
        const int fc = fi + F.nx * (fj + F.ny * fk);
        sum += b[fc] - s_apply<S>(F, x, fc, fi, fj, fk);
      }
  bc[c] = sum;
  xc[c] = lv_scale<S>(Cl, tau, c, i, j, k) * sum;
  END_CELLS
}
template <typename S>
__global__ void __launch_bounds__(kBlock)
k_mg_restrict(MgLevel F, MgLevel Cl, const S* __restrict__ x, const S* __restrict__ b,
              S* __restrict__ bc, S* __restrict__ xc, double tau, const PcgControl* ctl) {
  b_mg_restrict<S>(hw_block(), F, Cl, x, b, bc, xc, tau, ctl);
}

// Trilinear prolongation of the coarse correction, added to the fine x.
template <typename S>
__device__ __forceinline__ void b_mg_prolong(VBlock vb, const MgLevel& F, const MgLevel& Cl,
                                             const S* __restrict__ xc, const S* __restrict__ xin,
                                             S* __restrict__ xout, const PcgControl* ctl) {
  if (ctl->done) return;
  FOR_LEVEL_CELLS_VB(F, vb)
  xout[c] = xin[c] + prolong_value<S>(Cl, xc, i, j, k, lv_fine_fluid(F, c));
  END_CELLS
}
template <typename S>
__global__ void __launch_bounds__(kBlock)
k_mg_prolong(MgLevel F, MgLevel Cl, const S* __restrict__ xc, const S* __restrict__ xin,
             S* __restrict__ xout, const PcgControl* ctl) {
  b_mg_prolong<S>(hw_block(), F, Cl, xc, xin, xout, ctl);
}

// Coarsest level: one workgroup solves K x = b - mean_a(b) with f64 CG in
// LDS (n <= kMgCoarsestMax) over the active cells a, x mean-free over a and
// 0 elsewhere; tolerance relative 1e-12 or n+2 iterations.
constexpr int kCoarseThreads = 1024;

template <typename S>
__device__ __forceinline__ void b_mg_coarsest(const MgLevel& L, const S* __restrict__ b,
                                              S* __restrict__ xout, const PcgControl* ctl) {
  if (ctl->done) return;
  __shared__ double xs[kMgCoarsestMax], rs[kMgCoarsestMax], ds[kMgCoarsestMax], qs[kMgCoarsestMax];
  __shared__ double red[kCoarseThreads / 64];
  __shared__ double bcast;
  const int n = L.nx * L.ny * L.nz;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  auto block_sum = [&](double v) -> double {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (t == 0) {
      double a = 0.0;
      for (int w = 0; w < kCoarseThreads / 64; ++w) a += red[w];
      bcast = a;
    }
    __syncthreads();
    const double out = bcast;
    __syncthreads();
    return out;
  };
  double local = 0.0, active = 0.0;
  for (int c = t; c < n; c += kCoarseThreads)
    if (lv_active(L, c)) {
      local += (double)b[c];
      active += 1.0;
    }
  const double na = block_sum(active);
  const double mean = na > 0.0 ? block_sum(local) / na : 0.0;
  double rr_local = 0.0;
  for (int c = t; c < n; c += kCoarseThreads) {
    const double r = lv_active(L, c) ? (double)b[c] - mean : 0.0;
    xs[c] = 0.0;
    rs[c] = r;
    ds[c] = r;
    rr_local += r * r;
  }
  double rr = block_sum(rr_local);
  const double rr0 = rr;
  for (int it = 0; it < n + 2 && rr > 1e-24 * rr0 && rr > 0.0; ++it) {
    double dq_local = 0.0;
    for (int c = t; c < n; c += kCoarseThreads) {
      const int i = c % L.nx, j = (c / L.nx) % L.ny, k = c / (L.nx * L.ny);
      const double v = s_apply<double>(L, ds, c, i, j, k);
      qs[c] = v;
      dq_local += ds[c] * v;
    }
    const double dq = block_sum(dq_local);
    if (!(dq > 0.0)) break;
    const double alpha = rr / dq;
    double rn_local = 0.0;
    for (int c = t; c < n; c += kCoarseThreads) {
      xs[c] += alpha * ds[c];
      rs[c] -= alpha * qs[c];
      rn_local += rs[c] * rs[c];
    }
    const double rn = block_sum(rn_local);
    const double beta = rn / rr;
    rr = rn;
    for (int c = t; c < n; c += kCoarseThreads) ds[c] = rs[c] + beta * ds[c];
    __syncthreads();
  }
  double xm_local = 0.0;
  for (int c = t; c < n; c += kCoarseThreads)
    if (lv_active(L, c)) xm_local += xs[c];
  const double xmean = na > 0.0 ? block_sum(xm_local) / na : 0.0;
  for (int c = t; c < n; c += kCoarseThreads)
    xout[c] = lv_active(L, c) ? (S)(xs[c] - xmean) : (S)0;
}
template <typename S>
__global__ void __launch_bounds__(kCoarseThreads)
k_mg_coarsest(MgLevel L, const S* __restrict__ b, S* __restrict__ xout, const PcgControl* ctl) {
  b_mg_coarsest<S>(L, b, xout, ctl);
}

// Coarsest level with the dense projected inverse: x = G b, one workgroup.
// Row c of G b is split over P = min(kThreads, kDenseRed) / n threads,
// each summing a contiguous run of columns (loads issued together,
// coalesced over c); the P partial sums of a row are added in part order
// through LDS (`red`, kDenseRed doubles), so the result is deterministic.
constexpr int kDenseThreads = 512;
constexpr int kDenseRed = 512;

template <int kThreads, typename S>
__device__ __forceinline__ void dense_solve(int n, const double* __restrict__ G, const S* __restrict__ b,
                                            S* __restrict__ x, double* bs, double* red) {
  for (int c = threadIdx.x; c < n; c += kThreads) bs[c] = (double)b[c];
  __syncthreads();
  constexpr int kSpan = kThreads < kDenseRed ? kThreads : kDenseRed;
  const int parts = n > 0 && n <= kSpan ? kSpan / n : 1;
  const int t = threadIdx.x;
  if (parts > 1) {
    const int c = t % n, part = t / n;
    const int j0 = (int)((long long)part * n / parts), j1 = (int)((long long)(part + 1) * n / parts);
    double acc = 0.0;
    if (part < parts)
      for (int j2 = j0; j2 < j1; ++j2) acc += G[(size_t)j2 * n + c] * bs[j2];
    if (part < parts) red[part * n + c] = acc;
    __syncthreads();
    if (t < n) {
      double sum = 0.0;
      for (int q = 0; q < parts; ++q) sum += red[q * n + t];
      x[t] = (S)sum;
    }
  } else {
    for (int c = t; c < n; c += kThreads) {
      double acc = 0.0;
      for (int j2 = 0; j2 < n; ++j2) acc += G[(size_t)j2 * n + c] * bs[j2];
      x[c] = (S)acc;
    }
  }
  __syncthreads();
}

template <typename S>
__global__ void __launch_bounds__(kDenseThreads)
k_mg_coarsest_dense(MgLevel L, const double* __restrict__ minv, const S* __restrict__ b,
                    S* __restrict__ xout, const PcgControl* ctl) {
  if (ctl->done) return;
  __shared__ double bs[kMgDenseMax];
  __shared__ double red[kDenseRed];
  dense_solve<kDenseThreads, S>(L.nx * L.ny * L.nz, minv, b, xout, bs, red);
}
template <typename S>
__global__ void __launch_bounds__(kDenseThreads)
kb_mg_coarsest_dense(const BCoarseArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  const BCoarseArgs& a = A[batch_scene(prefix, n, blockIdx.x)];
  if (a.ctl->done) return;
  __shared__ double bs[kMgDenseMax];
  __shared__ double red[kDenseRed];
  dense_solve<kDenseThreads, S>(a.L->nx * a.L->ny * a.L->nz, a.minv, (const S*)a.b, (S*)a.x, bs, red);
}

// ------------------------------------------------------------- V-cycle tail
// Levels l0..L-1 (together within kMgTailLds bytes) run inside one
// workgroup of kTailThreads threads with every level array in LDS: the
// iterate x, its ping-pong partner y and the right-hand side b (V-cycle
// precision), and for obstacle levels the face-fraction words and the
// octant bytes; phases are separated by workgroup barriers instead of
// kernel launches.  The V-cycle runs as one loop over its phases with each
// phase kind written once, so every level and sweep executes the same code
// (separate inlined copies per call site each started with a cold
// instruction cache, several microseconds per phase).
constexpr int kTailThreads = 1024;

// Cells t, t + T, t + 2T, ... (T = kTailThreads) of an nx x ny level with
// (i, j, k) carried from cell to cell: the tail runs on one compute unit, so
// its phases are bound by instruction issue, and two runtime integer
// divisions per cell cost more than the stencil itself.
struct TailWalk {
  int c, i, j, k, di, dj, dk;
  __device__ __forceinline__ TailWalk(int nx, int ny) {
    c = threadIdx.x;
    const int row = c / nx;
    i = c - row * nx;
    k = row / ny;
    j = row - k * ny;
    const int r = kTailThreads / nx;
    di = kTailThreads - r * nx;
    dk = r / ny;
    dj = r - dk * ny;
  }
  // c += T: i + di < 2 nx and j + carry + dj < 2 ny, so one wrap each
  __device__ __forceinline__ void next(int nx, int ny) {
    c += kTailThreads;
    i += di;
    if (i >= nx) {
      i -= nx;
      ++j;
    }
    j += dj;
    if (j >= ny) {
      j -= ny;
      ++k;
    }
    k += dk;
  }
};

// 1 / D of a tail cell: the per-cell values staged in LDS (levels with
// fractions), the level's 64-entry table staged in LDS (box levels), or
// computed as in the level kernels when neither fitted.
__device__ __forceinline__ double tail_odiag(const MgLevel& L, const unsigned* fr, const double* od,
                                             const TailWalk& w) {
  if (od) return fr ? od[w.c] : od[box_mask6(L.nx, L.ny, L.nz, w.i, w.j, w.k)];
  return lv_odiag_fr(L, fr, w.c, w.i, w.j, w.k);
}

template <typename S>
__device__ __forceinline__ void tail_smooth(const MgLevel& L, const unsigned* fr, const double* od, const S* xin,
                                         const S* b, S* xout, double tau) {
  const int n = L.nx * L.ny * L.nz;
  for (TailWalk w(L.nx, L.ny); w.c < n; w.next(L.nx, L.ny)) {
    const S ax = s_apply_fr<S>(L, fr, xin, w.c, w.i, w.j, w.k);
    xout[w.c] = xin[w.c] + (S)(tau * tail_odiag(L, fr, od, w)) * (b[w.c] - ax);
  }
  __syncthreads();
}

// FP32 tail levels with nx % 4 == 0 take four cells per lane: 16-byte LDS
// reads of the cells, their y/z neighbours, b and the fraction words, and the
// level kernels' quad expressions (b_mg_smooth_qf, k_mg_smooth_q), so each
// value equals tail_smooth's.  kResidual: out = b - K x (the restriction's
// first pass) instead of the sweep.
template <bool kResidual>
__device__ __forceinline__ void tail_sweep_q(const MgLevel& L, const unsigned* fr, const double* od,
                                          const float* xin, const float* b, float* out, double tau) {
  const int qx = L.nx / 4;
  const int nq = qx * L.ny * L.nz;
  const float w0 = (float)L.w[0], w1 = (float)L.w[1], w2 = (float)L.w[2];
  for (TailWalk w(qx, L.ny); w.c < nq; w.next(qx, L.ny)) {
    const int i0 = 4 * w.i, j = w.j, k = w.k, c = 4 * w.c;
    Quad<float> Q;
    load_quad(xin, L.nx, L.ny, L.nz, i0, j, k, c, Q);
    float bb[4], xn[4];
    load4(b, c, bb);
    if (fr) {
      FracRow R;
      load_frac_row(fr, L, i0, j, k, c, R);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const Frac6 f = frac6_of(R, u, i0 + u + 1 < L.nx);
        const float ax = quad_apply_frac(L, Q, u, f);
        if (kResidual) {
          xn[u] = bb[u] - ax;
        } else {
          double o;
          if (od) {
            o = od[c + u];
          } else {
            const double d = L.wf[0] * (double)(f.xm + f.xp) + L.wf[1] * (double)(f.ym + f.yp) +
                             L.wf[2] * (double)(f.zm + f.zp);
            o = d > 0.0 ? 1.0 / d : 0.0;
          }
          xn[u] = Q.c[u] + (float)(tau * o) * (bb[u] - ax);
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float ax = quad_apply_f(Q, u, w0, w1, w2);
        if (kResidual) {
          xn[u] = bb[u] - ax;
        } else {
          const int m = box_mask6(L.nx, L.ny, L.nz, i0 + u, j, k);
          xn[u] = Q.c[u] + (float)(tau * (od ? od[m] : L.odiag[m])) * (bb[u] - ax);
        }
      }
    }
    store4(out, c, xn);
  }
  __syncthreads();
}

// Restriction in two passes: the fine residuals b - K x into r (every
// thread busy), then the eight-children sums in (dz, dy, dx) order (the
// same values and order as k_mg_restrict).
template <typename S>
__device__ __forceinline__ void tail_restrict(const MgLevel& F, const unsigned* ffr, const MgLevel& Cl,
                                           const unsigned* cfr, const double* cod, const S* x, const S* b,
                                           S* r, S* bc, S* xc, double tau) {
  if constexpr (sizeof(S) == 4) {
    if (F.nx % 4 == 0) {
      tail_sweep_q<true>(F, ffr, nullptr, x, b, r, 0.0);
    } else {
      const int nf = F.nx * F.ny * F.nz;
      for (TailWalk w(F.nx, F.ny); w.c < nf; w.next(F.nx, F.ny))
        r[w.c] = b[w.c] - s_apply_fr<S>(F, ffr, x, w.c, w.i, w.j, w.k);
      __syncthreads();
    }
  } else {
    const int nf = F.nx * F.ny * F.nz;
    for (TailWalk w(F.nx, F.ny); w.c < nf; w.next(F.nx, F.ny))
      r[w.c] = b[w.c] - s_apply_fr<S>(F, ffr, x, w.c, w.i, w.j, w.k);
    __syncthreads();
  }
  const int n = Cl.nx * Cl.ny * Cl.nz;
  for (TailWalk w(Cl.nx, Cl.ny); w.c < n; w.next(Cl.nx, Cl.ny)) {
    S sum = (S)0;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const int fi = 2 * w.i + dx, fj = 2 * w.j + dy, fk = 2 * w.k + dz;
          if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;  // padded coarse cell
          sum += r[fi + F.nx * (fj + F.ny * fk)];
        }
    bc[w.c] = sum;
    xc[w.c] = (S)(tau * tail_odiag(Cl, cfr, cod, w)) * sum;
  }
  __syncthreads();
}

// xout = xin + trilinear prolongation of xc (prolong_value with the coarse
// level's fractions and octants read from LDS)
template <typename S>
__device__ __forceinline__ void tail_prolong(const MgLevel& F, const unsigned* ffr, const MgLevel& Cl,
                                          const unsigned* cfr, const unsigned char* coct, const S* xc,
                                          const S* xin, S* xout) {
  const int n = F.nx * F.ny * F.nz;
  for (TailWalk w(F.nx, F.ny); w.c < n; w.next(F.nx, F.ny)) {
    const bool fine_fluid = ffr ? (ffr[w.c] & kFracFluid) != 0 : lv_fine_fluid(F, w.c);
    xout[w.c] = xin[w.c] + prolong_value_fr<S>(Cl, cfr, coct, xc, w.i, w.j, w.k, fine_fluid);
  }
  __syncthreads();
}

template <typename S>
__device__ __forceinline__ void tail_copy(int n, const S* src, S* dst) {
  for (int c = threadIdx.x; c < n; c += kTailThreads) dst[c] = src[c];
  __syncthreads();
}

// Dense coarsest solve x = G b inside the tail (dense_solve's partition and
// summation order); up to kTailDenseRegs entries of G per thread are loaded
// into registers when the tail starts, so their latency overlaps the
// staging instead of stalling the solve.
constexpr int kTailDenseRegs = 8;
struct TailDense {
  int n, c, part, parts, j0, j1;
  bool pre;  // every thread's share fits kTailDenseRegs
  double g[kTailDenseRegs];
  __device__ __forceinline__ void load(int cells, const double* __restrict__ G) {
    n = cells;
    parts = n > 0 && n <= kDenseRed ? kDenseRed / n : 1;
    pre = parts > 1 && n <= kDenseRed / kTailDenseRegs;
    const int t = threadIdx.x;
    c = pre ? t % n : 0;
    part = pre ? t / n : parts;
    j0 = (int)((long long)part * n / parts);
    j1 = (int)((long long)(part + 1) * n / parts);
#pragma unroll
    for (int q = 0; q < kTailDenseRegs; ++q)
      g[q] = (pre && part < parts && j0 + q < j1) ? G[(size_t)(j0 + q) * n + c] : 0.0;
  }
};

template <typename S>
__device__ __forceinline__ void tail_dense(const TailDense& D, const double* __restrict__ G, const S* b, S* x,
                                        double* bs, double* red) {
  if (!D.pre) {
    dense_solve<kTailThreads, S>(D.n, G, b, x, bs, red);
    return;
  }
  for (int c = threadIdx.x; c < D.n; c += kTailThreads) bs[c] = (double)b[c];
  __syncthreads();
  if (D.part < D.parts) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < kTailDenseRegs; ++q)
      if (D.j0 + q < D.j1) acc += D.g[q] * bs[D.j0 + q];
    red[D.part * D.n + D.c] = acc;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < D.n) {
    double sum = 0.0;
    for (int q = 0; q < D.parts; ++q) sum += red[q * D.n + t];
    x[t] = (S)sum;
  }
  __syncthreads();
}

// Lower tail levels staged in one pass (their loads in flight together).
constexpr int kTailStageLower = 4;

template <typename S>
__device__ __forceinline__ void b_mg_tail(const MgHierarchy* __restrict__ H, int l0,
                                          const PcgControl* ctl) {
  if (ctl->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double bs[kMgDenseMax];
  __shared__ double red[kDenseRed];
  const int levels = H->levels;
  auto at = [&](int l, int q) -> void* { return H->tail_off[l][q] >= 0 ? smem + H->tail_off[l][q] : nullptr; };
  auto X = [&](int l) { return (S*)at(l, 0); };
  auto Y = [&](int l) { return (S*)at(l, 1); };
  auto B = [&](int l) { return (S*)at(l, 2); };
  auto FR = [&](int l) { return (const unsigned*)at(l, 3); };
  auto OC = [&](int l) { return (const unsigned char*)at(l, 4); };
  auto OD = [&](int l) { return (const double*)at(l, 5); };
  auto cells = [&](int l) { return H->lv[l].nx * H->lv[l].ny * H->lv[l].nz; };
  TailDense dense;
  dense.load(cells(levels - 1), H->coarse_inv);
  // Stage from global memory: level l0's b and x (the parent's restriction
  // stored the first pre-smoothing sweep there), every level's fractions,
  // octants and per-cell 1/D, and the 64-entry 1/D tables of box levels.
  {
    const int n0 = cells(l0);
    const S* gb = (const S*)H->b[l0];
    const S* gx = (const S*)H->x[l0];
    const unsigned* gf = H->tail_off[l0][3] >= 0 ? H->lv[l0].frac : nullptr;
    const double* go = H->tail_off[l0][5] >= 0 ? H->tail_odc[l0] : nullptr;
    for (int c0 = threadIdx.x; c0 < n0; c0 += 4 * kTailThreads) {
      S vb[4], vx[4];
      unsigned vf[4];
      double vo[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = c0 + u * kTailThreads;
        if (c < n0) {
          vb[u] = gb[c];
          vx[u] = gx[c];
          vf[u] = gf ? gf[c] : 0u;
          vo[u] = go ? go[c] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = c0 + u * kTailThreads;
        if (c < n0) {
          B(l0)[c] = vb[u];
          X(l0)[c] = vx[u];
          if (gf) ((unsigned*)at(l0, 3))[c] = vf[u];
          if (go) ((double*)at(l0, 5))[c] = vo[u];
        }
      }
    }
    const int n1 = l0 + 1 < levels ? cells(l0 + 1) : 0;
    for (int c = threadIdx.x; c < n1; c += kTailThreads) {
      unsigned vf[kTailStageLower];
      unsigned char vc[kTailStageLower];
      double vo[kTailStageLower];
#pragma unroll
      for (int q = 0; q < kTailStageLower; ++q) {
        const int l = l0 + 1 + q;
        const bool in = l < levels && c < cells(l);
        vf[q] = in && H->tail_off[l][3] >= 0 ? H->lv[l].frac[c] : 0u;
        vc[q] = in && H->tail_off[l][4] >= 0 ? H->lv[l].octant[c] : (unsigned char)0;
        vo[q] = in && H->tail_odc[l] ? H->tail_odc[l][c] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kTailStageLower; ++q) {
        const int l = l0 + 1 + q;
        if (l < levels && c < cells(l)) {
          if (H->tail_off[l][3] >= 0) ((unsigned*)at(l, 3))[c] = vf[q];
          if (H->tail_off[l][4] >= 0) ((unsigned char*)at(l, 4))[c] = vc[q];
          if (H->tail_odc[l]) ((double*)at(l, 5))[c] = vo[q];
        }
      }
    }
    for (int l = l0 + 1 + kTailStageLower; l < levels; ++l) {
      const int n = cells(l);
      if (H->tail_off[l][3] >= 0)
        for (int c = threadIdx.x; c < n; c += kTailThreads) ((unsigned*)at(l, 3))[c] = H->lv[l].frac[c];
      if (H->tail_off[l][4] >= 0)
        for (int c = threadIdx.x; c < n; c += kTailThreads) ((unsigned char*)at(l, 4))[c] = H->lv[l].octant[c];
      if (H->tail_odc[l])
        for (int c = threadIdx.x; c < n; c += kTailThreads) ((double*)at(l, 5))[c] = H->tail_odc[l][c];
    }
    if (threadIdx.x < 64)
      for (int l = l0; l < levels; ++l)
        if (H->tail_off[l][5] >= 0 && !H->lv[l].frac) ((double*)at(l, 5))[threadIdx.x] = H->lv[l].odiag[threadIdx.x];
    __syncthreads();
  }
  // Every tail level's iterate starts in X(l); bit l of `in_y` says it has
  // moved to Y(l) (a bit set instead of per-level pointer arrays, which
  // would live in scratch memory).
  unsigned in_y = 0;
  auto cur = [&](int l) { return ((in_y >> l) & 1u) ? Y(l) : X(l); };
  auto nxt = [&](int l) { return ((in_y >> l) & 1u) ? X(l) : Y(l); };
  // Phase loop.  Going down, level l takes pre - 1 sweeps then restricts
  // to l + 1; the coarsest level is solved densely; going up, level l
  // takes the prolongation from l + 1, post sweeps and (below l0) a copy
  // of its result into X(l).
  int l = l0, step = 0;
  bool down = l0 < levels - 1;
  for (;;) {
    enum { kSmooth, kRestrict, kDense, kProlong, kCopy, kDone } op;
    double tau = 0.0;
    if (down) {
      if (step < H->cpre - 1) {
        op = kSmooth;
        tau = H->tau_cpre[step + 1];
      } else {
        op = kRestrict;
      }
    } else if (l == levels - 1) {
      op = kDense;
    } else if (step == 0) {
      op = kProlong;
    } else if (step <= H->cpost) {
      op = kSmooth;
      tau = H->tau_cpost[step - 1];
    } else {
      op = (l > l0 && ((in_y >> l) & 1u)) ? kCopy : kDone;
    }
    if (!down && l == levels - 1) op = kDense;
    const MgLevel& L = H->lv[l];
    switch (op) {
      case kSmooth:
        if constexpr (sizeof(S) == 4) {
          if (L.nx % 4 == 0)
            tail_sweep_q<false>(L, FR(l), OD(l), cur(l), B(l), nxt(l), tau);
          else
            tail_smooth<S>(L, FR(l), OD(l), cur(l), B(l), nxt(l), tau);
        } else {
          tail_smooth<S>(L, FR(l), OD(l), cur(l), B(l), nxt(l), tau);
        }
        break;
      case kRestrict:
        // nxt(l) is free until the prolongation: it holds the residuals
        tail_restrict<S>(L, FR(l), H->lv[l + 1], FR(l + 1), OD(l + 1), cur(l), B(l), nxt(l), B(l + 1),
                         X(l + 1), H->tau_cpre[0]);
        break;
      case kDense:
        tail_dense<S>(dense, H->coarse_inv, B(l), X(l), bs, red);
        break;
      case kProlong:
        tail_prolong<S>(L, FR(l), H->lv[l + 1], FR(l + 1), OC(l + 1), X(l + 1), cur(l), nxt(l));
        break;
      case kCopy:
        tail_copy<S>(cells(l), cur(l), X(l));
        break;
      default:
        break;
    }
    // advance
    if (op == kSmooth || op == kProlong) {
      in_y ^= 1u << l;
      ++step;
    } else if (op == kRestrict) {
      ++l;
      step = 0;
      if (l == levels - 1) down = false;
    } else if (op == kDense) {
      if (l == l0) break;  // the tail is only the coarsest level
      --l;
      step = 0;
      down = false;
    } else {  // kCopy or kDone: level l finished
      if (l == l0) break;
      --l;
      step = 0;
    }
  }
  // level l0's result back to global x[l0] for the parent's prolongation
  {
    S* gx = (S*)H->x[l0];
    const int n = cells(l0);
    const S* r = cur(l0);
    for (int c = threadIdx.x; c < n; c += kTailThreads) gx[c] = r[c];
  }
}
template <typename S>
__global__ void __launch_bounds__(kTailThreads)
k_mg_tail(const MgHierarchy* __restrict__ H, int l0, const PcgControl* ctl) {
  b_mg_tail<S>(H, l0, ctl);
}
template <typename S>
__global__ void __launch_bounds__(kTailThreads)
kb_mg_tail(const BTailArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  const BTailArgs& a = A[batch_scene(prefix, n, blockIdx.x)];
  b_mg_tail<S>(a.H, a.l0, a.ctl);
}

// p += x: a one-level hierarchy (no axis can be halved) solves the finest
// level directly, so the correction is added here.
template <typename S>
__global__ void __launch_bounds__(kBlock)
k_mg_add(DevGrid g, const S* __restrict__ x, double* __restrict__ p, const PcgControl* ctl) {
  if (ctl->done) return;
  FOR_CELLS(g)
  p[c] = p[c] + (double)x[c];
  END_CELLS
}

// p <- p - p0 with p0 = p[gauge] captured in ctl->spare[0] beforehand (a
// split fluid: each region's p at its own gauge, captured per region);
// solid cells (not unknowns) are set to 0.
// The gauge shift of the returned pressure, decided on the device from the
// solve's terminal state: a solve that ended OK or at its iteration limit
// has p[gauge] subtracted (solid cells set to 0); when no cycle ran and p
// was never written (p_unwritten: the first cycle would have written all
// of it), p becomes 0.  Any other outcome leaves p as it is.
__device__ __forceinline__ int regauge_mode(const PcgControl* ctl, bool p_unwritten) {
  if (p_unwritten && ctl->iteration == 0) return 2;  // zero-fill
  return ctl->done && (ctl->status == RH_OK || ctl->status == RH_ITERATION_LIMIT) ? 1 : 0;
}
__global__ void k_capture_gauge(const double* __restrict__ p, int gauge, PcgControl* ctl, bool p_unwritten) {
  if (threadIdx.x == 0 && threadIdx.y == 0 && regauge_mode(ctl, p_unwritten) == 1) ctl->spare[0] = p[gauge];
}
__global__ void k_capture_comp_gauges(DevGrid g, const double* __restrict__ p, const PcgControl* ctl,
                                      bool p_unwritten) {
  if (regauge_mode(ctl, p_unwritten) != 1) return;
  for (int k = threadIdx.x; k < g.ncomp; k += blockDim.x) g.comp_shift[k] = p[g.comp_gauge[k]];
}

__global__ void __launch_bounds__(kBlock)
k_regauge(DevGrid g, double* __restrict__ p, const PcgControl* ctl, bool p_unwritten) {
  const int mode = regauge_mode(ctl, p_unwritten);
  if (!mode) return;
  const double p0 = ctl->spare[0];
  FOR_CELLS(g)
  p[c] = mode == 2 ? 0.0 : (is_fluid(g, c) ? p[c] - (g.ncomp > 1 ? g.comp_shift[g.comp_label[c]] : p0) : 0.0);
  END_CELLS
}


// ------------------------------------------------------ batched (FP32)

#define RH_BATCH_PROLOGUE(ARGS)                                  \
  const int s_ = batch_scene(prefix, n, blockIdx.x);            \
  const ARGS& a = A[s_];                                        \
  const VBlock vb = batch_block(prefix, s_);

__global__ void __launch_bounds__(kBlock)
kb_mg_outer_residual(const BOuterArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BOuterArgs)
  b_mg_outer_residual<float>(vb, *a.g, *a.L0, a.p, a.b, a.b0, a.x0, a.tau, a.ctl, a.pdd, a.pmx);
}
__global__ void __launch_bounds__(kBlock)
kb_mg_smooth(const BSmoothArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BSmoothArgs)
  b_mg_smooth<float>(vb, *a.L, a.xin, a.b, a.xout, a.p, a.tau, a.ctl);
}
__global__ void __launch_bounds__(kBlock)
kb_mg_restrict(const BRestrictArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BRestrictArgs)
  b_mg_restrict<float>(vb, *a.F, *a.C, a.x, a.b, a.bc, a.xc, a.tau, a.ctl);
}
__global__ void __launch_bounds__(kBlock)
kb_mg_prolong(const BProlongArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BProlongArgs)
  b_mg_prolong<float>(vb, *a.F, *a.C, a.xc, a.xin, a.xout, a.ctl);
}

void launch_b_mg_outer(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_outer_residual<<<bl.blocks, block_shape(), 0, s>>>((const BOuterArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_smooth(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks) kb_mg_smooth<<<bl.blocks, block_shape(), 0, s>>>((const BSmoothArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_restrict(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_restrict<<<bl.blocks, block_shape(), 0, s>>>((const BRestrictArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_prolong(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_prolong<<<bl.blocks, block_shape(), 0, s>>>((const BProlongArgs*)bl.args, bl.prefix, bl.n);
}
// Dynamic LDS the tail kernels may use: kMgTailLds once the runtime has
// accepted that size for every tail kernel, else the 64 KB-workgroup share.
int mg_tail_lds_limit() {
  static const int limit = [] {
    const int want = kMgTailLds;
    const bool ok =
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_tail<float>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_tail<double>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&kb_mg_tail<float>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess;
    (void)hipGetLastError();
    return ok ? want : 56 * 1024;
  }();
  return limit;
}

void launch_b_mg_tail(const BatchLaunch& bl, int lds_bytes, hipStream_t s) {
  if (bl.blocks)
    kb_mg_tail<float><<<bl.blocks, kTailThreads, lds_bytes, s>>>((const BTailArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_coarsest_dense(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_coarsest_dense<float><<<bl.blocks, kDenseThreads, 0, s>>>((const BCoarseArgs*)bl.args, bl.prefix,
                                                                     bl.n);
}

// ------------------------------------------------------------ launchers

static inline dim3 level_grid(const MgLevel& L) { return dim3(level_blocks((long long)L.nx * L.ny * L.nz)); }

template <typename S>
int mg_outer_blocks(const DevGrid& g, const MgHierarchy& h, const double* p, const double* b) {
  const bool quad = sizeof(S) == 4 && h.quad && mg_quad_outer_ok(g, h.lv[0], p, b);
  return quad ? mg_quad_outer_blocks(g) : (int)cells_grid(g).x;
}

template <typename S>
void mg_outer_residual(const DevGrid& g, const MgHierarchy& h, const double* p, const double* b,
                       PcgControl* ctl, DD* pdd, double* pmx, int it, hipStream_t s, bool x0_implicit,
                       bool p_zero, bool finish, int* progress) {
  dim3 grid = cells_grid(g);
  {
    prof::Scope ps(prof::kOuterResidual, s, it);
    const bool quad = sizeof(S) == 4 && h.quad && mg_quad_outer_ok(g, h.lv[0], p, b);
    if (quad) grid = dim3(mg_quad_outer_blocks(g));
    if (quad)
      launch_mg_outer_residual_q(g, h.lv[0], p, b, (float*)h.b[0], x0_implicit ? nullptr : (float*)h.x[0],
                                 h.tau_pre[0], ctl, pdd, pmx, s, p_zero);
    else
      launch(k_mg_outer_residual<S>, grid, block_shape(), 0, s, g, h.lv[0], p, b, (S*)h.b[0], (S*)h.x[0],
                                                            h.tau_pre[0], ctl, pdd, pmx);
  }
  if (finish) launch_finish<1, 2>(pdd, pmx, grid.x, FinMgTrue{ctl, it, progress}, s);
}

// Profiler slot of a V-cycle kernel: finest-level kernels have their own
// slots (tagged with the iteration); all coarser work shares one slot.
static inline int mg_slot(int l, int finest_slot) { return l == 0 ? finest_slot : prof::kMgCoarse; }

bool mg_quad_update(const MgHierarchy& h, const double* p) {
  return h.quad && mg_quad_level(h.lv[0]) && ((uintptr_t)p & 15u) == 0;
}

template <typename S>
static void smooth(const MgHierarchy& h, int l, const S* xin, S* xout, double* p, double tau,
                   const PcgControl* ctl, int it, hipStream_t s, bool p_zero = false) {
  prof::Scope ps(mg_slot(l, p ? prof::kMgSmoothUpdate : prof::kMgSmooth), s, l == 0 ? it : -2);
  const bool sized = mg_quad_size_ok(h.lv[l], l);
  if (sizeof(S) == 4 && h.quad && sized && !p && mg_quad_frac_level(h.lv[l]))
    launch_mg_smooth_qf(h.lv[l], (const float*)xin, (const float*)h.b[l], (float*)xout, tau, ctl, s);
  else if (sizeof(S) == 4 && h.quad && sized && mg_quad_level(h.lv[l]) && ((uintptr_t)p & 15u) == 0)
    launch_mg_smooth_q(h.lv[l], (const float*)xin, (const float*)h.b[l], (float*)xout, p, tau, ctl, s, p_zero);
  else
    launch(k_mg_smooth<S>, level_grid(h.lv[l]), block_shape(), 0, s, h.lv[l], xin, (const S*)h.b[l], xout,
                                                                 tau, p, ctl);
}

// V-cycle from level l; on entry x[l] holds the first pre-smoothing sweep
// (x0_implicit, level 0: only b[0] is there and the second sweep
// recomputes the first from it).
template <typename S>
static void vcycle(const MgHierarchy& h, int l, double* p, const PcgControl* ctl, int it,
                   hipStream_t s, bool x0_implicit, bool p_zero, const MgFold* fold = nullptr) {
  S* x = (S*)h.x[l];
  S* y = (S*)h.y[l];
  const S* b = (const S*)h.b[l];
  if (h.tail_level > 0 && l == h.tail_level) {
    prof::Scope ps(prof::kMgCoarse, s, -2);
    launch(k_mg_tail<S>, 1, kTailThreads, h.tail_lds, s, h.d_self, l, ctl);
    return;
  }
  if (l == h.levels - 1) {
    prof::Scope ps(prof::kMgCoarse, s, -2);
    if (h.coarse_inv)
      launch(k_mg_coarsest_dense<S>, 1, kDenseThreads, 0, s, h.lv[l], h.coarse_inv, b, x, ctl);
    else
      launch(k_mg_coarsest<S>, 1, kCoarseThreads, 0, s, h.lv[l], b, x, (PcgControl*)ctl);
    if (l == 0) {
      DevGrid g{};
      g.nx = h.lv[0].nx;
      g.ny = h.lv[0].ny;
      g.nz = h.lv[0].nz;
      k_mg_add<S><<<cells_grid(g), block_shape(), 0, s>>>(g, x, p, ctl);
    }
    return;
  }
  // small FP32 coarse level: one launch per leg (mg_fused_kernels.hip)
  if (sizeof(S) == 4 && mg_leg_fusable(h, l)) {
    {
      prof::Scope ps(prof::kMgCoarse, s, -2);
      launch_mg_down_leg(h, l, ctl, s);
    }
    vcycle<S>(h, l + 1, p, ctl, it, s, false, false);
    prof::Scope ps(prof::kMgCoarse, s, -2);
    launch_mg_up_leg(h, l, ctl, s);
    return;
  }
  // pre-smoothing: sweeps 2..nu1 (ping-pong x <-> y), result in `cur`
  S* cur = mg_first_in_y(h, l) ? y : x;
  S* nxt = cur == x ? y : x;
  const int pre = l == 0 ? h.pre : h.cpre, post = l == 0 ? h.post : h.cpost;
  const double* tau_pre = l == 0 ? h.tau_pre : h.tau_cpre;
  const double* tau_post = l == 0 ? h.tau_post : h.tau_cpost;
  for (int sweep = 1; sweep < pre; ++sweep) {
    if (sweep == 1 && l == 0 && x0_implicit) {
      prof::Scope ps(prof::kMgSmoothFirst, s, it);
      launch_mg_smooth_first_q(h.lv[0], (const float*)b, (float*)nxt, h.tau_pre[0], h.tau_pre[1], ctl, s, fold);
    } else {
      smooth<S>(h, l, cur, nxt, nullptr, tau_pre[sweep], ctl, it, s);
    }
    S* t = cur;
    cur = nxt;
    nxt = t;
  }
  {
    S* first = (S*)(mg_first_in_y(h, l + 1) ? h.y[l + 1] : h.x[l + 1]);  // level l + 1's first sweep
    prof::Scope ps(mg_slot(l, prof::kMgRestrict), s, l == 0 ? it : -2);
    if (sizeof(S) == 4 && h.quad && mg_quad_size_ok(h.lv[l], l) && mg_quad_pair(h.lv[l], h.lv[l + 1]))
      launch_mg_restrict_q(h.lv[l], h.lv[l + 1], (const float*)cur, (const float*)b, (float*)h.b[l + 1],
                           (float*)first, h.tau_cpre[0], ctl, s);
    else
      launch(k_mg_restrict<S>, level_grid(h.lv[l + 1]), block_shape(), 0, s,
          h.lv[l], h.lv[l + 1], cur, b, (S*)h.b[l + 1], first, h.tau_cpre[0], ctl);
  }
  vcycle<S>(h, l + 1, p, ctl, it, s, false, false);
  {
    prof::Scope ps(mg_slot(l, prof::kMgProlong), s, l == 0 ? it : -2);
    if (sizeof(S) == 4 && h.quad && mg_quad_size_ok(h.lv[l], l) && mg_quad_prolong_pair(h.lv[l], h.lv[l + 1]))
      launch_mg_prolong_q(h.lv[l], h.lv[l + 1], (const float*)h.x[l + 1], (const float*)cur, (float*)nxt,
                          ctl, s);
    else
      launch(k_mg_prolong<S>, level_grid(h.lv[l]), block_shape(), 0, s, h.lv[l], h.lv[l + 1],
                                                                     (const S*)h.x[l + 1], cur, nxt, ctl);
  }
  {
    S* t = cur;
    cur = nxt;
    nxt = t;
  }
  for (int sweep = 0; sweep < post; ++sweep) {
    const bool last = sweep + 1 == post;
    if (last && l == 0) {
      smooth<S>(h, l, cur, nxt, p, tau_post[sweep], ctl, it, s, p_zero);  // p += x (finest, last sweep)
    } else {
      smooth<S>(h, l, cur, nxt, nullptr, tau_post[sweep], ctl, it, s);
      S* t = cur;
      cur = nxt;
      nxt = t;
    }
  }
  if (l > 0 && cur != x) {
    // keep the level result in x[l] for the parent's prolongation
    // (only reached when the sweep count leaves it in y)
    (void)hipMemcpyAsync(x, cur, sizeof(S) * (size_t)h.lv[l].nx * h.lv[l].ny * h.lv[l].nz,
                   hipMemcpyDeviceToDevice, s);
  }
}

template <typename S>
void mg_cycle(const MgHierarchy& h, double* p, const PcgControl* ctl, int it, hipStream_t s, bool x0_implicit,
              bool p_zero, const MgFold* fold) {
  vcycle<S>(h, 0, p, ctl, it, s, x0_implicit, p_zero, fold);
}

void mg_regauge(const DevGrid& g, double* p, PcgControl* ctl, hipStream_t s, bool p_unwritten) {
  if (g.ncomp > 1)
    k_capture_comp_gauges<<<1, 256, 0, s>>>(g, p, ctl, p_unwritten);
  else
    k_capture_gauge<<<1, 64, 0, s>>>(p, g.gauge, ctl, p_unwritten);
  k_regauge<<<cells_grid(g), block_shape(), 0, s>>>(g, p, ctl, p_unwritten);
}

template int mg_outer_blocks<float>(const DevGrid&, const MgHierarchy&, const double*, const double*);
template int mg_outer_blocks<double>(const DevGrid&, const MgHierarchy&, const double*, const double*);
template void mg_outer_residual<float>(const DevGrid&, const MgHierarchy&, const double*,
                                       const double*, PcgControl*, DD*, double*, int, hipStream_t, bool, bool,
                                       bool, int*);
template void mg_outer_residual<double>(const DevGrid&, const MgHierarchy&, const double*,
                                        const double*, PcgControl*, DD*, double*, int, hipStream_t, bool, bool,
                                        bool, int*);
template void mg_cycle<float>(const MgHierarchy&, double*, const PcgControl*, int, hipStream_t, bool, bool,
                              const MgFold*);
template void mg_cycle<double>(const MgHierarchy&, double*, const PcgControl*, int, hipStream_t, bool, bool,
                               const MgFold*);

}  // namespace rh
