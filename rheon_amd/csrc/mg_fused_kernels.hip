// Fused legs of the V-cycle on the small coarse levels (below the four-
// cells-per-lane size, above the LDS-resident tail).  There every sweep,
// restriction and prolongation is a launch of a few microseconds that moves
// a megabyte or less, so the level's time is launch latency.  Each leg runs
// as one launch of tile workgroups that stage their tile plus a halo in LDS
// and recompute the halo instead of exchanging it:
//
//   down leg  x1 = tau_1 b / D on the tile grown by `pre` cells, sweeps
//             2..pre on regions shrinking by one cell each, the residual
//             b - K x on the tile and its 8-children restriction to the
//             next level (b_c, and x_c = tau_1 b_c / D_c when the next
//             level is the tail); stores the pre-smoothed x in y[l].
//   up leg    x = y[l] + P x[l+1] on the tile grown by `post` cells, then
//             the post-sweeps on shrinking regions; stores the result in
//             x[l] for the parent's prolongation.
//
// Every value is formed by the expressions of the one-kernel-per-step path
// (s_apply_fr, lv_odiag_fr, lv_scale, prolong_value, b_mg_restrict), in the
// same order, so both paths give the same bits.
#include "common.hpp"
#include "kernels.hpp"
#include "mg.hpp"
#include "mg_stencil.hpp"
#include "profiler.hpp"
#include "stencil.hpp"

namespace rh {

namespace {

constexpr int kFT = 8;           // fine cells per tile edge
constexpr int kFThreads = 1024;

// A cubic LDS region of edge R at level origin (oi, oj, ok).
template <int R>
struct Reg {
  int oi, oj, ok;
  __device__ static constexpr int idx(int a, int b, int c) { return a + R * (b + R * c); }
};

// One stencil position of a region: K x and the cell's Jacobi scale tau / D,
// with the neighbour clamping of nbr_offsets (a wall neighbour is the cell
// itself) and the fraction words' upper faces (walls 0): s_apply_fr's and
// lv_odiag_fr's expressions on LDS copies.
template <bool kFrac, int R>
__device__ __forceinline__ float leg_apply(const MgLevel& L, const float* X, const unsigned* FR, int a, int b,
                                           int c, int gi, int gj, int gk, double* od) {
  constexpr int R1 = R + 1;
  const int lc = Reg<R>::idx(a, b, c);
  const bool lx = gi > 0, hx = gi + 1 < L.nx, ly = gj > 0, hy = gj + 1 < L.ny, lz = gk > 0, hz = gk + 1 < L.nz;
  const float xc = X[lc];
  const float xa = X[lx ? lc - 1 : lc], xb = X[hx ? lc + 1 : lc];
  const float ya = X[ly ? lc - R : lc], yb = X[hy ? lc + R : lc];
  const float za = X[lz ? lc - R * R : lc], zb = X[hz ? lc + R * R : lc];
  if (kFrac) {
    const int lf = a + R1 * (b + R1 * c);
    const unsigned q = FR[lf];
    const unsigned qx = FR[hx ? lf + 1 : lf], qy = FR[hy ? lf + R1 : lf], qz = FR[hz ? lf + R1 * R1 : lf];
    Frac6 f;
    f.xm = (int)(q & 1023u);
    f.ym = (int)((q >> 10) & 1023u);
    f.zm = (int)((q >> 20) & 1023u);
    f.xp = hx ? (int)(qx & 1023u) : 0;
    f.yp = hy ? (int)((qy >> 10) & 1023u) : 0;
    f.zp = hz ? (int)((qz >> 20) & 1023u) : 0;
    const double d = L.wf[0] * (double)(f.xm + f.xp) + L.wf[1] * (double)(f.ym + f.yp) +
                     L.wf[2] * (double)(f.zm + f.zp);
    *od = d > 0.0 ? 1.0 / d : 0.0;
    return (float)L.wf[0] * ((float)f.xm * (xc - xa) + (float)f.xp * (xc - xb)) +
           (float)L.wf[1] * ((float)f.ym * (xc - ya) + (float)f.yp * (xc - yb)) +
           (float)L.wf[2] * ((float)f.zm * (xc - za) + (float)f.zp * (xc - zb));
  }
  const int m = (int)lx | ((int)hx << 1) | ((int)ly << 2) | ((int)hy << 3) | ((int)lz << 4) | ((int)hz << 5);
  *od = L.odiag[m];
  const float w0 = (float)L.w[0], w1 = (float)L.w[1], w2 = (float)L.w[2];
  return w0 * ((xc - xa) + (xc - xb)) + w1 * ((xc - ya) + (xc - yb)) + w2 * ((xc - za) + (xc - zb));
}

// 1 / D of region cell (a, b, c) (the first sweep's scale).
template <bool kFrac, int R>
__device__ __forceinline__ double leg_odiag(const MgLevel& L, const unsigned* FR, int a, int b, int c, int gi,
                                            int gj, int gk) {
  if (kFrac) {
    constexpr int R1 = R + 1;
    const bool hx = gi + 1 < L.nx, hy = gj + 1 < L.ny, hz = gk + 1 < L.nz;
    const int lf = a + R1 * (b + R1 * c);
    const unsigned q = FR[lf];
    const unsigned qx = FR[hx ? lf + 1 : lf], qy = FR[hy ? lf + R1 : lf], qz = FR[hz ? lf + R1 * R1 : lf];
    const int xm = (int)(q & 1023u), ym = (int)((q >> 10) & 1023u), zm = (int)((q >> 20) & 1023u);
    const int xp = hx ? (int)(qx & 1023u) : 0, yp = hy ? (int)((qy >> 10) & 1023u) : 0,
              zp = hz ? (int)((qz >> 20) & 1023u) : 0;
    const double d = L.wf[0] * (double)(xm + xp) + L.wf[1] * (double)(ym + yp) + L.wf[2] * (double)(zm + zp);
    return d > 0.0 ? 1.0 / d : 0.0;
  }
  return L.odiag[box_mask6(L.nx, L.ny, L.nz, gi, gj, gk)];
}

// Calls f(a, b, c, gi, gj, gk) for the cells of sub-cube [O, R - O)^3 of
// region G that lie in level L, thread t taking cells t, t + kFThreads, ...
// (compile-time trip count, so the calls' loads can be in flight together).
template <int R, int O, typename F>
__device__ __forceinline__ void reg_cells(const MgLevel& L, const Reg<R>& G, F&& f) {
  constexpr int E = R - 2 * O, N = E * E * E;
#pragma unroll
  for (int q = 0; q < (N + kFThreads - 1) / kFThreads; ++q) {
    const int t = (int)threadIdx.x + q * kFThreads;
    const int a = O + t % E, b = O + (t / E) % E, c = O + t / (E * E);
    const int gi = G.oi + a, gj = G.oj + b, gk = G.ok + c;
    if (t < N && gi >= 0 && gj >= 0 && gk >= 0 && gi < L.nx && gj < L.ny && gk < L.nz) f(a, b, c, gi, gj, gk);
  }
}

// Copies level array v (cells of the region grown by G1 - R cells on its
// upper sides, in region order of edge G1) into LDS: every load of the
// thread is issued before any store.
template <int R, int E, typename T>
__device__ __forceinline__ void stage(const MgLevel& L, const Reg<R>& G, const T* __restrict__ v, T* out) {
  constexpr int N = E * E * E, PER = (N + kFThreads - 1) / kFThreads;
  T w[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = (int)threadIdx.x + q * kFThreads;
    const int a = t % E, b = (t / E) % E, c = t / (E * E);
    const int gi = G.oi + a, gj = G.oj + b, gk = G.ok + c;
    const bool in = t < N && gi >= 0 && gj >= 0 && gk >= 0 && gi < L.nx && gj < L.ny && gk < L.nz;
    w[q] = in ? v[gi + L.nx * (gj + L.ny * gk)] : T(0);
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = (int)threadIdx.x + q * kFThreads;
    if (t < N) out[t] = w[q];
  }
}

// One sweep x <- x + tau (b - K x) / D over sub-cube [O, R - O)^3.
template <bool kFrac, int R, int O>
__device__ __forceinline__ void leg_sweep(const MgLevel& L, const Reg<R>& G, const float* cur, float* nxt,
                                          const float* sb, const unsigned* sfr, double tau) {
  reg_cells<R, O>(L, G, [&](int a, int b, int c, int gi, int gj, int gk) {
    const int lc = Reg<R>::idx(a, b, c);
    double od;
    const float ax = leg_apply<kFrac, R>(L, cur, sfr, a, b, c, gi, gj, gk, &od);
    nxt[lc] = cur[lc] + (float)(tau * od) * (sb[lc] - ax);
  });
}

}  // namespace

// Down leg of level L (fine) onto Cl: one workgroup per 4^3 tile of Cl.
template <bool kFrac, int kH>
__global__ void __launch_bounds__(kFThreads)
k_mg_down_fused(MgLevel L, MgLevel Cl, const float* __restrict__ b, float* __restrict__ xpre,
                float* __restrict__ bc, float* __restrict__ xc, MgLegTaus taus, double tau_c1,
                const PcgControl* ctl) {
  if (ctl->done) return;
  constexpr int R = kFT + 2 * kH, R1 = R + 1;
  __shared__ float sb[R * R * R], sx0[R * R * R], sx1[R * R * R];
  __shared__ unsigned sfr[kFrac ? R1 * R1 * R1 : 1];
  const int I0 = blockIdx.x * (kFT / 2), J0 = blockIdx.y * (kFT / 2), K0 = blockIdx.z * (kFT / 2);
  const Reg<R> G{2 * I0 - kH, 2 * J0 - kH, 2 * K0 - kH};
  if constexpr (kFrac) stage<R, R1>(L, G, L.frac, sfr);
  stage<R, R>(L, G, b, sb);
  __syncthreads();
  // first sweep from zero: x = (S)(tau_1 / D) b
  reg_cells<R, 0>(L, G, [&](int a, int bb, int c, int gi, int gj, int gk) {
    const int lc = Reg<R>::idx(a, bb, c);
    sx0[lc] = (float)(taus.t[0] * leg_odiag<kFrac, R>(L, sfr, a, bb, c, gi, gj, gk)) * sb[lc];
  });
  __syncthreads();
  float* cur = sx0;
  float* nxt = sx1;
  if constexpr (kH >= 2) {
    leg_sweep<kFrac, R, 1>(L, G, cur, nxt, sb, sfr, taus.t[1]);
    __syncthreads();
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  if constexpr (kH >= 3) {
    leg_sweep<kFrac, R, 2>(L, G, cur, nxt, sb, sfr, taus.t[2]);
    __syncthreads();
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  // the pre-smoothed tile, then the residual b - K x of every tile cell
  reg_cells<R, kH>(L, G, [&](int a, int bb, int c, int gi, int gj, int gk) {
    const int lc = Reg<R>::idx(a, bb, c);
    xpre[gi + L.nx * (gj + L.ny * gk)] = cur[lc];
    double od;
    nxt[lc] = sb[lc] - leg_apply<kFrac, R>(L, cur, sfr, a, bb, c, gi, gj, gk, &od);
  });
  __syncthreads();
  // restriction: sum of the 8 children in (dz, dy, dx) order
  for (int t = threadIdx.x; t < (kFT / 2) * (kFT / 2) * (kFT / 2); t += kFThreads) {
    const int di = t % (kFT / 2), dj = (t / (kFT / 2)) % (kFT / 2), dk = t / ((kFT / 2) * (kFT / 2));
    const int I = I0 + di, J = J0 + dj, K = K0 + dk;
    if (I >= Cl.nx || J >= Cl.ny || K >= Cl.nz) continue;
    float sum = 0.0f;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const int fi = 2 * I + dx, fj = 2 * J + dy, fk = 2 * K + dz;
          if (fi >= L.nx || fj >= L.ny || fk >= L.nz) continue;  // padded coarse cell
          sum += nxt[Reg<R>::idx(fi - G.oi, fj - G.oj, fk - G.ok)];
        }
    const int cc = I + Cl.nx * (J + Cl.ny * K);
    bc[cc] = sum;
    if (xc) xc[cc] = lv_scale<float>(Cl, tau_c1, cc, I, J, K) * sum;
  }
}

// Up leg of level L from Cl: one workgroup per 8^3 tile of L.
template <bool kFrac, int kH>
__global__ void __launch_bounds__(kFThreads)
k_mg_up_fused(MgLevel L, MgLevel Cl, const float* __restrict__ xc, const float* __restrict__ xpre,
              const float* __restrict__ b, float* __restrict__ xout, MgLegTaus taus, const PcgControl* ctl) {
  if (ctl->done) return;
  constexpr int R = kFT + 2 * kH, R1 = R + 1;
  __shared__ float sb[R * R * R], sx0[R * R * R], sx1[R * R * R];
  __shared__ unsigned sfr[kFrac ? R1 * R1 * R1 : 1];
  const Reg<R> G{(int)blockIdx.x * kFT - kH, (int)blockIdx.y * kFT - kH, (int)blockIdx.z * kFT - kH};
  if constexpr (kFrac) stage<R, R1>(L, G, L.frac, sfr);
  stage<R, R>(L, G, b, sb);
  // prolongation of the coarse result onto the grown tile
  reg_cells<R, 0>(L, G, [&](int a, int bb, int c, int gi, int gj, int gk) {
    const int gc = gi + L.nx * (gj + L.ny * gk);
    sx0[Reg<R>::idx(a, bb, c)] = xpre[gc] + prolong_value<float>(Cl, xc, gi, gj, gk, lv_fine_fluid(L, gc));
  });
  __syncthreads();
  float* cur = sx0;
  float* nxt = sx1;
  if constexpr (kH >= 2) {
    leg_sweep<kFrac, R, 1>(L, G, cur, nxt, sb, sfr, taus.t[0]);
    __syncthreads();
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  if constexpr (kH >= 3) {
    leg_sweep<kFrac, R, 2>(L, G, cur, nxt, sb, sfr, taus.t[1]);
    __syncthreads();
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  // the last post-sweep on the tile, stored
  reg_cells<R, kH>(L, G, [&](int a, int bb, int c, int gi, int gj, int gk) {
    const int lc = Reg<R>::idx(a, bb, c);
    double od;
    const float ax = leg_apply<kFrac, R>(L, cur, sfr, a, bb, c, gi, gj, gk, &od);
    xout[gi + L.nx * (gj + L.ny * gk)] = cur[lc] + (float)(taus.t[kH - 1] * od) * (sb[lc] - ax);
  });
}

bool mg_leg_fusable(const MgHierarchy& h, int l) {
  const int bottom = h.tail_level > 0 ? h.tail_level : h.levels - 1;
  if (l < 1 || l >= bottom) return false;
  const MgLevel& L = h.lv[l];
  if (L.mask) return false;
  if (h.quad && mg_quad_size_ok(L, l)) return false;  // the four-cell kernels' size
  return h.cpre >= 1 && h.cpre <= kMgLegMaxSweeps && h.cpost >= 1 && h.cpost <= kMgLegMaxSweeps;
}

template <bool kFrac>
static void down_leg(const MgHierarchy& h, int l, const PcgControl* ctl, hipStream_t s) {
  const MgLevel& L = h.lv[l];
  const MgLevel& Cl = h.lv[l + 1];
  const int bottom = h.tail_level > 0 ? h.tail_level : h.levels - 1;
  // the tail stages the first sweep of its level from x; a fused level
  // below recomputes it from b
  float* first = l + 1 == bottom ? (float*)h.x[l + 1] : nullptr;
  MgLegTaus t{};
  for (int q = 0; q < h.cpre; ++q) t.t[q] = h.tau_cpre[q];
  const dim3 grid((Cl.nx + 3) / 4, (Cl.ny + 3) / 4, (Cl.nz + 3) / 4);
  auto go = [&](auto kernel) {
    launch(kernel, grid, dim3(kFThreads), 0, s, L, Cl, (const float*)h.b[l], (float*)h.y[l], (float*)h.b[l + 1],
           first, t, h.tau_cpre[0], ctl);
  };
  if (h.cpre == 1) go(k_mg_down_fused<kFrac, 1>);
  else if (h.cpre == 2) go(k_mg_down_fused<kFrac, 2>);
  else go(k_mg_down_fused<kFrac, 3>);
}

template <bool kFrac>
static void up_leg(const MgHierarchy& h, int l, const PcgControl* ctl, hipStream_t s) {
  const MgLevel& L = h.lv[l];
  const MgLevel& Cl = h.lv[l + 1];
  MgLegTaus t{};
  for (int q = 0; q < h.cpost; ++q) t.t[q] = h.tau_cpost[q];
  const dim3 grid((L.nx + kFT - 1) / kFT, (L.ny + kFT - 1) / kFT, (L.nz + kFT - 1) / kFT);
  auto go = [&](auto kernel) {
    launch(kernel, grid, dim3(kFThreads), 0, s, L, Cl, (const float*)h.x[l + 1], (const float*)h.y[l],
           (const float*)h.b[l], (float*)h.x[l], t, ctl);
  };
  if (h.cpost == 1) go(k_mg_up_fused<kFrac, 1>);
  else if (h.cpost == 2) go(k_mg_up_fused<kFrac, 2>);
  else go(k_mg_up_fused<kFrac, 3>);
}

void launch_mg_down_leg(const MgHierarchy& h, int l, const PcgControl* ctl, hipStream_t s) {
  if (h.lv[l].frac)
    down_leg<true>(h, l, ctl, s);
  else
    down_leg<false>(h, l, ctl, s);
}

void launch_mg_up_leg(const MgHierarchy& h, int l, const PcgControl* ctl, hipStream_t s) {
  if (h.lv[l].frac)
    up_leg<true>(h, l, ctl, s);
  else
    up_leg<false>(h, l, ctl, s);
}

}  // namespace rh
