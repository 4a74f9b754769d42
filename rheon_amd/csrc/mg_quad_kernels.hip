// Multigrid kernels with four consecutive x-cells per lane (levels whose nx
// is a multiple of 4 and whose operator is the box or the level-0 obstacle
// mask; FP32 V-cycle storage).  Each lane loads its cells and their y/z
// neighbours as 16-byte vectors and the two x-edge neighbours as scalars,
// so a stencil over four cells costs 7 vector loads instead of 28 scalar
// ones.  Arithmetic and operation order per cell are those of the scalar
// kernels (s_apply, apply_at, prolong_value), so both give identical values.
#include "common.hpp"
#include "kernels.hpp"
#include "mg.hpp"
#include "mg_stencil.hpp"
#include "batch.hpp"
#include "profiler.hpp"
#include "stencil.hpp"
#include "mg_quad.hpp"

namespace rh {

// Outer FP64 residual r = b - K p with the first pre-smoothing sweep fused
// (as k_mg_outer_residual); x0 = null: the sweep's result is not stored
// (b_mg_smooth_first_q recomputes it where it is needed).  kZeroP: p is
// known to be 0 (the solve's first residual), so it is not read; K 0 =
// +0.0 exactly, the values of the general form.  kReport: only the
// reductions (the residual report of a returned pressure: b_residual_report's
// sums and maxima, four cells per lane), no stores, no control block.
template <bool kZeroP = false, bool kReport = false>
__device__ __forceinline__ void
b_mg_outer_residual_q(VBlock vb, const DevGrid& g, const MgLevel& L0, const double* __restrict__ p, const double* __restrict__ b,
                      float* __restrict__ b0, float* __restrict__ x0, double tau, const PcgControl* ctl,
                      DD* pdd, double* pmx) {
  if (!kReport && ctl->done) return;
  DD s[1] = {{0.0, 0.0}};
  double m[2] = {0.0, 0.0};
  const int nz_ = g.nz;
  FOR_QUADS(g.nx, g.ny, g.div_q, g.div_y)
  Quad<double> Q;
  if (kZeroP) {
#pragma unroll
    for (int u = 0; u < 4; ++u) Q.c[u] = Q.xm[u] = Q.xp[u] = Q.ym[u] = Q.yp[u] = Q.zm[u] = Q.zp[u] = 0.0;
  } else {
    load_quad(p, g.nx, g.ny, g.nz, i0, j, k, c, Q);
  }
  double bb[4];
  load4(b, c, bb);
  const unsigned char* const mrow = row_mask(g.mask, g.nx, g.ny, g.nz, row, kRowPlain);
  const unsigned m4 = mrow ? *reinterpret_cast<const unsigned*>(mrow + c) : 0u;
  float rb[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (mrow) close_faces(Q, u, (int)((m4 >> (8 * u)) & 63u));
    const double v = quad_apply_d(Q, u, g);
    if (!isfinite(v)) m[1] = 1.0;
    const double r = bb[u] - v;
    dd_acc(s[0], r * r);
    m[0] = fmax(m[0], fabs(r));
    if (!kReport) rb[u] = (float)r;
  }
  if (!kReport) {
    store4(b0, c, rb);
    if (x0) {  // the first sweep's value, where it is stored
      const int yz = box_yz(g.ny, g.nz, j, k);
      float rx[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        rx[u] = (float)(tau * L0.odiag[quad_mask(mrow, m4, g.nx, i0, yz, u)]) * rb[u];
      store4(x0, c, rx);
    }
  }
  END_QUADS
  block_publish<1, 2>(s, m, pdd, pmx, vb);
}

// Sweep scales of a box-rule level: an open cell (mask 63) and the x-wall
// cells of a box row (62: x- closed, 61: x+ closed).
struct QuadScales {
  float w0, w1, w2, a63, a62, a61;
  __device__ QuadScales(const MgLevel& L, double tau)
      : w0((float)L.w[0]), w1((float)L.w[1]), w2((float)L.w[2]), a63((float)(tau * L.odiag[63])),
        a62((float)(tau * L.odiag[62])), a61((float)(tau * L.odiag[61])) {}
};

// One sweep of the four cells i0..i0+3 of row (j, k) (cell index c, row
// index row) whose neighbourhood Q is loaded: xn = x + (tau / D)(b - K x).
// Closed obstacle faces are resolved in Q.
__device__ __forceinline__ void smooth_quad(const MgLevel& L, const QuadScales& S, double tau, Quad<float>& Q,
                                            const float bb[4], int i0, int j, int k, int c, int row,
                                            float xn[4]) {
  const unsigned char* const mrow = row_mask(L.mask, L.nx, L.ny, L.nz, row, kRowPlain);
  // an interior box-rule row: the cells' masks are 63 but for the x-wall
  // cells (quad_box_mask), so the scales are the three entries above
  const bool plain = !mrow && j > 0 && j + 1 < L.ny && k > 0 && k + 1 < L.nz;
  float sc[4];
  if (plain) {
    sc[0] = i0 > 0 ? S.a63 : S.a62;
    sc[1] = S.a63;
    sc[2] = S.a63;
    sc[3] = i0 + 4 < L.nx ? S.a63 : S.a61;
  } else {
    const unsigned m4 = mrow ? *reinterpret_cast<const unsigned*>(mrow + c) : 0u;
    const int yz = box_yz(L.ny, L.nz, j, k);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int mk = quad_mask(mrow, m4, L.nx, i0, yz, u);
      if (mrow) close_faces(Q, u, mk);
      sc[u] = (float)(tau * L.odiag[mk]);
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const float ax = quad_apply_f(Q, u, S.w0, S.w1, S.w2);
    xn[u] = Q.c[u] + sc[u] * (bb[u] - ax);
  }
}

// Residual b - K x of four cells of a level-0 (box or obstacle mask) row
// whose neighbourhood Q is loaded.
__device__ __forceinline__ void residual_quad(const MgLevel& F, Quad<float>& Q, const float bb[4], int fc, int row,
                                              float w0, float w1, float w2, float r[4]) {
  const unsigned char* const mrow = row_mask(F.mask, F.nx, F.ny, F.nz, row, kRowPlain);
  const unsigned m4 = mrow ? *reinterpret_cast<const unsigned*>(mrow + fc) : 0u;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (mrow) close_faces(Q, u, (int)((m4 >> (8 * u)) & 63u));
    r[u] = bb[u] - quad_apply_f(Q, u, w0, w1, w2);
  }
}

// One smoothing sweep (as k_mg_smooth); kUpdate: p += x_out instead of
// storing x_out (the finest level's last sweep).  kZeroP: p is known to be
// 0 (the first cycle), so p = 0.0 + x_out is formed without reading it.
template <bool kUpdate, bool kZeroP = false>
__device__ __forceinline__ void
b_mg_smooth_q(VBlock vb, const MgLevel& L, const float* __restrict__ xin, const float* __restrict__ b,
              float* __restrict__ xout, double* __restrict__ p, double tau, const PcgControl* ctl) {
  if (ctl->done) return;
  const QuadScales S(L, tau);
  const int nz_ = L.nz;
  FOR_QUADS(L.nx, L.ny, L.div_q, L.div_y)
  Quad<float> Q;
  load_quad(xin, L.nx, L.ny, L.nz, i0, j, k, c, Q);
  float bb[4];
  load4(b, c, bb);
  double pv[4];
  if (kUpdate) {
    if (kZeroP) {
#pragma unroll
      for (int u = 0; u < 4; ++u) pv[u] = 0.0;
    } else {
      load4(p, c, pv);
    }
  }
  float xn[4];
  smooth_quad(L, S, tau, Q, bb, i0, j, k, c, row, xn);
  if (kUpdate) {
#pragma unroll
    for (int u = 0; u < 4; ++u) pv[u] = pv[u] + (double)xn[u];
    store4(p, c, pv);
  } else {
    store4(xout, c, xn);
  }
  END_QUADS
}


// One smoothing sweep on a level with face fractions (coarse levels with
// obstacles or padding), four cells per lane: the fraction words of the
// cells and of their +x/+y/+z neighbours come in as 16-byte vectors.  The
// stencil and the scale are s_apply's and lv_odiag's fraction forms
// verbatim, so the values equal the one-cell kernel's.
__device__ __forceinline__ void
b_mg_smooth_qf(VBlock vb, const MgLevel& L, const float* __restrict__ xin, const float* __restrict__ b,
               float* __restrict__ xout, double tau, const PcgControl* ctl) {
  if (ctl->done) return;
  const int nz_ = L.nz;
  FOR_QUADS(L.nx, L.ny, L.div_q, L.div_y)
  Quad<float> Q;
  load_quad(xin, L.nx, L.ny, L.nz, i0, j, k, c, Q);
  float bb[4];
  load4(b, c, bb);
  FracRow R;
  load_frac_row(L, i0, j, k, c, R);
  float xn[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const Frac6 f = frac6_of(R, u, i0 + u + 1 < L.nx);
    const float xc = Q.c[u];
    const float ax = quad_apply_frac(L, Q, u, f);
    const double d = L.wf[0] * (double)(f.xm + f.xp) + L.wf[1] * (double)(f.ym + f.yp) +
                     L.wf[2] * (double)(f.zm + f.zp);
    const double od = d > 0.0 ? 1.0 / d : 0.0;
    xn[u] = xc + (float)(tau * od) * (bb[u] - ax);
  }
  store4(xout, c, xn);
  END_QUADS
}

// The finest level's second pre-smoothing sweep reading b0 only: the first
// sweep's value x0 = (float)(tau1 / D) b0, which the outer residual no
// longer stores, is recomputed at every stencil position from b0 and the
// position's open-face mask (the same expression, so the same bits), then
// x1 = x0 + (float)(tau2 / D) (b0 - K x0).  Saves the x0 store and load
// (8 bytes per cell and cycle).
__device__ __forceinline__ void
b_mg_smooth_first_q(VBlock vb, const MgLevel& L, const float* __restrict__ b, float* __restrict__ xout,
                    double tau1, double tau2, const PcgControl* ctl) {
  if (ctl->done) return;
  __shared__ float s1[64], s2[64];  // (float)(tau / D) per open-face mask
  {
    const int t = flat_tid();
    if (t < 64) {
      s1[t] = (float)(tau1 * L.odiag[t]);
      s2[t] = (float)(tau2 * L.odiag[t]);
    }
    __syncthreads();
  }
  const float w0 = (float)L.w[0], w1 = (float)L.w[1], w2 = (float)L.w[2];
  // the scales of an open cell (mask 63) and of the x-wall cells of a box
  // row (62: x- closed, 61: x+ closed), the table's entries
  const float a63 = (float)(tau1 * L.odiag[63]), a62 = (float)(tau1 * L.odiag[62]);
  const float a61 = (float)(tau1 * L.odiag[61]);
  const float c63 = (float)(tau2 * L.odiag[63]), c62 = (float)(tau2 * L.odiag[62]);
  const float c61 = (float)(tau2 * L.odiag[61]);
  const int nz_ = L.nz;
  const int sy = L.nx, sz = L.nx * L.ny;
  FOR_QUADS(L.nx, L.ny, L.div_q, L.div_y)
  Quad<float> Q;
  load_quad(b, L.nx, L.ny, L.nz, i0, j, k, c, Q);
  float bb[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) bb[u] = Q.c[u];
  const unsigned char* const mrow = row_mask(L.mask, L.nx, L.ny, L.nz, row, kRowPlainNear);
  // b0 -> x0 = (float)(tau1 / D) b0 at every loaded position (cells, y/z
  // neighbours, the two x-edge neighbours), scaled in place; t2 = the
  // cells' (float)(tau2 / D), mc = their masks (obstacle rows)
  float t2[4], xl, xr;
  int mc[4] = {63, 63, 63, 63};
  // A box-rule row whose four neighbour rows are interior too (2 <= j <
  // ny - 2, 2 <= k < nz - 2): every position's mask is 63 but for the
  // x-wall cells (62 at i0 = 0, 61 at i0 + 4 = nx), which the same
  // positions of the neighbour rows share; the clamped x-edge neighbour of
  // a wall cell is the cell itself.  The box-rule branch below gives these
  // same masks, so both give the same values.
  const bool plain = !mrow && j >= 2 && j + 2 < L.ny && k >= 2 && k + 2 < L.nz;
  if (plain) {
    const float e0 = i0 > 0 ? a63 : a62, e3 = i0 + 4 < L.nx ? a63 : a61;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float e = u == 0 ? e0 : (u == 3 ? e3 : a63);
      Q.c[u] = e * Q.c[u];
      Q.ym[u] = e * Q.ym[u];
      Q.yp[u] = e * Q.yp[u];
      Q.zm[u] = e * Q.zm[u];
      Q.zp[u] = e * Q.zp[u];
    }
    xl = e0 * Q.xm[0];
    xr = e3 * Q.xp[3];
    t2[0] = i0 > 0 ? c63 : c62;
    t2[1] = c63;
    t2[2] = c63;
    t2[3] = i0 + 4 < L.nx ? c63 : c61;
  } else {
    // positions of the loaded values (wall neighbours clamp to the cell)
    const int jm = j > 0 ? j - 1 : j, jp = j + 1 < L.ny ? j + 1 : j;
    const int km = k > 0 ? k - 1 : k, kp = k + 1 < L.nz ? k + 1 : k;
    const int il = i0 > 0 ? i0 - 1 : i0, ir = i0 + 4 < L.nx ? i0 + 4 : i0 + 3;
    if (mrow) {
      const unsigned a = *reinterpret_cast<const unsigned*>(mrow + c);
      const unsigned ym4 = *reinterpret_cast<const unsigned*>(mrow + (c + (jm - j) * sy));
      const unsigned yp4 = *reinterpret_cast<const unsigned*>(mrow + (c + (jp - j) * sy));
      const unsigned zm4 = *reinterpret_cast<const unsigned*>(mrow + (c + (km - k) * sz));
      const unsigned zp4 = *reinterpret_cast<const unsigned*>(mrow + (c + (kp - k) * sz));
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        mc[u] = (int)((a >> (8 * u)) & 63u);
        Q.c[u] = s1[mc[u]] * Q.c[u];
        Q.ym[u] = s1[(ym4 >> (8 * u)) & 63u] * Q.ym[u];
        Q.yp[u] = s1[(yp4 >> (8 * u)) & 63u] * Q.yp[u];
        Q.zm[u] = s1[(zm4 >> (8 * u)) & 63u] * Q.zm[u];
        Q.zp[u] = s1[(zp4 >> (8 * u)) & 63u] * Q.zp[u];
        t2[u] = s2[mc[u]];
      }
      xl = s1[mrow[c + (il - i0)] & 63] * Q.xm[0];
      xr = s1[mrow[c + (ir - i0)] & 63] * Q.xp[3];
    } else {
      // the box rule per row (quad_box_mask); the x-edge neighbours il = i0 - 1
      // and ir = i0 + 4 are interior in x unless they clamp onto cells 0 / 3
      const int yc = box_yz(L.ny, L.nz, j, k), yym = box_yz(L.ny, L.nz, jm, k), yyp = box_yz(L.ny, L.nz, jp, k);
      const int yzm = box_yz(L.ny, L.nz, j, km), yzp = box_yz(L.ny, L.nz, j, kp);
      int m0 = 63, m3 = 63;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int mcu = quad_box_mask(L.nx, i0, yc, u);
        if (u == 0) m0 = mcu;
        if (u == 3) m3 = mcu;
        Q.c[u] = s1[mcu] * Q.c[u];
        Q.ym[u] = s1[quad_box_mask(L.nx, i0, yym, u)] * Q.ym[u];
        Q.yp[u] = s1[quad_box_mask(L.nx, i0, yyp, u)] * Q.yp[u];
        Q.zm[u] = s1[quad_box_mask(L.nx, i0, yzm, u)] * Q.zm[u];
        Q.zp[u] = s1[quad_box_mask(L.nx, i0, yzp, u)] * Q.zp[u];
        t2[u] = s2[mcu];
      }
      xl = s1[i0 > 0 ? (yc | 3) : m0] * Q.xm[0];
      xr = s1[i0 + 4 < L.nx ? (yc | 3) : m3] * Q.xp[3];
    }
  }
  Q.xm[0] = xl;
  Q.xm[1] = Q.c[0];
  Q.xm[2] = Q.c[1];
  Q.xm[3] = Q.c[2];
  Q.xp[0] = Q.c[1];
  Q.xp[1] = Q.c[2];
  Q.xp[2] = Q.c[3];
  Q.xp[3] = xr;
  float xn[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (mrow) close_faces(Q, u, mc[u]);
    const float ax = quad_apply_f(Q, u, w0, w1, w2);
    xn[u] = Q.c[u] + t2[u] * (bb[u] - ax);
  }
  store4(xout, c, xn);
  END_QUADS
}

// Restriction (as k_mg_restrict): each lane owns two coarse cells
// (2 I0, 2 I0 + 1) of coarse row (J, K), i.e. one four-cell chunk in each of
// the four fine rows (2J + dy, 2K + dz).  Sums keep the (dz, dy, dx) order.
template <bool kFrac>
__device__ __forceinline__ void
b_mg_restrict_q(VBlock vb, const MgLevel& F, const MgLevel& Cl, const float* __restrict__ x, const float* __restrict__ b,
                float* __restrict__ bc, float* __restrict__ xc, double tau, const PcgControl* ctl) {
  if (ctl->done) return;
  const float w0 = (float)F.w[0], w1 = (float)F.w[1], w2 = (float)F.w[2];
  const int nz_ = Cl.nz;
  // coarse pairs: Cl.nx = F.nx / 2 is even, two coarse cells per lane
  FOR_QUADS(2 * Cl.nx, Cl.ny, Cl.div_h, Cl.div_y)
  (void)c;
  const int J = j, K = k, I0 = i0 >> 1;
  float s0 = 0.0f, s1 = 0.0f;
  // dz rolled: the two rows of one dz in flight at a time, so the pass runs
  // at higher occupancy
#pragma unroll 1
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
      const int fj = 2 * J + dy, fk = 2 * K + dz;
      const int fc = i0 + F.nx * (fj + F.ny * fk);
      Quad<float> Q;
      load_quad(x, F.nx, F.ny, F.nz, i0, fj, fk, fc, Q);
      float bb[4];
      load4(b, fc, bb);
      float r[4];
      if (kFrac) {  // fine level with face fractions
        FracRow R;
        load_frac_row(F, i0, fj, fk, fc, R);
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] = bb[u] - quad_apply_frac(F, Q, u, frac6_of(R, u, i0 + u + 1 < F.nx));
      } else {
        residual_quad(F, Q, bb, fc, fj + F.ny * fk, w0, w1, w2, r);
      }
      s0 += r[0];
      s0 += r[1];
      s1 += r[2];
      s1 += r[3];
    }
  const int cc = I0 + Cl.nx * (J + Cl.ny * K);
  *reinterpret_cast<float2*>(bc + cc) = make_float2(s0, s1);
  const float o0 = lv_scale<float>(Cl, tau, cc, I0, J, K), o1 = lv_scale<float>(Cl, tau, cc + 1, I0 + 1, J, K);
  *reinterpret_cast<float2*>(xc + cc) = make_float2(o0 * s0, o1 * s1);
  END_QUADS
}

// Prolongation (as k_mg_prolong) for four fine cells: parents I0 = i0/2 and
// I0 + 1 and their x-neighbours I0 - 1 and I0 + 2 (mirrored at walls).
// out = xin + P xc for cells i0..i0+3 of row (j, k) (cell c, row index row).
__device__ __forceinline__ void prolong_quad(const MgLevel& F, const MgLevel& Cl, const float* __restrict__ xc,
                                             const float* __restrict__ xin, int i0, int j, int k, int c, int row,
                                             float out[4]) {
  const int sy = Cl.nx, sz = Cl.nx * Cl.ny;
  const float a = 0.75f, bq = 0.25f;
  const int I0 = i0 >> 1, J = j >> 1, K = k >> 1;
  int Jn = J + ((j & 1) ? 1 : -1), Kn = K + ((k & 1) ? 1 : -1);
  if (Jn < 0 || Jn >= Cl.ny) Jn = J;
  if (Kn < 0 || Kn >= Cl.nz) Kn = K;
  const int Im = I0 > 0 ? I0 - 1 : I0;              // x-neighbour of the even child
  const int Ip = I0 + 2 < Cl.nx ? I0 + 2 : I0 + 1;  // x-neighbour of the last odd child
  // every load is issued up front (all in bounds); the renormalised
  // obstacle form replaces the result only next to inactive coarse cells
  float xi[4];
  load4(xin, c, xi);
  float v[4][4];
  const int rows[4] = {sy * J + sz * K, sy * Jn + sz * K, sy * J + sz * Kn, sy * Jn + sz * Kn};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float2 mid = *reinterpret_cast<const float2*>(xc + I0 + rows[r]);
    v[r][0] = xc[Im + rows[r]];
    v[r][1] = mid.x;
    v[r][2] = mid.y;
    v[r][3] = xc[Ip + rows[r]];
  }
  // cells needing the renormalised obstacle form: fluid fine cells whose
  // octant stencil reaches an inactive coarse cell
  unsigned slow = 0;
  if (Cl.frac) {
    const unsigned oc = *reinterpret_cast<const unsigned short*>(Cl.octant + I0 + rows[0]);
    // fluid flags of the four fine cells (level-0 mask, fraction words, or all)
    unsigned fl = 0xFu;
    const unsigned char* const mrow = row_mask(F.mask, F.nx, F.ny, F.nz, row, kRowPlain);
    if (mrow) {
      const unsigned m4 = *reinterpret_cast<const unsigned*>(mrow + c);
      fl = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) fl |= ((m4 >> (8 * u)) & kFluidBit) ? 1u << u : 0u;
    } else if (!F.mask && F.frac) {
      const uint4 f4 = *reinterpret_cast<const uint4*>(F.frac + c);
      fl = ((f4.x & kFracFluid) ? 1u : 0u) | ((f4.y & kFracFluid) ? 2u : 0u) | ((f4.z & kFracFluid) ? 4u : 0u) |
           ((f4.w & kFracFluid) ? 8u : 0u);
    }
    const int oyz = ((j & 1) << 1) | ((k & 1) << 2);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const unsigned bits = (oc >> (8 * (u >> 1))) & 0xffu;
      const bool fluid = (fl >> u) & 1u;
      if (fluid && !((bits >> ((u & 1) | oyz)) & 1u)) slow |= 1u << u;
    }
  }
  // fine cell u: parent index P = 1 + (u >> 1), neighbour N = u even ? P-1 : P+1
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int P = 1 + (u >> 1), N = (u & 1) ? P + 1 : P - 1;
    const float v00 = a * v[0][P] + bq * v[0][N];
    const float v10 = a * v[1][P] + bq * v[1][N];
    const float v01 = a * v[2][P] + bq * v[2][N];
    const float v11 = a * v[3][P] + bq * v[3][N];
    out[u] = xi[u] + (a * (a * v00 + bq * v10) + bq * (a * v01 + bq * v11));
  }
  if (slow) {
    // renormalised form (prolong_value's order: dz, dy, dx) from the values
    // already loaded; one batch of activity-flag loads
    bool act[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint2 mid = *reinterpret_cast<const uint2*>(Cl.frac + I0 + rows[r]);
      act[r][0] = (Cl.frac[Im + rows[r]] & kFracFluid) != 0;
      act[r][1] = (mid.x & kFracFluid) != 0;
      act[r][2] = (mid.y & kFracFluid) != 0;
      act[r][3] = (Cl.frac[Ip + rows[r]] & kFracFluid) != 0;
    }
    const float wt[2] = {a, bq};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!((slow >> u) & 1)) continue;
      const int P = 1 + (u >> 1), N = (u & 1) ? P + 1 : P - 1;
      const int xs[2] = {P, N};
      float num = 0.0f, den = 0.0f;
#pragma unroll
      for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
          for (int dx = 0; dx < 2; ++dx) {
            const int r = dy + 2 * dz;
            const float w = wt[dx] * wt[dy] * wt[dz];
            const bool on = act[r][xs[dx]];
            num += on ? w * v[r][xs[dx]] : 0.0f;
            den += on ? w : 0.0f;
          }
      out[u] = xi[u] + (den > 0.0f ? num / den : 0.0f);
    }
  }
}

__device__ __forceinline__ void
b_mg_prolong_q(VBlock vb, const MgLevel& F, const MgLevel& Cl, const float* __restrict__ xc, const float* __restrict__ xin,
               float* __restrict__ xout, const PcgControl* ctl) {
  if (ctl->done) return;
  const int nz_ = F.nz;
  FOR_QUADS(F.nx, F.ny, F.div_q, F.div_y)
  float out[4];
  prolong_quad(F, Cl, xc, xin, i0, j, k, c, row, out);
  store4(xout, c, out);
  END_QUADS
}

// ------------------------------------------------------- kernel wrappers

template <bool kZeroP>
__global__ void __launch_bounds__(kBlock)
k_mg_outer_residual_q(DevGrid g, MgLevel L0, const double* __restrict__ p, const double* __restrict__ b,
                      float* __restrict__ b0, float* __restrict__ x0, double tau, const PcgControl* ctl,
                      DD* pdd, double* pmx) {
  b_mg_outer_residual_q<kZeroP>(hw_block(), g, L0, p, b, b0, x0, tau, ctl, pdd, pmx);
}
__global__ void __launch_bounds__(kBlock)
k_residual_report_q(DevGrid g, const double* __restrict__ p, const double* __restrict__ b, DD* pdd, double* pmx) {
  const MgLevel unused{};
  b_mg_outer_residual_q<false, true>(hw_block(), g, unused, p, b, nullptr, nullptr, 0.0, nullptr, pdd, pmx);
}
template <bool kUpdate, bool kZeroP>
__global__ void __launch_bounds__(kBlock)
k_mg_smooth_q(MgLevel L, const float* __restrict__ xin, const float* __restrict__ b,
              float* __restrict__ xout, double* __restrict__ p, double tau, const PcgControl* ctl) {
  b_mg_smooth_q<kUpdate, kZeroP>(hw_block(), L, xin, b, xout, p, tau, ctl);
}
__global__ void __launch_bounds__(kBlock)
k_mg_smooth_qf(MgLevel L, const float* __restrict__ xin, const float* __restrict__ b, float* __restrict__ xout,
               double tau, const PcgControl* ctl) {
  b_mg_smooth_qf(hw_block(), L, xin, b, xout, tau, ctl);
}
__global__ void __launch_bounds__(kBlock)
k_mg_smooth_first_q(MgLevel L, const float* __restrict__ b, float* __restrict__ xout, double tau1, double tau2,
                    const PcgControl* ctl, MgFold fold) {
  if (fold.fin.ctl && blockIdx.x == 0) {
    // the previous outer residual's acceptance; this block then reads the
    // control block after the decision (the barrier orders it), the others
    // read it as it stands when they start
    finish_body<1, 2, FinMgTrue, 4>(fold.pdd, fold.pmx, fold.nb, fold.fin);
    __syncthreads();
  }
  b_mg_smooth_first_q(hw_block(), L, b, xout, tau1, tau2, ctl);
}
template <bool kFrac>
__global__ void __launch_bounds__(kBlock)
k_mg_restrict_q(MgLevel F, MgLevel Cl, const float* __restrict__ x, const float* __restrict__ b,
                float* __restrict__ bc, float* __restrict__ xc, double tau, const PcgControl* ctl) {
  b_mg_restrict_q<kFrac>(hw_block(), F, Cl, x, b, bc, xc, tau, ctl);
}
__global__ void __launch_bounds__(kBlock)
k_mg_prolong_q(MgLevel F, MgLevel Cl, const float* __restrict__ xc, const float* __restrict__ xin,
               float* __restrict__ xout, const PcgControl* ctl) {
  b_mg_prolong_q(hw_block(), F, Cl, xc, xin, xout, ctl);
}

// Batched: block -> scene, scene-local block, the scene's argument record.
#define RH_BATCH_PROLOGUE(ARGS)                                  \
  const int s_ = batch_scene(prefix, n, blockIdx.x);            \
  const ARGS& a = A[s_];                                        \
  const VBlock vb = batch_block(prefix, s_);

__global__ void __launch_bounds__(kBlock)
kb_mg_outer_residual_q(const BOuterArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BOuterArgs)
  b_mg_outer_residual_q(vb, *a.g, *a.L0, a.p, a.b, a.b0, a.x0, a.tau, a.ctl, a.pdd, a.pmx);
}
template <bool kUpdate>
__global__ void __launch_bounds__(kBlock)
kb_mg_smooth_q(const BSmoothArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BSmoothArgs)
  b_mg_smooth_q<kUpdate>(vb, *a.L, a.xin, a.b, a.xout, a.p, a.tau, a.ctl);
}
__global__ void __launch_bounds__(kBlock)
kb_mg_smooth_qf(const BSmoothArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BSmoothArgs)
  b_mg_smooth_qf(vb, *a.L, a.xin, a.b, a.xout, a.tau, a.ctl);
}
__global__ void __launch_bounds__(kBlock)
kb_mg_smooth_first_q(const BSmoothArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BSmoothArgs)
  b_mg_smooth_first_q(vb, *a.L, a.b, a.xout, a.tau_first, a.tau, a.ctl);
}
template <bool kFrac>
__global__ void __launch_bounds__(kBlock)
kb_mg_restrict_q(const BRestrictArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BRestrictArgs)
  b_mg_restrict_q<kFrac>(vb, *a.F, *a.C, a.x, a.b, a.bc, a.xc, a.tau, a.ctl);
}
__global__ void __launch_bounds__(kBlock)
kb_mg_prolong_q(const BProlongArgs* __restrict__ A, const int* __restrict__ prefix, int n) {
  RH_BATCH_PROLOGUE(BProlongArgs)
  b_mg_prolong_q(vb, *a.F, *a.C, a.xc, a.xin, a.xout, a.ctl);
}

void launch_b_mg_outer_q(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_outer_residual_q<<<bl.blocks, block_shape(), 0, s>>>((const BOuterArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_smooth_q(const BatchLaunch& bl, bool update, hipStream_t s) {
  if (!bl.blocks) return;
  if (update)
    kb_mg_smooth_q<true><<<bl.blocks, block_shape(), 0, s>>>((const BSmoothArgs*)bl.args, bl.prefix, bl.n);
  else
    kb_mg_smooth_q<false><<<bl.blocks, block_shape(), 0, s>>>((const BSmoothArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_smooth_qf(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_smooth_qf<<<bl.blocks, block_shape(), 0, s>>>((const BSmoothArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_smooth_first_q(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_smooth_first_q<<<bl.blocks, block_shape(), 0, s>>>((const BSmoothArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_restrict_q(const BatchLaunch& bl, bool frac, hipStream_t s) {
  if (!bl.blocks) return;
  if (frac)
    kb_mg_restrict_q<true><<<bl.blocks, block_shape(), 0, s>>>((const BRestrictArgs*)bl.args, bl.prefix, bl.n);
  else
    kb_mg_restrict_q<false><<<bl.blocks, block_shape(), 0, s>>>((const BRestrictArgs*)bl.args, bl.prefix, bl.n);
}
void launch_b_mg_prolong_q(const BatchLaunch& bl, hipStream_t s) {
  if (bl.blocks)
    kb_mg_prolong_q<<<bl.blocks, block_shape(), 0, s>>>((const BProlongArgs*)bl.args, bl.prefix, bl.n);
}

// ------------------------------------------------------------- launchers

bool mg_quad_level(const MgLevel& L) { return L.nx % 4 == 0 && !L.frac; }
bool mg_quad_frac_level(const MgLevel& L) { return L.nx % 4 == 0 && L.frac; }
bool mg_quad_prolong_pair(const MgLevel& F, const MgLevel& C) {
  return F.nx % 4 == 0 && F.nx == 2 * C.nx && F.ny == 2 * C.ny && F.nz == 2 * C.nz;
}
bool mg_quad_pair(const MgLevel& F, const MgLevel& C) { return mg_quad_prolong_pair(F, C); }

static inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

bool mg_quad_outer_ok(const DevGrid& g, const MgLevel& L0, const double* p, const double* b) {
  return mg_quad_level(L0) && g.nx == L0.nx && aligned16(p) && aligned16(b);
}

int mg_quad_outer_blocks(const DevGrid& g) { return quad_blocks(g.nx / 4, g.ny * g.nz); }

void launch_mg_outer_residual_q(const DevGrid& g, const MgLevel& L0, const double* p, const double* b,
                                float* b0, float* x0, double tau, const PcgControl* ctl, DD* pdd,
                                double* pmx, hipStream_t s, bool p_zero) {
  if (p_zero)
    launch(k_mg_outer_residual_q<true>, dim3(mg_quad_outer_blocks(g)), block_shape(), 0, s, g, L0, p, b, b0, x0,
           tau, ctl, pdd, pmx);
  else
    launch(k_mg_outer_residual_q<false>, dim3(mg_quad_outer_blocks(g)), block_shape(), 0, s, g, L0, p, b, b0, x0,
           tau, ctl, pdd, pmx);
}

void launch_residual_report_q(const DevGrid& g, const double* p, const double* b, const Reduce& red,
                              hipStream_t s) {
  const int nb = mg_quad_outer_blocks(g);
  launch(k_residual_report_q, dim3(nb), block_shape(), 0, s, g, p, b, red.pdd, red.pmx);
  launch_finish<1, 2>(red.pdd, red.pmx, nb, FinStore{red.out, 2, 0.0, 0.0, 0.0}, s);
}

void launch_mg_smooth_q(const MgLevel& L, const float* xin, const float* b, float* xout, double* p,
                        double tau, const PcgControl* ctl, hipStream_t s, bool p_zero) {
  const dim3 grid(quad_blocks(L.nx / 4, L.ny * L.nz));
  if (p && p_zero)
    launch(k_mg_smooth_q<true, true>, grid, block_shape(), 0, s, L, xin, b, xout, p, tau, ctl);
  else if (p)
    launch(k_mg_smooth_q<true, false>, grid, block_shape(), 0, s, L, xin, b, xout, p, tau, ctl);
  else
    launch(k_mg_smooth_q<false, false>, grid, block_shape(), 0, s, L, xin, b, xout, p, tau, ctl);
}

void launch_mg_smooth_qf(const MgLevel& L, const float* xin, const float* b, float* xout, double tau,
                         const PcgControl* ctl, hipStream_t s) {
  launch(k_mg_smooth_qf, dim3(quad_blocks(L.nx / 4, L.ny * L.nz)), block_shape(), 0, s, L, xin, b, xout, tau,
         ctl);
}

void launch_mg_smooth_first_q(const MgLevel& L, const float* b, float* xout, double tau1, double tau2,
                              const PcgControl* ctl, hipStream_t s, const MgFold* fold) {
  MgFold f{};
  if (fold) f = *fold;
  launch(k_mg_smooth_first_q, dim3(quad_blocks(L.nx / 4, L.ny * L.nz)), block_shape(), 0, s, L, b, xout, tau1,
         tau2, ctl, f);
}

bool mg_first_from_b(const DevGrid& g, const MgHierarchy& h, const double* p, const double* b) {
  return h.quad && h.pre >= 2 && h.levels > 1 && mg_quad_outer_ok(g, h.lv[0], p, b);
}

void launch_mg_restrict_q(const MgLevel& F, const MgLevel& Cl, const float* x, const float* b, float* bc,
                          float* xc, double tau, const PcgControl* ctl, hipStream_t s) {
  const dim3 grid(quad_blocks(Cl.nx / 2, Cl.ny * Cl.nz));
  if (F.frac)
    launch(k_mg_restrict_q<true>, grid, block_shape(), 0, s, F, Cl, x, b, bc, xc, tau, ctl);
  else
    launch(k_mg_restrict_q<false>, grid, block_shape(), 0, s, F, Cl, x, b, bc, xc, tau, ctl);
}

void launch_mg_prolong_q(const MgLevel& F, const MgLevel& Cl, const float* xc, const float* xin,
                         float* xout, const PcgControl* ctl, hipStream_t s) {
  launch(k_mg_prolong_q, dim3(quad_blocks(F.nx / 4, F.ny * F.nz)), block_shape(), 0, s, F, Cl, xc, xin, xout,
         ctl);
}

}  // namespace rh
