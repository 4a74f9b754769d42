// Level-local stencil helpers of the multigrid kernels (box, level-0
// obstacle mask, coarse face fractions), shared by the scalar and the
// four-cells-per-lane kernels.
#pragma once
#include "common.hpp"
#include "mg.hpp"
#include "stencil.hpp"

namespace rh {

// Active (fluid-holding) cell of a level; every cell of the plain box.
__device__ __forceinline__ bool lv_active(const MgLevel& L, int c) {
  if (L.frac) return (L.frac[c] & kFracFluid) != 0;
  if (L.mask) return (L.mask[c] & kFluidBit) != 0;
  return true;
}

// Open fractions (in units of fscale) of the six faces of cell c on a level
// with face fractions `fr` (the level's `frac` array, or a copy of it);
// upper faces are the neighbours' lower faces, walls are 0.
struct Frac6 {
  int xm, xp, ym, yp, zm, zp;
};
__device__ __forceinline__ Frac6 lv_frac6(const MgLevel& L, const unsigned* fr, int c, int i, int j,
                                          int k) {
  const int sy = L.nx, sz = L.nx * L.ny;
  const bool hx = i + 1 < L.nx, hy = j + 1 < L.ny, hz = k + 1 < L.nz;
  const unsigned q = fr[c];
  const unsigned qx = fr[hx ? c + 1 : c];
  const unsigned qy = fr[hy ? c + sy : c];
  const unsigned qz = fr[hz ? c + sz : c];
  Frac6 f;
  f.xm = (int)(q & 1023u);
  f.ym = (int)((q >> 10) & 1023u);
  f.zm = (int)((q >> 20) & 1023u);
  f.xp = hx ? (int)(qx & 1023u) : 0;
  f.yp = hy ? (int)((qy >> 10) & 1023u) : 0;
  f.zp = hz ? (int)((qz >> 20) & 1023u) : 0;
  return f;
}

__device__ __forceinline__ int lv_mask(const MgLevel& L, int c, int i, int j, int k) {
  return L.mask ? (L.mask[c] & 63) : box_mask6(L.nx, L.ny, L.nz, i, j, k);
}

// Level-local stencil in the V-cycle precision S (arithmetic in S).
// Branch-free: all six neighbours are loaded at once from clamped
// addresses (a wall or closed-face neighbour is the cell itself and
// contributes 0).  `fr` = the level's face fractions (null for box/mask
// levels).
template <typename S>
__device__ __forceinline__ S s_apply_fr(const MgLevel& L, const unsigned* fr, const S* __restrict__ x, int c,
                                        int i, int j, int k) {
  if (fr) {
    const Frac6 f = lv_frac6(L, fr, c, i, j, k);
    const Nbr n = nbr_offsets(L.nx, L.ny, L.nz, i, j, k);
    const S xc = x[c];
    const S a = x[c + n.xm], b = x[c + n.xp], cy = x[c + n.ym], d = x[c + n.yp],
            e = x[c + n.zm], ff = x[c + n.zp];
    return (S)L.wf[0] * ((S)f.xm * (xc - a) + (S)f.xp * (xc - b)) +
           (S)L.wf[1] * ((S)f.ym * (xc - cy) + (S)f.yp * (xc - d)) +
           (S)L.wf[2] * ((S)f.zm * (xc - e) + (S)f.zp * (xc - ff));
  }
  // box-clamped loads first (independent of the mask), then closed faces
  // select x_c (exact 0 contribution)
  const Nbr n = nbr_offsets(L.nx, L.ny, L.nz, i, j, k);
  const S xc = x[c];
  S a = x[c + n.xm], b = x[c + n.xp], cy = x[c + n.ym], d = x[c + n.yp], e = x[c + n.zm],
    f = x[c + n.zp];
  if (L.mask) {
    const int m = L.mask[c];
    a = (m & 1) ? a : xc;
    b = (m & 2) ? b : xc;
    cy = (m & 4) ? cy : xc;
    d = (m & 8) ? d : xc;
    e = (m & 16) ? e : xc;
    f = (m & 32) ? f : xc;
  }
  const S w0 = (S)L.w[0], w1 = (S)L.w[1], w2 = (S)L.w[2];
  return w0 * ((xc - a) + (xc - b)) + w1 * ((xc - cy) + (xc - d)) + w2 * ((xc - e) + (xc - f));
}
template <typename S>
__device__ __forceinline__ S s_apply(const MgLevel& L, const S* __restrict__ x, int c, int i, int j,
                                     int k) {
  return s_apply_fr<S>(L, L.frac, x, c, i, j, k);
}

// 1 / D (Jacobi scale) of the cell; 0 where D = 0 (solid).
__device__ __forceinline__ double lv_odiag_fr(const MgLevel& L, const unsigned* fr, int c, int i, int j,
                                              int k) {
  if (fr) {
    const Frac6 f = lv_frac6(L, fr, c, i, j, k);
    const double d = L.wf[0] * (double)(f.xm + f.xp) + L.wf[1] * (double)(f.ym + f.yp) +
                     L.wf[2] * (double)(f.zm + f.zp);
    return d > 0.0 ? 1.0 / d : 0.0;
  }
  return L.odiag[lv_mask(L, c, i, j, k)];
}
__device__ __forceinline__ double lv_odiag(const MgLevel& L, int c, int i, int j, int k) {
  return lv_odiag_fr(L, L.frac, c, i, j, k);
}

// Sweep scale tau / D in the V-cycle precision.
template <typename S>
__device__ __forceinline__ S lv_scale(const MgLevel& L, double tau, int c, int i, int j, int k) {
  return (S)(tau * lv_odiag(L, c, i, j, k));
}

// Fine cell c of level F holds fluid (always in the plain box).
__device__ __forceinline__ bool lv_fine_fluid(const MgLevel& F, int c) { return lv_active(F, c); }

// Trilinear prolongation value of fine cell (i, j, k) from the coarse
// iterate (cell-centred weights 27/64, 9/64, 3/64, 1/64; a missing
// neighbour at a wall mirrors the parent).  With obstacles, for a fluid
// fine cell whose stencil reaches an inactive coarse cell, the weights of
// inactive cells are dropped and the rest renormalised (a solid fine
// cell's value is never read by a fluid equation, so it keeps the plain
// form).
template <typename S>
__device__ __forceinline__ S prolong_value_fr(const MgLevel& Cl, const unsigned* cfr, const unsigned char* coct,
                                              const S* __restrict__ xc, int i, int j, int k, bool fine_fluid) {
  const int I = i >> 1, J = j >> 1, K = k >> 1;
  int In = I + ((i & 1) ? 1 : -1), Jn = J + ((j & 1) ? 1 : -1), Kn = K + ((k & 1) ? 1 : -1);
  if (In < 0 || In >= Cl.nx) In = I;
  if (Jn < 0 || Jn >= Cl.ny) Jn = J;
  if (Kn < 0 || Kn >= Cl.nz) Kn = K;
  const int sy = Cl.nx, sz = Cl.nx * Cl.ny;
  const S a = (S)0.75, bq = (S)0.25;
  // box form with every load issued up front; with obstacles it stands
  // unless an inactive coarse cell is near the parent
  const S v00 = a * xc[I + sy * J + sz * K] + bq * xc[In + sy * J + sz * K];
  const S v10 = a * xc[I + sy * Jn + sz * K] + bq * xc[In + sy * Jn + sz * K];
  const S v01 = a * xc[I + sy * J + sz * Kn] + bq * xc[In + sy * J + sz * Kn];
  const S v11 = a * xc[I + sy * Jn + sz * Kn] + bq * xc[In + sy * Jn + sz * Kn];
  const S box = a * (a * v00 + bq * v10) + bq * (a * v01 + bq * v11);
  if (!cfr || !fine_fluid) return box;
  const int o = (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
  if ((coct[I + sy * J + sz * K] >> o) & 1) return box;
  const int ci[2] = {I, In}, cj[2] = {J, Jn}, ck[2] = {K, Kn};
  const S wt[2] = {a, bq};
  S num = (S)0, den = (S)0;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int idx = ci[dx] + sy * cj[dy] + sz * ck[dz];
        const S w = wt[dx] * wt[dy] * wt[dz];
        const bool act = (cfr[idx] & kFracFluid) != 0;
        num += act ? w * xc[idx] : (S)0;
        den += act ? w : (S)0;
      }
  return den > (S)0 ? num / den : (S)0;
}
template <typename S>
__device__ __forceinline__ S prolong_value(const MgLevel& Cl, const S* __restrict__ xc, int i, int j, int k,
                                           bool fine_fluid) {
  return prolong_value_fr<S>(Cl, Cl.frac, Cl.octant, xc, i, j, k, fine_fluid);
}

#define FOR_LEVEL_CELLS(L) FOR_LEVEL_CELLS_VB(L, hw_block())
// Flat partition of a level's cells: consecutive threads take consecutive
// cells (every lane busy however short the rows of a coarse level are).
#define FOR_LEVEL_CELLS_VB(L, VB)                                                    \
  {                                                                                  \
    const RowRange cr_ = block_span((L).nx * (L).ny * (L).nz, (VB));                 \
    for (int c = cr_.begin + flat_tid(); c < cr_.end; c += kBlock) {                 \
      {                                                                              \
        const int row = c / (L).nx, i = c - row * (L).nx;                            \
        const int k = row / (L).ny, j = row - k * (L).ny;


}  // namespace rh
