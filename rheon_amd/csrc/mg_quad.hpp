// Four-cells-per-lane building blocks shared by the multigrid kernels
// (mg_quad_kernels.hip) and the LDS-resident V-cycle tail (mg_kernels.hip):
// row-chunk loads, the six-neighbour quad, and the box / fraction stencils
// written with the one-cell kernels' expressions (so both give the same bits).
#pragma once
#include "common.hpp"
#include "mg.hpp"
#include "mg_stencil.hpp"
#include "stencil.hpp"

namespace rh {

// Chunk loop over this block's rows: chunk t covers cells 4q..4q+3 of one
// row; consecutive threads take consecutive chunks (coalesced).  DQ and DY
// divide by the chunks per row ((NX) / 4) and by NY.
#define FOR_QUADS(NX, NY, DQ, DY)                                           \
  {                                                                         \
    const int cpr_ = (NX) >> 2;                                             \
    const RowRange qr_ = block_span(cpr_ * (NY) * nz_, vb);                 \
    for (int t_ = qr_.begin + flat_tid(); t_ < qr_.end; t_ += kBlock) {     \
      const int row = fast_div(t_, DQ), q = t_ - row * cpr_;                \
      const int k = fast_div(row, DY), j = row - k * (NY);                  \
      const int i0 = q << 2;                                                \
      const int c = i0 + (NX) * row;
#define END_QUADS \
  }               \
  }

__device__ __forceinline__ void load4(const float* __restrict__ p, int c, float v[4]) {
  const float4 t = *reinterpret_cast<const float4*>(p + c);
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
__device__ __forceinline__ void load4(const double* __restrict__ p, int c, double v[4]) {
  const double2 a = *reinterpret_cast<const double2*>(p + c);
  const double2 b = *reinterpret_cast<const double2*>(p + c + 2);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void store4(float* __restrict__ p, int c, const float v[4]) {
  *reinterpret_cast<float4*>(p + c) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void store4(double* __restrict__ p, int c, const double v[4]) {
  *reinterpret_cast<double2*>(p + c) = make_double2(v[0], v[1]);
  *reinterpret_cast<double2*>(p + c + 2) = make_double2(v[2], v[3]);
}

// Four cells and their six neighbours; a missing wall neighbour is the cell
// itself (box clamping), a closed obstacle face is resolved by `close`.
template <typename T>
struct Quad {
  T c[4], xm[4], xp[4], ym[4], yp[4], zm[4], zp[4];
};

template <typename T>
__device__ __forceinline__ void load_quad(const T* __restrict__ x, int nx, int ny, int nz, int i0,
                                          int j, int k, int c, Quad<T>& Q) {
  const int sy = nx, sz = nx * ny;
  load4(x, c, Q.c);
  load4(x, j > 0 ? c - sy : c, Q.ym);
  load4(x, j + 1 < ny ? c + sy : c, Q.yp);
  load4(x, k > 0 ? c - sz : c, Q.zm);
  load4(x, k + 1 < nz ? c + sz : c, Q.zp);
  const T left = x[i0 > 0 ? c - 1 : c];
  const T right = x[i0 + 4 < nx ? c + 4 : c + 3];
  Q.xm[0] = left;
  Q.xm[1] = Q.c[0];
  Q.xm[2] = Q.c[1];
  Q.xm[3] = Q.c[2];
  Q.xp[0] = Q.c[1];
  Q.xp[1] = Q.c[2];
  Q.xp[2] = Q.c[3];
  Q.xp[3] = right;
}

// box_mask6's y/z bits of row (j, k).
__device__ __forceinline__ int box_yz(int ny, int nz, int j, int k) {
  return ((j > 0) << 2) | ((j + 1 < ny) << 3) | ((k > 0) << 4) | ((k + 1 < nz) << 5);
}

// box_mask6 of cell i0 + u of a four-cell chunk (nx % 4 == 0, i0 % 4 == 0)
// in a row with y/z bits yz: cells 1 and 2 have both x faces open; only cell
// 0 (x- at i0 = 0) and cell 3 (x+ at i0 + 4 = nx) can touch an x wall.
__device__ __forceinline__ int quad_box_mask(int nx, int i0, int yz, int u) {
  const int xb = u == 0 ? ((i0 > 0) | 2) : (u == 3 ? (1 | ((i0 + 4 < nx) << 1)) : 3);
  return yz | xb;
}

// Open-face bits of cell i0+u: obstacle mask byte or the box rule (yz =
// box_yz of the chunk's row).
__device__ __forceinline__ int quad_mask(const unsigned char* mask, unsigned m4, int nx, int i0, int yz, int u) {
  return mask ? (int)((m4 >> (8 * u)) & 63u) : quad_box_mask(nx, i0, yz, u);
}

// Row flags of the level-0 obstacle mask, stored after it (one byte per
// (j, k) row at mask + nx ny nz): kRowPlain = every cell of the row is fluid
// with the plain box's open faces; kRowPlainNear = kRowPlain holds for the
// row and its four neighbour rows (j +- 1, k +- 1, clamped at walls).  A row
// with the flag takes the box form of a stencil (no mask load), which gives
// the same values: its masks are box_mask6, and at walls the box form's
// clamped neighbour is the cell itself, as close_faces makes it.
constexpr unsigned kRowPlain = 1u, kRowPlainNear = 2u;
__device__ __forceinline__ const unsigned char* row_mask(const unsigned char* mask, int nx, int ny, int nz,
                                                         int row, unsigned flag) {
  if (!mask) return nullptr;
  return (mask[(long long)nx * ny * nz + row] & flag) ? nullptr : mask;
}

// Closed faces take the cell's own value (contribution exactly 0).
template <typename T>
__device__ __forceinline__ void close_faces(Quad<T>& Q, int u, int m) {
  const T xc = Q.c[u];
  Q.xm[u] = (m & 1) ? Q.xm[u] : xc;
  Q.xp[u] = (m & 2) ? Q.xp[u] : xc;
  Q.ym[u] = (m & 4) ? Q.ym[u] : xc;
  Q.yp[u] = (m & 8) ? Q.yp[u] : xc;
  Q.zm[u] = (m & 16) ? Q.zm[u] : xc;
  Q.zp[u] = (m & 32) ? Q.zp[u] : xc;
}

// s_apply in S = float for cell u of the quad (same expression).
__device__ __forceinline__ float quad_apply_f(const Quad<float>& Q, int u, float w0, float w1, float w2) {
  const float xc = Q.c[u];
  return w0 * ((xc - Q.xm[u]) + (xc - Q.xp[u])) + w1 * ((xc - Q.ym[u]) + (xc - Q.yp[u])) +
         w2 * ((xc - Q.zm[u]) + (xc - Q.zp[u]));
}

// apply_at in f64 for cell u (the reference operator's order).
__device__ __forceinline__ double quad_apply_d(const Quad<double>& Q, int u, const DevGrid& g) {
  const double xc = Q.c[u];
  double v = 0.0;
  v += g.w[0] * (xc - Q.xm[u]);
  v += g.w[0] * (xc - Q.xp[u]);
  v += g.w[1] * (xc - Q.ym[u]);
  v += g.w[1] * (xc - Q.yp[u]);
  v += g.w[2] * (xc - Q.zm[u]);
  v += g.w[2] * (xc - Q.zp[u]);
  return v;
}

// Fraction words of a four-cell row chunk: the cells' own (lower faces),
// the next cell's along x, and the rows above in y and z (upper faces).
struct FracRow {
  unsigned fw[5];  // cells 0..3 and the cell after the chunk
  unsigned fy[4], fz[4];
  bool hy, hz;
};
// `fr` = the level's fraction words (global memory or a copy in LDS).
__device__ __forceinline__ void load_frac_row(const unsigned* fr, const MgLevel& L, int i0, int j, int k, int c,
                                              FracRow& R) {
  const int sy = L.nx, sz = L.nx * L.ny;
  R.hy = j + 1 < L.ny;
  R.hz = k + 1 < L.nz;
  const uint4 q4 = *reinterpret_cast<const uint4*>(fr + c);
  const uint4 y4 = *reinterpret_cast<const uint4*>(fr + (R.hy ? c + sy : c));
  const uint4 z4 = *reinterpret_cast<const uint4*>(fr + (R.hz ? c + sz : c));
  R.fw[0] = q4.x;
  R.fw[1] = q4.y;
  R.fw[2] = q4.z;
  R.fw[3] = q4.w;
  R.fw[4] = fr[i0 + 4 < L.nx ? c + 4 : c + 3];
  R.fy[0] = y4.x;
  R.fy[1] = y4.y;
  R.fy[2] = y4.z;
  R.fy[3] = y4.w;
  R.fz[0] = z4.x;
  R.fz[1] = z4.y;
  R.fz[2] = z4.z;
  R.fz[3] = z4.w;
}
__device__ __forceinline__ void load_frac_row(const MgLevel& L, int i0, int j, int k, int c, FracRow& R) {
  load_frac_row(L.frac, L, i0, j, k, c, R);
}
// lv_frac6 of cell u of the chunk.
__device__ __forceinline__ Frac6 frac6_of(const FracRow& R, int u, bool hx) {
  Frac6 f;
  f.xm = (int)(R.fw[u] & 1023u);
  f.ym = (int)((R.fw[u] >> 10) & 1023u);
  f.zm = (int)((R.fw[u] >> 20) & 1023u);
  f.xp = hx ? (int)(R.fw[u + 1] & 1023u) : 0;
  f.yp = R.hy ? (int)((R.fy[u] >> 10) & 1023u) : 0;
  f.zp = R.hz ? (int)((R.fz[u] >> 20) & 1023u) : 0;
  return f;
}
// s_apply's fraction form for cell u of the quad.
__device__ __forceinline__ float quad_apply_frac(const MgLevel& L, const Quad<float>& Q, int u, const Frac6& f) {
  const float xc = Q.c[u];
  return (float)L.wf[0] * ((float)f.xm * (xc - Q.xm[u]) + (float)f.xp * (xc - Q.xp[u])) +
         (float)L.wf[1] * ((float)f.ym * (xc - Q.ym[u]) + (float)f.yp * (xc - Q.yp[u])) +
         (float)L.wf[2] * ((float)f.zm * (xc - Q.zm[u]) + (float)f.zp * (xc - Q.zp[u]));
}

}  // namespace rh
