/*
 * rheon_oracle.c -- CPU restatement of Rheon's fixed-box MAC projection step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Build with
 * -ffp-contract=off so every multiply and add rounds exactly as in the
 * reference.  The OpenMP paths mirror src/parallel.rs: element-wise passes are
 * split into 2048-value chunks once a pass has >= 16 Ki values and more than
 * one thread is requested; reductions always run serially on the caller in
 * index order.  Results are bit-identical for every thread count.
 */
#include "rheon_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MIN_PARALLEL_VALUES (16 * 1024) /* src/parallel.rs:11 */
#define CHUNK_VALUES 2048               /* src/parallel.rs:12 */

/* Rust f64::clamp (panics if min > max; callers guarantee min <= max). */
static inline double clampd(double x, double lo, double hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
/* Rust f64::max / f64::min: IEEE maxNum/minNum (NaN-ignoring). */
static inline double maxd(double a, double b) { return fmax(a, b); }
static inline double mind(double a, double b) { return fmin(a, b); }

static int use_parallel(int64_t len, int threads) {
  return threads > 1 && len >= MIN_PARALLEL_VALUES;
}

/* ------------------------------------------------------------------------ */
/* Geometry: src/geometry.rs:114-191, :235-258, :284-315                      */
/* ------------------------------------------------------------------------ */

int or_grid_init(or_grid* g, const uint64_t counts[3], const double spacing[3],
                 const double origin[3]) {
  memset(g, 0, sizeof(*g));
  for (int a = 0; a < 3; ++a) {
    if (counts[a] == 0) return OR_GEOMETRY;
    if (counts[a] > (uint64_t)INT64_MAX) return OR_GEOMETRY;
    if (!isfinite(spacing[a]) || spacing[a] <= 0.0) return OR_GEOMETRY;
    if (!isfinite(origin[a])) return OR_GEOMETRY;
  }
  /* product with overflow checks (geometry.rs:76-81) */
  __int128 cells = (__int128)counts[0] * counts[1];
  if (cells > INT64_MAX) return OR_GEOMETRY;
  cells *= counts[2];
  if (cells > INT64_MAX) return OR_GEOMETRY;
  for (int a = 0; a < 3; ++a) g->n[a] = (int64_t)counts[a];
  g->cell_len = (int64_t)cells;
  for (int a = 0; a < 3; ++a) {
    __int128 f = 1;
    for (int d = 0; d < 3; ++d) {
      g->fn[a][d] = g->n[d] + (d == a ? 1 : 0);
      f *= g->fn[a][d];
      if (f > INT64_MAX) return OR_GEOMETRY;
    }
    g->face_len[a] = (int64_t)f;
  }
  g->volume = spacing[0] * spacing[1] * spacing[2];
  if (!isfinite(g->volume) || g->volume <= 0.0) return OR_GEOMETRY;
  for (int a = 0; a < 3; ++a) {
    if (counts[a] > (1ULL << 52)) return OR_GEOMETRY;
    double n = (double)counts[a];
    double extent = n * spacing[a];
    double upper = origin[a] + extent;
    double scale = maxd(maxd(fabs(origin[a]), fabs(upper)), extent);
    double half = 0.5 * spacing[a];
    int half_normal = isnormal(half);
    if (!isfinite(extent) || !isfinite(scale) || !half_normal ||
        spacing[a] / scale < 32.0 * DBL_EPSILON)
      return OR_GEOMETRY;
    double first_center = origin[a] + 0.5 * spacing[a];
    double last_center = origin[a] + (n - 0.5) * spacing[a];
    if (!isfinite(upper) || upper <= origin[a] || first_center <= origin[a] ||
        last_center >= upper || (counts[a] > 1 && origin[a] + spacing[a] <= origin[a]))
      return OR_GEOMETRY;
    g->h[a] = spacing[a];
    g->origin[a] = origin[a];
    g->upper[a] = upper;
  }
  return OR_OK;
}

int64_t or_buffer_plan(const or_grid* g, int64_t* velocity_pair, int64_t* tracer_pair,
                       int64_t* pressure_workspace) {
  int64_t faces = g->face_len[0] + g->face_len[1] + g->face_len[2];
  *velocity_pair = faces * 8;
  *tracer_pair = g->cell_len * 8;
  *pressure_workspace = g->cell_len * 48;
  return *velocity_pair + *tracer_pair + *pressure_workspace;
}

void or_face_position(const or_grid* g, int axis, const int64_t p[3], double out[3]) {
  for (int d = 0; d < 3; ++d) {
    double offset = d == axis ? 0.0 : 0.5;
    out[d] = g->origin[d] + ((double)p[d] + offset) * g->h[d];
  }
}

void or_cell_position(const or_grid* g, const int64_t p[3], double out[3]) {
  for (int d = 0; d < 3; ++d) out[d] = g->origin[d] + ((double)p[d] + 0.5) * g->h[d];
}

/* Sampling support of one component: face lattice or cell lattice
 * (src/sampling.rs:31-51). */
static void support(const or_grid* g, int component, int64_t dims[3], double off[3]) {
  for (int d = 0; d < 3; ++d) {
    if (component < 3) {
      dims[d] = g->fn[component][d];
      off[d] = d == component ? 0.0 : 0.5;
    } else {
      dims[d] = g->n[d];
      off[d] = 0.5;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Sampling and advection: src/sampling.rs:74-175, src/advection.rs:28-102    */
/* ------------------------------------------------------------------------ */

#define DEFINE_SAMPLING(T, SUF)                                                            \
  int or_sample_##SUF(const or_grid* g, int component, const T* values, const double pos[3], \
                      double* out) {                                                        \
    int64_t dims[3], lo[3], hi[3];                                                          \
    double off[3], t[3];                                                                    \
    support(g, component, dims, off);                                                       \
    for (int d = 0; d < 3; ++d)                                                             \
      if (!isfinite(pos[d])) return OR_NONFINITE_POSITION;                                  \
    for (int d = 0; d < 3; ++d) {                                                           \
      double x = clampd(pos[d], g->origin[d], g->upper[d]);                                 \
      double q = clampd((x - g->origin[d]) / g->h[d] - off[d], 0.0, (double)(dims[d] - 1)); \
      if (!isfinite(q)) return OR_ARITHMETIC_FAILURE;                                       \
      lo[d] = (int64_t)floor(q);                                                            \
      hi[d] = lo[d] + 1 < dims[d] - 1 ? lo[d] + 1 : dims[d] - 1;                            \
      t[d] = q - (double)lo[d];                                                             \
    }                                                                                       \
    double result = 0.0, minimum = INFINITY, maximum = -INFINITY;                           \
    for (int z = 0; z < 2; ++z)                                                             \
      for (int y = 0; y < 2; ++y)                                                           \
        for (int x = 0; x < 2; ++x) {                                                       \
          int bits[3] = {x, y, z};                                                          \
          int64_t p[3];                                                                     \
          double weight = 1.0;                                                              \
          for (int d = 0; d < 3; ++d) {                                                     \
            p[d] = bits[d] == 0 ? lo[d] : hi[d];                                            \
            weight = weight * (bits[d] == 0 ? 1.0 - t[d] : t[d]);                           \
          }                                                                                 \
          double value = (double)values[p[0] + dims[0] * (p[1] + dims[1] * p[2])];          \
          minimum = mind(minimum, value);                                                   \
          maximum = maxd(maximum, value);                                                   \
          result += weight * value;                                                         \
        }                                                                                   \
    if (!isfinite(result)) return OR_ARITHMETIC_FAILURE;                                    \
    *out = clampd(result, minimum, maximum);                                                \
    return OR_OK;                                                                           \
  }                                                                                         \
  static int departure_##SUF(const or_grid* g, const double pos[3], const double vel[3],    \
                             double dt, double out[3]) {                                    \
    for (int d = 0; d < 3; ++d) {                                                           \
      double value = pos[d] - dt * vel[d];                                                  \
      if (!isfinite(value)) return OR_ARITHMETIC_FAILURE;                                   \
      out[d] = clampd(value, g->origin[d], g->upper[d]);                                    \
    }                                                                                       \
    return OR_OK;                                                                           \
  }                                                                                         \
  static int sample_velocity_##SUF(const or_grid* g, const T* u, const T* v, const T* w,   \
                                   const double pos[3], double out[3]) {                    \
    int s;                                                                                  \
    if ((s = or_sample_##SUF(g, 0, u, pos, &out[0]))) return s;                             \
    if ((s = or_sample_##SUF(g, 1, v, pos, &out[1]))) return s;                             \
    return or_sample_##SUF(g, 2, w, pos, &out[2]);                                          \
  }                                                                                         \
  int or_backtrace_##SUF(const or_grid* g, const T* u, const T* v, const T* w,              \
                         const double pos[3], double dt, double out[3]) {                   \
    if (!isfinite(dt) || dt <= 0.0) return OR_INVALID_TIME_STEP;                            \
    double initial[3], midpoint[3], velocity[3];                                            \
    int s;                                                                                  \
    if ((s = sample_velocity_##SUF(g, u, v, w, pos, initial))) return s;                    \
    if ((s = departure_##SUF(g, pos, initial, 0.5 * dt, midpoint))) return s;               \
    if ((s = sample_velocity_##SUF(g, u, v, w, midpoint, velocity))) return s;              \
    return departure_##SUF(g, pos, velocity, dt, out);                                      \
  }                                                                                         \
  static int finite_all_##SUF(const T* x, int64_t n) {                                      \
    for (int64_t i = 0; i < n; ++i)                                                         \
      if (!isfinite((double)x[i])) return 0;                                                \
    return 1;                                                                               \
  }

DEFINE_SAMPLING(float, f32)
DEFINE_SAMPLING(double, f64)

#define DEFINE_ADVECT(T, SUF)                                                                \
  int or_advect_velocity_##SUF(const or_grid* g, const T* u, const T* v, const T* w,        \
                               double dt, T* ou, T* ov, T* ow, int threads) {                \
    if (!isfinite(dt) || dt <= 0.0) return OR_INVALID_TIME_STEP;                            \
    if (!finite_all_##SUF(u, g->face_len[0]) || !finite_all_##SUF(v, g->face_len[1]) ||     \
        !finite_all_##SUF(w, g->face_len[2]))                                               \
      return OR_NONFINITE_INPUT;                                                            \
    const T* old[3] = {u, v, w};                                                            \
    T* out[3] = {ou, ov, ow};                                                               \
    for (int d = 0; d < 3; ++d) {                                                           \
      const int64_t* dims = g->fn[d];                                                       \
      int64_t len = g->face_len[d], plane = dims[0] * dims[1];                              \
      int status = OR_OK;                                                                   \
      int par = use_parallel(len, threads);                                                 \
      int64_t chunks = (len + 1023) / 1024;                                                 \
      _Pragma("omp parallel for schedule(dynamic, 1) num_threads(threads) if(par)")          \
      for (int64_t c = 0; c < chunks; ++c) {                                                \
        int64_t end = (c + 1) * 1024 < len ? (c + 1) * 1024 : len;                          \
        for (int64_t row = c * 1024; row < end; ++row) {                                    \
          int64_t p[3] = {row % dims[0], (row / dims[0]) % dims[1], row / plane};           \
          if (p[d] == 0 || p[d] + 1 == dims[d]) {                                           \
            out[d][row] = (T)0.0;                                                           \
            continue;                                                                       \
          }                                                                                 \
          double pos[3], dep[3], value;                                                     \
          or_face_position(g, d, p, pos);                                                   \
          int s = or_backtrace_##SUF(g, u, v, w, pos, dt, dep);                             \
          if (!s) s = or_sample_##SUF(g, d, old[d], dep, &value);                           \
          if (s) {                                                                          \
            _Pragma("omp atomic write") status = s;                                         \
            continue;                                                                       \
          }                                                                                 \
          out[d][row] = (T)value;                                                           \
        }                                                                                   \
      }                                                                                     \
      if (status) return status;                                                            \
    }                                                                                       \
    return OR_OK;                                                                           \
  }

DEFINE_ADVECT(float, f32)
DEFINE_ADVECT(double, f64)

int or_advect_tracer_f32(const or_grid* g, const float* old, const float* u, const float* v,
                         const float* w, double dt, float* out, int threads) {
  if (!isfinite(dt) || dt <= 0.0) return OR_INVALID_TIME_STEP;
  if (!finite_all_f32(old, g->cell_len) || !finite_all_f32(u, g->face_len[0]) ||
      !finite_all_f32(v, g->face_len[1]) || !finite_all_f32(w, g->face_len[2]))
    return OR_NONFINITE_INPUT;
  int64_t len = g->cell_len, plane = g->n[0] * g->n[1];
  int status = OR_OK;
  int par = use_parallel(len, threads);
  int64_t chunks = (len + 1023) / 1024;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (par)
  for (int64_t c = 0; c < chunks; ++c) {
    int64_t end = (c + 1) * 1024 < len ? (c + 1) * 1024 : len;
    for (int64_t row = c * 1024; row < end; ++row) {
      int64_t p[3] = {row % g->n[0], (row / g->n[0]) % g->n[1], row / plane};
      double pos[3], dep[3], value;
      or_cell_position(g, p, pos);
      int s = or_backtrace_f32(g, u, v, w, pos, dt, dep);
      if (!s) s = or_sample_f32(g, 3, old, dep, &value);
      if (s) {
#pragma omp atomic write
        status = s;
        continue;
      }
      out[row] = (float)value;
    }
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* Operator: src/operator.rs                                                  */
/* ------------------------------------------------------------------------ */

/* PressureOperator::new coefficients (src/operator.rs:58-97). */
/* Optional obstacle: per-cell open-face mask (bits x-,x+,y-,y+,z-,z+ and
 * 64 = fluid) and the eliminated gauge cell (0 in the plain box). */
typedef struct op_coeff {
  double area[3], acc[3], w[3], diag[64];
  const uint8_t* mask;
  int64_t gauge;
  /* several fluid components: one gauge per component (mask bit 128 on each
   * component's gauge cell), labels[row] = component or -1 (solid) */
  const int32_t* labels;
  int32_t components;
} op_coeff;

#define OR_GAUGE_BIT 128

/* The eliminated rows: the gauge cell, and with an obstacle every cell that
 * carries the gauge bit (one per connected fluid region). */
static inline int is_gauge(const op_coeff* c, int64_t row) {
  return row == c->gauge || (c->mask && (c->mask[row] & OR_GAUGE_BIT));
}

/* r = z = 0 on the eliminated rows (src/pressure.rs:534-541; per component,
 * src/obstacle_pressure.rs:385-388, :415-425). */
static void zero_gauge_rows(const or_grid* g, const op_coeff* c, double* r, double* z) {
  if (!c->mask) {
    r[c->gauge] = 0.0;
    z[c->gauge] = 0.0;
    return;
  }
  for (int64_t row = 0; row < g->cell_len; ++row)
    if (is_gauge(c, row)) r[row] = z[row] = 0.0;
}

static int op_init(const or_grid* g, double density, op_coeff* c) {
  c->mask = NULL;
  c->gauge = 0;
  c->labels = NULL;
  c->components = 1;
  if (!isfinite(density) || density <= 0.0) return OR_INVALID_DENSITY;
  const double* h = g->h;
  c->area[0] = h[1] * h[2];
  c->area[1] = h[0] * h[2];
  c->area[2] = h[0] * h[1];
  for (int d = 0; d < 3; ++d) {
    c->acc[d] = 1.0 / (density * h[d]);
    c->w[d] = c->area[d] * c->acc[d];
    if (!isfinite(c->area[d]) || c->area[d] <= 0.0 || !isfinite(c->acc[d]) ||
        c->acc[d] <= 0.0 || !isfinite(c->w[d]) || c->w[d] <= 0.0)
      return OR_INVALID_ARGUMENT;
  }
  for (int mask = 0; mask < 64; ++mask) {
    double value = 0.0;
    for (int side = 0; side < 6; ++side)
      if (mask & (1 << side)) value += c->w[side / 2];
    c->diag[mask] = value;
  }
  return OR_OK;
}

static inline int neighbor_mask(const or_grid* g, const int64_t p[3]) {
  int mask = 0;
  for (int d = 0; d < 3; ++d) {
    mask |= (p[d] > 0) << (2 * d);
    mask |= (p[d] + 1 < g->n[d]) << (2 * d + 1);
  }
  return mask;
}

/* Open-face bits of a row: the neighbour bits in the plain box, the mask
 * with an obstacle (a face is open when both cells are fluid). */
static inline int row_mask(const or_grid* g, const op_coeff* c, int64_t row, const int64_t p[3]) {
  return c->mask ? (c->mask[row] & 63) : neighbor_mask(g, p);
}

double or_diagonal(const or_grid* g, double density, const int64_t p[3]) {
  op_coeff c;
  if (op_init(g, density, &c)) return NAN;
  return c.diag[neighbor_mask(g, p)];
}

#define DEFINE_VALIDATE(T, SUF)                                                         \
  static int validate_velocity_##SUF(const or_grid* g, const T* const vel[3]) {        \
    for (int d = 0; d < 3; ++d) {                                                      \
      if (!finite_all_##SUF(vel[d], g->face_len[d])) return OR_NONFINITE_INPUT;        \
      const int64_t* dims = g->fn[d];                                                  \
      for (int64_t k = 0; k < dims[2]; ++k)                                            \
        for (int64_t j = 0; j < dims[1]; ++j)                                          \
          for (int64_t i = 0; i < dims[0]; ++i) {                                      \
            int64_t p[3] = {i, j, k};                                                  \
            if (p[d] == 0 || p[d] == g->n[d]) {                                        \
              if (vel[d][i + dims[0] * (j + dims[1] * k)] != 0.0)                      \
                return OR_NONZERO_WALL_VELOCITY;                                       \
            }                                                                          \
          }                                                                            \
    }                                                                                  \
    return OR_OK;                                                                      \
  }

DEFINE_VALIDATE(float, f32)
DEFINE_VALIDATE(double, f64)

static inline int64_t face_index(const or_grid* g, int axis, int64_t i, int64_t j, int64_t k) {
  return i + g->fn[axis][0] * (j + g->fn[axis][1] * k);
}

#define DEFINE_RHS(T, SUF)                                                                    \
  int or_build_rhs_##SUF(const or_grid* g, const T* u, const T* v, const T* w, double dt,    \
                         double* out) {                                                       \
    if (!isfinite(dt) || dt <= 0.0) return OR_INVALID_TIME_STEP;                             \
    const T* vel[3] = {u, v, w};                                                             \
    int s = validate_velocity_##SUF(g, vel);                                                 \
    if (s) return s;                                                                         \
    double area[3] = {g->h[1] * g->h[2], g->h[0] * g->h[2], g->h[0] * g->h[1]};              \
    for (int64_t k = 0; k < g->n[2]; ++k)                                                    \
      for (int64_t j = 0; j < g->n[1]; ++j)                                                  \
        for (int64_t i = 0; i < g->n[0]; ++i) {                                              \
          int64_t p[3] = {i, j, k};                                                          \
          double flux = 0.0;                                                                 \
          for (int d = 0; d < 3; ++d) {                                                      \
            int64_t q[3] = {i, j, k};                                                        \
            q[d] += 1;                                                                       \
            int64_t lo = face_index(g, d, p[0], p[1], p[2]);                                 \
            int64_t hi = face_index(g, d, q[0], q[1], q[2]);                                 \
            flux += area[d] * ((double)vel[d][lo] - (double)vel[d][hi]);                     \
          }                                                                                  \
          double value = flux / dt;                                                          \
          if (!isfinite(value)) return OR_ARITHMETIC_FAILURE;                                \
          out[i + g->n[0] * (j + g->n[1] * k)] = value;                                      \
        }                                                                                    \
    return OR_OK;                                                                            \
  }

DEFINE_RHS(float, f32)
DEFINE_RHS(double, f64)

static int finite_cells(const double* x, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(x[i])) return 0;
  return 1;
}

/* apply_full box fast path body (src/operator.rs:400-416). */
static int apply_range(const or_grid* g, const op_coeff* c, const double* x, double* y,
                       int64_t begin, int64_t end) {
  const int64_t nx = g->n[0], ny = g->n[1];
  const int64_t stride[3] = {1, nx, nx * ny};
  for (int64_t row = begin; row < end; ++row) {
    int64_t p[3] = {row % nx, (row / nx) % ny, row / (nx * ny)};
    const int m = row_mask(g, c, row, p);
    double value = 0.0;
    for (int d = 0; d < 3; ++d) {
      if (m & (1 << (2 * d))) value += c->w[d] * (x[row] - x[row - stride[d]]);
      if (m & (2 << (2 * d))) value += c->w[d] * (x[row] - x[row + stride[d]]);
    }
    if (!isfinite(value)) return OR_ARITHMETIC_FAILURE;
    y[row] = value;
  }
  return OR_OK;
}

static int apply_full_c(const or_grid* g, const op_coeff* c, const double* x, double* y,
                        int threads) {
  int64_t n = g->cell_len;
  if (!use_parallel(n, threads)) return apply_range(g, c, x, y, 0, n);
  int status = OR_OK;
  int64_t chunks = (n + CHUNK_VALUES - 1) / CHUNK_VALUES;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int64_t ch = 0; ch < chunks; ++ch) {
    int64_t e = (ch + 1) * CHUNK_VALUES < n ? (ch + 1) * CHUNK_VALUES : n;
    int s = apply_range(g, c, x, y, ch * CHUNK_VALUES, e);
    if (s) {
#pragma omp atomic write
      status = s;
    }
  }
  return status;
}

int or_apply_full(const or_grid* g, double density, const double* x, double* y, int threads) {
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  if (!finite_cells(x, g->cell_len)) return OR_NONFINITE_INPUT;
  return apply_full_c(g, &c, x, y, threads);
}

int or_divergence_f32(const or_grid* g, const float* u, const float* v, const float* w,
                      double* out) {
  const float* vel[3] = {u, v, w};
  int s = validate_velocity_f32(g, vel);
  if (s) return s;
  for (int64_t k = 0; k < g->n[2]; ++k)
    for (int64_t j = 0; j < g->n[1]; ++j)
      for (int64_t i = 0; i < g->n[0]; ++i) {
        double value = ((double)u[face_index(g, 0, i + 1, j, k)] -
                        (double)u[face_index(g, 0, i, j, k)]) /
                           g->h[0] +
                       ((double)v[face_index(g, 1, i, j + 1, k)] -
                        (double)v[face_index(g, 1, i, j, k)]) /
                           g->h[1] +
                       ((double)w[face_index(g, 2, i, j, k + 1)] -
                        (double)w[face_index(g, 2, i, j, k)]) /
                           g->h[2];
        if (!isfinite(value)) return OR_ARITHMETIC_FAILURE;
        out[i + g->n[0] * (j + g->n[1] * k)] = value;
      }
  return OR_OK;
}

int or_divergence_f64(const or_grid* g, const double* u, const double* v, const double* w,
                      double* out) {
  const double* vel[3] = {u, v, w};
  int s = validate_velocity_f64(g, vel);
  if (s) return s;
  for (int64_t k = 0; k < g->n[2]; ++k)
    for (int64_t j = 0; j < g->n[1]; ++j)
      for (int64_t i = 0; i < g->n[0]; ++i) {
        int64_t p[3] = {i, j, k};
        double value = 0.0;
        for (int d = 0; d < 3; ++d) {
          int64_t q[3] = {i, j, k};
          q[d] += 1;
          value += (vel[d][face_index(g, d, q[0], q[1], q[2])] -
                    vel[d][face_index(g, d, p[0], p[1], p[2])]) /
                   g->h[d];
        }
        if (!isfinite(value)) return OR_ARITHMETIC_FAILURE;
        out[i + g->n[0] * (j + g->n[1] * k)] = value;
      }
  return OR_OK;
}

int or_correct_f32(const or_grid* g, double density, const double* p, double dt, float* u,
                   float* v, float* w) {
  if (!isfinite(dt) || dt <= 0.0) return OR_INVALID_TIME_STEP;
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  float* vel[3] = {u, v, w};
  const float* cvel[3] = {u, v, w};
  if ((s = validate_velocity_f32(g, cvel))) return s;
  if (!finite_cells(p, g->cell_len)) return OR_NONFINITE_INPUT;
  const int64_t stride[3] = {1, g->n[0], g->n[0] * g->n[1]};
  for (int d = 0; d < 3; ++d) {
    const int64_t* dims = g->fn[d];
    for (int64_t k = 0; k < dims[2]; ++k)
      for (int64_t j = 0; j < dims[1]; ++j)
        for (int64_t i = 0; i < dims[0]; ++i) {
          int64_t q[3] = {i, j, k};
          if (q[d] == 0 || q[d] == g->n[d]) continue;
          int64_t index = face_index(g, d, i, j, k);
          int64_t right = i + g->n[0] * (j + g->n[1] * k);
          int64_t left = right - stride[d];
          double difference = p[right] - p[left];
          double value = (double)vel[d][index] - dt * c.acc[d] * difference;
          if (!isfinite(value) || fabs(value) > (double)FLT_MAX) return OR_ARITHMETIC_FAILURE;
          vel[d][index] = (float)value;
        }
  }
  return OR_OK;
}

int or_correct_f64(const or_grid* g, double density, const double* p, double dt, double* u,
                   double* v, double* w) {
  if (!isfinite(dt) || dt <= 0.0) return OR_INVALID_TIME_STEP;
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  double* vel[3] = {u, v, w};
  const double* cvel[3] = {u, v, w};
  if ((s = validate_velocity_f64(g, cvel))) return s;
  if (!finite_cells(p, g->cell_len)) return OR_NONFINITE_INPUT;
  const int64_t stride[3] = {1, g->n[0], g->n[0] * g->n[1]};
  for (int d = 0; d < 3; ++d) {
    const int64_t* dims = g->fn[d];
    for (int64_t k = 0; k < dims[2]; ++k)
      for (int64_t j = 0; j < dims[1]; ++j)
        for (int64_t i = 0; i < dims[0]; ++i) {
          int64_t q[3] = {i, j, k};
          if (q[d] == 0 || q[d] == g->n[d]) continue;
          int64_t index = face_index(g, d, i, j, k);
          int64_t right = i + g->n[0] * (j + g->n[1] * k);
          int64_t left = right - stride[d];
          /* correction(): acceleration_factor * (p_R - p_L) (src/operator.rs:357-358) */
          double gradient = c.acc[d] * (p[right] - p[left]);
          double value = vel[d][index] - dt * gradient;
          if (!isfinite(value)) return OR_ARITHMETIC_FAILURE;
          vel[d][index] = value;
        }
  }
  return OR_OK;
}

int or_actual_divergence_max_f32(const or_grid* g, const float* u, const float* v,
                                 const float* w, double* out) {
  double* div = malloc(sizeof(double) * g->cell_len);
  if (!div) return OR_INVALID_ARGUMENT;
  int s = or_divergence_f32(g, u, v, w, div);
  double m = 0.0;
  if (!s)
    for (int64_t i = 0; i < g->cell_len; ++i) m = maxd(m, fabs(div[i]));
  free(div);
  *out = m;
  return s;
}

int or_actual_divergence_max_f64(const or_grid* g, const double* u, const double* v,
                                 const double* w, double* out) {
  double* div = malloc(sizeof(double) * g->cell_len);
  if (!div) return OR_INVALID_ARGUMENT;
  int s = or_divergence_f64(g, u, v, w, div);
  double m = 0.0;
  if (!s)
    for (int64_t i = 0; i < g->cell_len; ++i) m = maxd(m, fabs(div[i]));
  free(div);
  *out = m;
  return s;
}

#define DEFINE_RATE(T, SUF)                                                                   \
  double or_velocity_rate_##SUF(const or_grid* g, const T* u, const T* v, const T* w) {      \
    const T* vel[3] = {u, v, w};                                                             \
    double rate = 0.0;                                                                       \
    for (int d = 0; d < 3; ++d) {                                                            \
      double m = 0.0;                                                                        \
      for (int64_t i = 0; i < g->face_len[d]; ++i) m = maxd(m, fabs((double)vel[d][i]));     \
      rate += m / g->h[d];                                                                   \
    }                                                                                        \
    return rate;                                                                             \
  }

DEFINE_RATE(float, f32)
DEFINE_RATE(double, f64)

/* ------------------------------------------------------------------------ */
/* Reductions and preconditioners: src/pressure.rs:135-161, :465-575          */
/* ------------------------------------------------------------------------ */

static inline void neumaier_add(double* sum, double* correction, double value) {
  double next = *sum + value;
  *correction += fabs(*sum) >= fabs(value) ? (*sum - next) + value : (value - next) + *sum;
  *sum = next;
}

double or_compensated_sum(const double* x, int64_t n) {
  double sum = 0.0, corr = 0.0;
  for (int64_t i = 0; i < n; ++i) neumaier_add(&sum, &corr, x[i]);
  return sum + corr;
}

double or_dot(const double* a, const double* b, int64_t n) {
  double sum = 0.0, corr = 0.0;
  for (int64_t i = 0; i < n; ++i) neumaier_add(&sum, &corr, a[i] * b[i]);
  return sum + corr;
}

static int jacobi_c(const or_grid* g, const op_coeff* c, double* r, double* z, double* rz,
                    int threads) {
  int64_t n = g->cell_len;
  zero_gauge_rows(g, c, r, z);
  const int64_t nx = g->n[0], ny = g->n[1];
  int status = OR_OK;
  int64_t chunks = (n + CHUNK_VALUES - 1) / CHUNK_VALUES;
  int par = use_parallel(n, threads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (par)
  for (int64_t ch = 0; ch < chunks; ++ch) {
    int64_t e = (ch + 1) * CHUNK_VALUES < n ? (ch + 1) * CHUNK_VALUES : n;
    for (int64_t row = ch * CHUNK_VALUES; row < e; ++row) {
      if (is_gauge(c, row)) continue;
      if (c->mask && !(c->mask[row] & 64)) {  // solid cell: not an unknown
        z[row] = 0.0;
        continue;
      }
      int64_t p[3] = {row % nx, (row / nx) % ny, row / (nx * ny)};
      double diagonal = c->diag[row_mask(g, c, row, p)];
      if (!isfinite(diagonal) || diagonal <= 0.0) {
#pragma omp atomic write
        status = OR_BREAKDOWN;
        continue;
      }
      z[row] = r[row] / diagonal;
    }
  }
  if (status) return status;
  *rz = or_dot(r, z, n);
  return OR_OK;
}

int or_precondition_jacobi(const or_grid* g, double density, double* r, double* z,
                           double* rz, int threads) {
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  return jacobi_c(g, &c, r, z, rz, threads);
}

static int sgs_c(const or_grid* g, const op_coeff* c, double* r, double* z, double* rz) {
  r[0] = 0.0;
  z[0] = 0.0;
  const int64_t nx = g->n[0], ny = g->n[1], n = g->cell_len;
  const int64_t stride[3] = {1, nx, nx * ny};
  for (int64_t row = 1; row < n; ++row) {
    int64_t p[3] = {row % nx, (row / nx) % ny, row / (nx * ny)};
    double diagonal = c->diag[neighbor_mask(g, p)];
    double value = r[row];
    for (int d = 0; d < 3; ++d)
      if (p[d] > 0) value += c->w[d] * z[row - stride[d]];
    z[row] = value / diagonal;
    if (!isfinite(z[row]) || diagonal <= 0.0 || !isfinite(diagonal)) return OR_BREAKDOWN;
  }
  for (int64_t row = n - 1; row >= 1; --row) {
    int64_t p[3] = {row % nx, (row / nx) % ny, row / (nx * ny)};
    double diagonal = c->diag[neighbor_mask(g, p)];
    double correction = 0.0;
    for (int d = 0; d < 3; ++d)
      if (p[d] + 1 < g->n[d]) correction += c->w[d] * z[row + stride[d]];
    z[row] = z[row] + correction / diagonal;
    if (!isfinite(z[row])) return OR_BREAKDOWN;
  }
  *rz = or_dot(r, z, n);
  return OR_OK;
}

int or_precondition_sgs(const or_grid* g, double density, double* r, double* z, double* rz) {
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  return sgs_c(g, &c, r, z, rz);
}

/* Red-black symmetric Gauss-Seidel (rheon_amd's own preconditioner, not in
 * the reference).  Colour (i+j+k)&1, red = 0.  M = (D+L)D^-1(D+U) in R,B
 * order on the gauge-eliminated system (row 0 is red and is skipped):
 *   red:   y = r / D
 *   black: y = (r + sum_nbr w*y) / D          (neighbours are red)
 *   black reverse: z = y                      (no later unknowns)
 *   red reverse:   z = y + (sum_nbr w*z) / D  (neighbours are black)
 * Neighbour order: axis 0..2, lower then upper. */
static int rbsgs_c(const or_grid* g, const op_coeff* c, double* r, double* z, double* rz) {
  const int64_t nx = g->n[0], ny = g->n[1], n = g->cell_len;
  zero_gauge_rows(g, c, r, z);
  const int64_t stride[3] = {1, nx, nx * ny};
  for (int pass = 0; pass < 3; ++pass) {
    int colour = pass == 1 ? 1 : 0;
    for (int64_t row = 0; row < n; ++row) {
      if (is_gauge(c, row)) continue;
      int64_t p[3] = {row % nx, (row / nx) % ny, row / (nx * ny)};
      if (((p[0] + p[1] + p[2]) & 1) != colour) continue;
      if (c->mask && !(c->mask[row] & 64)) {  // solid cell: not an unknown
        z[row] = 0.0;
        continue;
      }
      const int m = row_mask(g, c, row, p);
      double diagonal = c->diag[m];
      if (pass == 0) {
        z[row] = r[row] / diagonal;
      } else if (pass == 1) {
        double value = r[row];
        for (int d = 0; d < 3; ++d) {
          if (m & (1 << (2 * d))) value += c->w[d] * z[row - stride[d]];
          if (m & (2 << (2 * d))) value += c->w[d] * z[row + stride[d]];
        }
        z[row] = value / diagonal;
      } else {
        double correction = 0.0;
        for (int d = 0; d < 3; ++d) {
          if (m & (1 << (2 * d))) correction += c->w[d] * z[row - stride[d]];
          if (m & (2 << (2 * d))) correction += c->w[d] * z[row + stride[d]];
        }
        z[row] = z[row] + correction / diagonal;
      }
      if (!isfinite(z[row]) || !isfinite(diagonal) || diagonal <= 0.0) return OR_BREAKDOWN;
    }
  }
  *rz = or_dot(r, z, n);
  return OR_OK;
}

int or_precondition_rbsgs(const or_grid* g, double density, double* r, double* z, double* rz) {
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  return rbsgs_c(g, &c, r, z, rz);
}

/* ------------------------------------------------------------------------ */
/* PCG: src/pressure.rs:432-456 (true_report), :577-740 (solve_loaded)        */
/* ------------------------------------------------------------------------ */

typedef struct ws {
  double *r, *d, *q, *z;
} ws;

static int true_report(const or_grid* g, const op_coeff* c, const double* rhs,
                       const double* p, double* product, double dt, int64_t iterations,
                       int threads, or_report* out) {
  int s = apply_full_c(g, c, p, product, threads);
  if (s) return s;
  int64_t n = g->cell_len;
  int par = use_parallel(n, threads);
  int64_t chunks = (n + CHUNK_VALUES - 1) / CHUNK_VALUES;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (par)
  for (int64_t ch = 0; ch < chunks; ++ch) {
    int64_t e = (ch + 1) * CHUNK_VALUES < n ? (ch + 1) * CHUNK_VALUES : n;
    for (int64_t i = ch * CHUNK_VALUES; i < e; ++i) product[i] = rhs[i] - product[i];
  }
  double squared = or_dot(product, product, n);
  double maximum = 0.0;
  for (int64_t i = 0; i < n; ++i) maximum = maxd(maximum, fabs(product[i]));
  double divergence = maximum * dt / g->volume;
  if (!isfinite(squared) || !isfinite(divergence)) {
    out->error_iteration = iterations;
    return OR_BREAKDOWN;
  }
  out->iterations = iterations;
  out->true_residual_l2 = sqrt(squared);
  out->true_residual_max = maximum;
  out->predicted_divergence_max = divergence;
  return OR_OK;
}

static int precondition(const or_grid* g, const op_coeff* c, int method, double* r, double* z,
                        double* rz, int threads) {
  switch (method) {
    case OR_JACOBI: return jacobi_c(g, c, r, z, rz, threads);
    case OR_SGS: return sgs_c(g, c, r, z, rz);
    case OR_RBSGS: return rbsgs_c(g, c, r, z, rz);
    default: return OR_INVALID_ARGUMENT;
  }
}

static int solve_c(const or_grid* g, const op_coeff* c, const double* rhs, const double* initial, double dt,
                   const or_settings* st, int method, int threads, double* p, or_report* rep) {
  const int64_t n = g->cell_len;
  if (!finite_cells(rhs, n)) return OR_NONFINITE_INPUT;
  if (!isfinite(dt) || dt <= 0.0 || !isfinite(st->relative_residual) ||
      st->relative_residual < 0.0 || !isfinite(st->absolute_residual) ||
      st->absolute_residual < 0.0 || !isfinite(st->divergence_limit) ||
      st->divergence_limit < 0.0 || st->max_iterations < 0)
    return OR_INVALID_SETTINGS;
  double sum = or_compensated_sum(rhs, n);
  double sum_abs = 0.0, corr = 0.0;
  for (int64_t i = 0; i < n; ++i) neumaier_add(&sum_abs, &corr, fabs(rhs[i]));
  sum_abs += corr;
  double rounding_budget = (64.0 * DBL_EPSILON) * sum_abs;
  if (!isfinite(sum) || !isfinite(sum_abs)) return OR_BREAKDOWN;
  if (fabs(sum) > rounding_budget) {
    rep->error_value = sum;
    return OR_INCOMPATIBLE_RHS;
  }
  if (c->components > 1) {
    /* each component's sum within 64 eps of its absolute sum, components in
     * label order, compensated sums in index order
     * (src/obstacle_pressure.rs:455-472 with :164-183) */
    for (int32_t comp = 0; comp < c->components; ++comp) {
      double cs = 0.0, cc = 0.0, ca = 0.0, cac = 0.0;
      for (int64_t i = 0; i < n; ++i)
        if (c->labels[i] == comp) {
          neumaier_add(&cs, &cc, rhs[i]);
          neumaier_add(&ca, &cac, fabs(rhs[i]));
        }
      cs += cc;
      ca += cac;
      if (!isfinite(cs) || !isfinite(ca)) return OR_BREAKDOWN;
      if (fabs(cs) > (64.0 * DBL_EPSILON) * ca) {
        rep->error_value = cs;
        return OR_INCOMPATIBLE_RHS;
      }
    }
  }
  if (initial) {
    /* caller-supplied sealed pressure guess (src/pressure.rs:615-623) */
    if (!finite_cells(initial, n)) return OR_NONFINITE_INPUT;
    if (initial[c->gauge] != 0.0) return OR_INVALID_INITIAL_PRESSURE;
    for (int64_t i = 0; i < n; ++i) p[i] = initial[i];
  } else {
    for (int64_t i = 0; i < n; ++i) p[i] = 0.0;
  }
  double rhs_l2 = sqrt(or_dot(rhs, rhs, n));
  double threshold = maxd(st->absolute_residual, st->relative_residual * rhs_l2);
  if (!isfinite(rhs_l2) || !isfinite(threshold)) return OR_BREAKDOWN;

  ws w;
  w.r = malloc(sizeof(double) * n);
  w.d = malloc(sizeof(double) * n);
  w.q = malloc(sizeof(double) * n);
  w.z = malloc(sizeof(double) * n);
  if (!w.r || !w.d || !w.q || !w.z) {
    free(w.r), free(w.d), free(w.q), free(w.z);
    return OR_INVALID_ARGUMENT;
  }
  int status = OR_OK;
  or_report report;
  memset(&report, 0, sizeof(report));
#define ACCEPTED(R) ((R).true_residual_l2 <= threshold && (R).predicted_divergence_max <= st->divergence_limit)
  if ((status = true_report(g, c, rhs, p, w.q, dt, 0, threads, &report))) goto done;
  *rep = report;
  if (ACCEPTED(report)) goto done;
  if (st->max_iterations == 0) {
    rep->error_iteration = 0;
    rep->error_value = report.true_residual_max;
    status = OR_ITERATION_LIMIT;
    goto done;
  }
  memcpy(w.r, w.q, sizeof(double) * n);
  double rz;
  if ((status = precondition(g, c, method, w.r, w.z, &rz, threads))) goto done;
  memcpy(w.d, w.z, sizeof(double) * n);
  /* element-wise updates skip the gauge rows (src/pressure.rs:669-679) */
  const int64_t gauge = c->gauge;
  const int multi = c->mask && c->components > 1;
  int par = use_parallel(n - 1, threads);
  int64_t chunks = (n + CHUNK_VALUES - 1) / CHUNK_VALUES;
  for (int64_t it = 1; it <= st->max_iterations; ++it) {
    if ((status = apply_full_c(g, c, w.d, w.q, threads))) goto done;
    w.q[gauge] = 0.0;
    if (multi)
      for (int64_t i = 0; i < n; ++i)
        if (is_gauge(c, i)) w.q[i] = 0.0;
    double curvature = or_dot(w.d, w.q, n);
    if (!isfinite(curvature) || curvature <= 0.0 || !isfinite(rz) || rz <= 0.0) {
      rep->error_iteration = it;
      status = OR_BREAKDOWN;
      goto done;
    }
    double alpha = rz / curvature;
    if (!isfinite(alpha)) {
      rep->error_iteration = it;
      status = OR_BREAKDOWN;
      goto done;
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (par)
    for (int64_t ch = 0; ch < chunks; ++ch) {
      int64_t b = ch * CHUNK_VALUES;
      int64_t e = b + CHUNK_VALUES < n ? b + CHUNK_VALUES : n;
      for (int64_t i = b; i < e; ++i) {
        if (i == gauge || (multi && is_gauge(c, i))) continue;
        p[i] += alpha * w.d[i];
        w.r[i] -= alpha * w.q[i];
      }
    }
    double recursive_l2 = sqrt(or_dot(w.r, w.r, n));
    int candidate = recursive_l2 <= threshold;
    int replaced = 0;
    if (candidate || it == st->max_iterations) {
      if ((status = true_report(g, c, rhs, p, w.q, dt, it, threads, &report))) {
        rep->error_iteration = it;
        goto done;
      }
      *rep = report;
      if (ACCEPTED(report)) goto done;
      /* a seeded solve decides the replacement on the residual without the
       * eliminated gauge row (src/pressure.rs:689-707) */
      double replacement_l2 = report.true_residual_l2;
      if (initial) {
        double sum = 0.0, corr = 0.0;
        for (int64_t i = 0; i < n; ++i)
          if (!(i == gauge || (multi && is_gauge(c, i)))) neumaier_add(&sum, &corr, w.q[i] * w.q[i]);
        replacement_l2 = sqrt(sum + corr);
      }
      if (replacement_l2 > threshold) {
        memcpy(w.r, w.q, sizeof(double) * n);
        replaced = 1;
      }
    }
    if (it == st->max_iterations) {
      rep->error_iteration = it;
      rep->error_value = report.true_residual_max;
      status = OR_ITERATION_LIMIT;
      goto done;
    }
    double next_rz;
    if ((status = precondition(g, c, method, w.r, w.z, &next_rz, threads))) goto done;
    if (!isfinite(next_rz) || next_rz <= 0.0) {
      rep->error_iteration = it;
      status = OR_BREAKDOWN;
      goto done;
    }
    if (replaced) {
      memcpy(w.d, w.z, sizeof(double) * n);
    } else {
      double beta = next_rz / rz;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (par)
      for (int64_t ch = 0; ch < chunks; ++ch) {
        int64_t b = ch * CHUNK_VALUES;
        int64_t e = b + CHUNK_VALUES < n ? b + CHUNK_VALUES : n;
        for (int64_t i = b; i < e; ++i)
          if (!(i == gauge || (multi && is_gauge(c, i)))) w.d[i] = w.z[i] + beta * w.d[i];
      }
    }
    rz = next_rz;
  }
#undef ACCEPTED
done:
  free(w.r), free(w.d), free(w.q), free(w.z);
  return status;
}

int or_solve_rhs(const or_grid* g, double density, const double* rhs, double dt,
                 const or_settings* st, int method, int threads, double* p, or_report* rep) {
  memset(rep, 0, sizeof(*rep));
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  return solve_c(g, &c, rhs, NULL, dt, st, method, threads, p, rep);
}

int or_solve_rhs_initial(const or_grid* g, double density, const double* rhs, const double* initial,
                         double dt, const or_settings* st, int method, int threads, double* p,
                         or_report* rep) {
  memset(rep, 0, sizeof(*rep));
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  return solve_c(g, &c, rhs, initial, dt, st, method, threads, p, rep);
}

int or_apply_acceleration(const or_grid* g, double dt, const double value[3], const double* lower,
                          const double* upper, float* u, float* v, float* w) {
  float* vel[3] = {u, v, w};
  for (int d = 0; d < 3; ++d) {
    const int64_t* dims = g->fn[d];
    for (int64_t k = 0; k < dims[2]; ++k)
      for (int64_t j = 0; j < dims[1]; ++j)
        for (int64_t i = 0; i < dims[0]; ++i) {
          const int64_t q[3] = {i, j, k};
          if (q[d] == 0 || q[d] == g->n[d]) continue;
          double pos[3];
          for (int e = 0; e < 3; ++e) pos[e] = g->origin[e] + ((double)q[e] + (e == d ? 0.0 : 0.5)) * g->h[e];
          int inside = 1;
          if (lower)
            for (int e = 0; e < 3; ++e) inside = inside && pos[e] >= lower[e] && pos[e] < upper[e];
          double a = 0.0;
          if (inside) a += value[d];
          const int64_t index = face_index(g, d, i, j, k);
          const double old = (double)vel[d][index];
          const double updated = old + dt * a;
          if (!isfinite(updated) || fabs(updated) > (double)FLT_MAX) return OR_ARITHMETIC_FAILURE;
          vel[d][index] = (float)updated;
        }
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Solid obstacles (builder-defined sphere scenes).  Discretisation of      */
/* src/obstacle_pressure.rs:224-284 with open areas in {0, A} and fluid      */
/* volumes in {0, V}: a face is open iff both cells are fluid; weights      */
/* A/(rho h) on open faces; solid cells are not unknowns.  The fluid cells  */
/* are labelled by connected component (src/static_obstacle.rs:165-210:     */
/* seeds in index order, breadth-first over open faces); each component    */
/* has one gauge, its largest-volume cell with the lowest index on ties     */
/* (src/obstacle_pressure.rs:310-323), i.e. its seed when all fluid volumes */
/* tie, marked with OR_GAUGE_BIT in the mask.  Summation order is the box   */
/* operator's.                                                              */
/* ------------------------------------------------------------------------ */

/* Component labels of the mask's fluid cells (-1 for solid cells); returns
 * the number of components. */
int32_t or_obstacle_components(const or_grid* g, const uint8_t* mask, int32_t* labels) {
  const int64_t nx = g->n[0], ny = g->n[1], n = g->cell_len;
  const int64_t off[6] = {-1, 1, -nx, nx, -nx * ny, nx * ny};
  int64_t* queue = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  if (!queue) return -1;
  for (int64_t row = 0; row < n; ++row) labels[row] = -1;
  int32_t components = 0;
  for (int64_t seed = 0; seed < n; ++seed) {
    if (!(mask[seed] & 64) || labels[seed] >= 0) continue;
    labels[seed] = components;
    int64_t head = 0, tail = 0;
    queue[tail++] = seed;
    while (head < tail) {
      const int64_t cell = queue[head++];
      for (int side = 0; side < 6; ++side)
        if ((mask[cell] >> side) & 1) {
          const int64_t other = cell + off[side];
          if (labels[other] < 0) {
            labels[other] = components;
            queue[tail++] = other;
          }
        }
    }
    ++components;
  }
  free(queue);
  return components;
}

int or_obstacle_mask(const or_grid* g, const uint8_t* solid, uint8_t* mask, int64_t* gauge) {
  const int64_t nx = g->n[0], ny = g->n[1], n = g->cell_len;
  const int64_t stride[3] = {1, nx, nx * ny};
  *gauge = -1;
  for (int64_t row = 0; row < n; ++row) {
    int64_t p[3] = {row % nx, (row / nx) % ny, row / (nx * ny)};
    uint8_t m = 0;
    if (!solid[row]) {
      m = 64;
      if (*gauge < 0) *gauge = row;
      for (int d = 0; d < 3; ++d) {
        if (p[d] > 0 && !solid[row - stride[d]]) m |= (uint8_t)(1 << (2 * d));
        if (p[d] + 1 < g->n[d] && !solid[row + stride[d]]) m |= (uint8_t)(2 << (2 * d));
      }
    }
    mask[row] = m;
  }
  if (*gauge < 0) return OR_INVALID_ARGUMENT;
  int32_t* labels = malloc(sizeof(int32_t) * (size_t)n);
  if (!labels) return OR_INVALID_ARGUMENT;
  const int32_t components = or_obstacle_components(g, mask, labels);
  /* the seed of each component (its lowest index) is its gauge */
  int32_t next = 0;
  for (int64_t row = 0; row < n && next < components; ++row)
    if (labels[row] == next) {
      mask[row] |= OR_GAUGE_BIT;
      ++next;
    }
  free(labels);
  return components < 1 ? OR_INVALID_ARGUMENT : OR_OK;
}

/* op_coeff of a masked operator: the labels of a mask with several gauges. */
static int masked_coeff(const or_grid* g, const uint8_t* mask, int64_t gauge, op_coeff* c, int32_t** labels) {
  c->mask = mask;
  c->gauge = gauge;
  *labels = NULL;
  int64_t gauges = 0;
  for (int64_t row = 0; row < g->cell_len; ++row) gauges += (mask[row] & OR_GAUGE_BIT) != 0;
  if (gauges <= 1) return OR_OK;
  *labels = malloc(sizeof(int32_t) * (size_t)g->cell_len);
  if (!*labels) return OR_INVALID_ARGUMENT;
  c->components = or_obstacle_components(g, mask, *labels);
  c->labels = *labels;
  return OR_OK;
}

int or_apply_masked(const or_grid* g, double density, const uint8_t* mask, const double* x,
                    double* y) {
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  c.mask = mask;
  return apply_full_c(g, &c, x, y, 1);
}

int or_precondition_masked(const or_grid* g, double density, const uint8_t* mask, int64_t gauge,
                           int method, double* r, double* z, double* rz) {
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  if (method == OR_SGS) return OR_INVALID_ARGUMENT;
  int32_t* labels = NULL;
  if ((s = masked_coeff(g, mask, gauge, &c, &labels))) return s;
  s = precondition(g, &c, method, r, z, rz, 1);
  free(labels);
  return s;
}

int or_solve_rhs_masked(const or_grid* g, double density, const uint8_t* mask, int64_t gauge,
                        const double* rhs, double dt, const or_settings* st, int method,
                        int threads, double* p, or_report* rep) {
  memset(rep, 0, sizeof(*rep));
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  if (method == OR_SGS) return OR_INVALID_ARGUMENT;
  int32_t* labels = NULL;
  if ((s = masked_coeff(g, mask, gauge, &c, &labels))) return s;
  s = solve_c(g, &c, rhs, NULL, dt, st, method, threads, p, rep);
  free(labels);
  return s;
}

/* Corrected f32 faces with closed faces left at zero; max |div| over fluid
 * cells only (src/obstacle_pressure.rs:566-578 keeps solid faces at 0). */
int or_correct_f32_masked(const or_grid* g, double density, const uint8_t* mask,
                          const double* p, double dt, float* u, float* v, float* w) {
  if (!isfinite(dt) || dt <= 0.0) return OR_INVALID_TIME_STEP;
  op_coeff c;
  int s = op_init(g, density, &c);
  if (s) return s;
  float* vel[3] = {u, v, w};
  const int64_t stride[3] = {1, g->n[0], g->n[0] * g->n[1]};
  for (int d = 0; d < 3; ++d) {
    const int64_t* dims = g->fn[d];
    for (int64_t k = 0; k < dims[2]; ++k)
      for (int64_t j = 0; j < dims[1]; ++j)
        for (int64_t i = 0; i < dims[0]; ++i) {
          int64_t q[3] = {i, j, k};
          int64_t index = face_index(g, d, i, j, k);
          if (q[d] == 0 || q[d] == g->n[d]) continue;
          int64_t right = i + g->n[0] * (j + g->n[1] * k);
          if (!(mask[right] & (1 << (2 * d)))) {
            vel[d][index] = 0.0f;
            continue;
          }
          int64_t left = right - stride[d];
          double difference = p[right] - p[left];
          double value = (double)vel[d][index] - dt * c.acc[d] * difference;
          if (!isfinite(value) || fabs(value) > (double)FLT_MAX) return OR_ARITHMETIC_FAILURE;
          vel[d][index] = (float)value;
        }
  }
  return OR_OK;
}

int or_actual_divergence_max_f32_masked(const or_grid* g, const uint8_t* mask, const float* u,
                                        const float* v, const float* w, double* out) {
  double* div = malloc(sizeof(double) * g->cell_len);
  if (!div) return OR_INVALID_ARGUMENT;
  int s = or_divergence_f32(g, u, v, w, div);
  double m = 0.0;
  if (!s)
    for (int64_t i = 0; i < g->cell_len; ++i)
      if (mask[i] & 64) m = maxd(m, fabs(div[i]));
  free(div);
  *out = m;
  return s;
}
