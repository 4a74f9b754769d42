/*
 * rheon_oracle.h -- CPU restatement of Rheon's fixed-box MAC projection step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * rheon_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product never links, calls or falls back to it.
 *
 * Every function restates one reference routine (file:line given beside each
 * declaration; paths relative to the Rheon source tree).  Arithmetic order
 * follows the reference exactly (no FMA contraction: build with
 * -ffp-contract=off), so f64 results are bit-comparable with a faithful port.
 *
 * Storage suffixes: _f32 reads/writes float faces (the reference's own
 * storage), _f64 reads/writes double faces (the f64 variant of the same
 * formula; sampling/advection in f64 storage is a template instantiation of
 * sampling.rs/advection.rs, which the reference only has for f32).
 */
#ifndef RHEON_ORACLE_H
#define RHEON_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared with include/rheon_gpu.h (same numbering). */
enum {
  OR_OK = 0,
  OR_INVALID_ARGUMENT = 7,
  OR_INVALID_SETTINGS = 8,
  OR_INCOMPATIBLE_RHS = 9,
  OR_BREAKDOWN = 10,
  OR_ITERATION_LIMIT = 11,
  OR_NONFINITE_INPUT = 12,
  OR_NONZERO_WALL_VELOCITY = 13,
  OR_ARITHMETIC_FAILURE = 14,
  OR_INVALID_TIME_STEP = 15,
  OR_INVALID_DENSITY = 16,
  OR_NONFINITE_POSITION = 17,
  OR_DIVERGENCE_LIMIT = 18,
  OR_GEOMETRY = 20,
  OR_INVALID_INITIAL_PRESSURE = 24,
  OR_CANCELLED = 6
};

/* GridGeometry (src/geometry.rs:84-191): counts, spacing, origin, upper. */
typedef struct or_grid {
  int64_t n[3];
  int64_t fn[3][3];   /* face counts per axis: fn[a][d] */
  int64_t cell_len;
  int64_t face_len[3];
  double h[3];
  double origin[3];
  double upper[3];
  double volume;
} or_grid;

/* PressureSettings (src/pressure.rs:33-39). */
typedef struct or_settings {
  double relative_residual;
  double absolute_residual;
  double divergence_limit;
  int64_t max_iterations;
} or_settings;

/* PressureReport (src/pressure.rs:43-49) plus the error payload. */
typedef struct or_report {
  int64_t iterations;
  double true_residual_l2;
  double true_residual_max;
  double predicted_divergence_max;
  int64_t error_iteration;   /* Breakdown/IterationLimit/Cancelled payload */
  double error_value;        /* IncompatibleRhs sum / IterationLimit residual_max */
} or_report;

enum { OR_JACOBI = 0, OR_SGS = 1, OR_RBSGS = 2 };

/* GridGeometry::new (src/geometry.rs:114-191). */
int or_grid_init(or_grid* g, const uint64_t counts[3], const double spacing[3],
                 const double origin[3]);
/* BufferPlan::for_grid (src/geometry.rs:284-315): returns total bytes,
 * fills the three parts. */
int64_t or_buffer_plan(const or_grid* g, int64_t* velocity_pair, int64_t* tracer_pair,
                       int64_t* pressure_workspace);
/* GridGeometry::face_position (src/geometry.rs:240-248). */
void or_face_position(const or_grid* g, int axis, const int64_t p[3], double out[3]);
void or_cell_position(const or_grid* g, const int64_t p[3], double out[3]);

/* ScalarSampler::sample (src/sampling.rs:74-122). component: 0..2 faces,
 * 3 = cell-centred values. */
int or_sample_f32(const or_grid* g, int component, const float* values, const double pos[3],
                  double* out);
int or_sample_f64(const or_grid* g, int component, const double* values, const double pos[3],
                  double* out);
/* VelocitySampler::backtrace (src/sampling.rs:151-175). */
int or_backtrace_f32(const or_grid* g, const float* u, const float* v, const float* w,
                     const double pos[3], double dt, double out[3]);
int or_backtrace_f64(const or_grid* g, const double* u, const double* v, const double* w,
                     const double pos[3], double dt, double out[3]);
/* advect_velocity (src/advection.rs:28-68), dense (no band skipping). */
int or_advect_velocity_f32(const or_grid* g, const float* u, const float* v, const float* w,
                           double dt, float* ou, float* ov, float* ow, int threads);
int or_advect_velocity_f64(const or_grid* g, const double* u, const double* v,
                           const double* w, double dt, double* ou, double* ov, double* ow,
                           int threads);
/* advect_tracer (src/advection.rs:72-102), dense evaluation of every cell. */
int or_advect_tracer_f32(const or_grid* g, const float* old, const float* u, const float* v,
                         const float* w, double dt, float* out, int threads);

/* PressureOperator::build_rhs (src/operator.rs:460-509) / build_rhs_f64 (:672-708). */
int or_build_rhs_f32(const or_grid* g, const float* u, const float* v, const float* w,
                     double dt, double* out);
int or_build_rhs_f64(const or_grid* g, const double* u, const double* v, const double* w,
                     double dt, double* out);
/* PressureOperator::apply_full, box fast path (src/operator.rs:378-416). */
int or_apply_full(const or_grid* g, double density, const double* x, double* y, int threads);
/* PressureOperator::diagonal with the cached mask table (src/operator.rs:81-91, :813-838). */
double or_diagonal(const or_grid* g, double density, const int64_t p[3]);
/* divergence (src/operator.rs:513-568) / divergence_f64 (:711-742). */
int or_divergence_f32(const or_grid* g, const float* u, const float* v, const float* w,
                      double* out);
int or_divergence_f64(const or_grid* g, const double* u, const double* v, const double* w,
                      double* out);
/* correct_candidate_in_place (src/operator.rs:627-668) / correct_candidate_f64 (:746-781). */
int or_correct_f32(const or_grid* g, double density, const double* p, double dt, float* u,
                   float* v, float* w);
int or_correct_f64(const or_grid* g, double density, const double* p, double dt, double* u,
                   double* v, double* w);
/* compensated_sum / dot (src/pressure.rs:135-161). */
double or_compensated_sum(const double* x, int64_t n);
double or_dot(const double* a, const double* b, int64_t n);
/* precondition_sgs, serial lexicographic (src/pressure.rs:467-532). Sets r[0]=z[0]=0. */
int or_precondition_sgs(const or_grid* g, double density, double* r, double* z, double* rz);
/* Symmetric red-black Gauss-Seidel preconditioner (R,B forward then B,R
 * reverse) on the gauge-eliminated system.  NOT in the reference: it is the
 * CPU statement of the rheon_amd "rbsgs-pcg" preconditioner. */
int or_precondition_rbsgs(const or_grid* g, double density, double* r, double* z, double* rz);
/* precondition_jacobi (src/pressure.rs:534-575). */
int or_precondition_jacobi(const or_grid* g, double density, double* r, double* z,
                           double* rz, int threads);
/* PressureWorkspace::solve_rhs -> solve_loaded (src/pressure.rs:365-393, :577-740).
 * p (len N) receives the last iterate also on IterationLimit. */
/* solve_rhs_with_initial_pressure (src/pressure.rs:413-429, :615-623,
 * :689-707): the solve seeded with a caller-supplied pressure (finite, 0 at
 * the gauge). */
int or_solve_rhs_initial(const or_grid* g, double density, const double* rhs, const double* initial,
                         double dt, const or_settings* s, int method, int threads, double* p,
                         or_report* rep);
/* One body force on the interior faces (src/forces.rs:97-205): u = f32(f64(u)
 * + dt a(x)) with a(x) = value when the face position x lies in the half-open
 * region [lower, upper) (lower = upper = NULL: the whole box), else 0. */
int or_apply_acceleration(const or_grid* g, double dt, const double value[3], const double* lower,
                          const double* upper, float* u, float* v, float* w);
int or_solve_rhs(const or_grid* g, double density, const double* rhs, double dt,
                 const or_settings* s, int method, int threads, double* p, or_report* rep);
/* PressureWorkspace::actual_divergence_max (src/pressure.rs:330-342). */
int or_actual_divergence_max_f32(const or_grid* g, const float* u, const float* v,
                                 const float* w, double* out);
int or_actual_divergence_max_f64(const or_grid* g, const double* u, const double* v,
                                 const double* w, double* out);
/* Solid obstacles: per-cell open-face mask from a solid-cell flag array
 * (bits 0..5 = faces x-,x+,y-,y+,z-,z+ open, bit 6 = fluid, bit 7 = the
 * gauge of a connected fluid region, its lowest-index cell) and the first
 * gauge.  Discretisation of src/obstacle_pressure.rs:224-323 for
 * all-or-nothing face apertures; summation order of the box operator.
 * or_obstacle_components labels the fluid regions (src/static_obstacle.rs:
 * 165-210; -1 = solid) and returns their count. */
int or_obstacle_mask(const or_grid* g, const uint8_t* solid, uint8_t* mask, int64_t* gauge);
int32_t or_obstacle_components(const or_grid* g, const uint8_t* mask, int32_t* labels);
int or_apply_masked(const or_grid* g, double density, const uint8_t* mask, const double* x,
                    double* y);
int or_precondition_masked(const or_grid* g, double density, const uint8_t* mask, int64_t gauge,
                           int method, double* r, double* z, double* rz);
int or_solve_rhs_masked(const or_grid* g, double density, const uint8_t* mask, int64_t gauge,
                        const double* rhs, double dt, const or_settings* s, int method,
                        int threads, double* p, or_report* rep);
int or_correct_f32_masked(const or_grid* g, double density, const uint8_t* mask,
                          const double* p, double dt, float* u, float* v, float* w);
int or_actual_divergence_max_f32_masked(const or_grid* g, const uint8_t* mask, const float* u,
                                        const float* v, const float* w, double* out);
/* velocity_rate (src/simulation.rs:1415-1429). */
double or_velocity_rate_f32(const or_grid* g, const float* u, const float* v, const float* w);
double or_velocity_rate_f64(const or_grid* g, const double* u, const double* v,
                            const double* w);

/* ---- dense3d frames (dense3d_oracle.c; src/dense3d_sequence.rs:415-564) */
/* The frame header and diagnostics values (already checked). */
typedef struct or_d3_meta {
  uint64_t frame;
  double time;   /* state.carrier.time */
  int32_t has_report;
  int32_t pad;
  double dt;     /* report.carrier.step.dt */
  uint64_t carrier_id, carrier_version, liquid_id, liquid_version;
  uint64_t pressure_iterations;
  double diag[11]; /* pressure_residual, carrier and liquid divergence, volume before/after, mass after,
                      inward, outward, source, balance, rounding budget (the diagnostics object's order) */
} or_d3_meta;
/* Rust's `{:.Pe}` of a finite double with `digits` = P + 1 significant digits. */
int or_format_sci(double v, int digits, char* out);
/* write_dense3d_frame's line (src/dense3d_sequence.rs:497-563); returns its
   length, or minus the length when out is NULL or cap is smaller.  threads
   > 1 formats the field arrays in 64 Ki-value chunks on that many OpenMP
   threads (the same bytes). */
int64_t or_dense3d_frame(const int64_t n[3], const float* u, const float* v, const float* w, const float* tracer,
                         const double* fraction, const double* pressure, const or_d3_meta* meta, char* out,
                         int64_t cap, int threads);

#ifdef __cplusplus
}
#endif
#endif
