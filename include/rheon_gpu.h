/*
 * rheon_gpu.h -- C ABI of rheon_amd: Rheon's fixed-box MAC projection step on
 * one AMD Instinct MI355X (gfx950).
 *
 * The boundary mirrors the reference's pressure-provider seam
 * (src/solver_provider.rs:43-76, PressureRequest / PressureProvider /
 * ProviderRejection), its host qualification + fallback workspace
 * (src/solver_provider.rs:94-327, ProviderWorkspace), the pressure workspace
 * (src/pressure.rs:243-740, PressureWorkspace::solve_rhs/solve_velocity) and
 * the per-step operator calls of the carrier simulation
 * (src/simulation.rs:926-1277: advect_velocity, build_rhs, solve,
 * correct_candidate_in_place, actual_divergence_max, gates).
 *
 * Conventions
 *  - Plain pointers and sizes only.  Arrays are x-fastest flat arrays exactly
 *    as in the reference (cell i + nx*(j + ny*k); face arrays of axis a have
 *    counts n + e_a, src/geometry.rs:250-258).
 *  - Functions taking `d_` pointers expect device memory from rh_device_alloc
 *    (or any hipMalloc'd memory on the current device); all others take host
 *    memory.
 *  - Every entry point is synchronous with respect to the caller (results are
 *    complete on return), never throws, and reports failures as rh_status.
 *    rh_last_error() returns a thread-local message for the last failure.
 *  - Cancellation callbacks are polled only on the calling thread, between
 *    device batches (the reference polls once per iteration,
 *    src/pressure.rs:652-656; see rh_pressure_settings.poll_interval).
 *  - Not thread-safe per object: one thread at a time may use a given
 *    solver, provider, workspace or scene.
 */
#ifndef RHEON_GPU_H
#define RHEON_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RH_ABI_VERSION 2

/* Status codes.  0-6 mirror ProviderRejection (src/solver_provider.rs:54-62);
 * 7-19 mirror PressureError / OperatorError / SimulationError variants
 * (src/pressure.rs:51-77, src/operator.rs:6-17, src/simulation.rs gates). */
typedef enum rh_status {
  RH_OK = 0,
  RH_UNSUPPORTED = 1,          /* ProviderRejection::Unsupported */
  RH_EXECUTION = 2,            /* ProviderRejection::Execution */
  RH_NONFINITE_OUTPUT = 3,     /* ProviderRejection::NonFiniteOutput */
  RH_NONZERO_GAUGE = 4,        /* ProviderRejection::NonZeroGauge */
  RH_QUALITY_GATE = 5,         /* ProviderRejection::QualityGate */
  RH_CANCELLED = 6,            /* ProviderRejection::Cancelled / PressureError::Cancelled */
  RH_INVALID_ARGUMENT = 7,     /* null pointer, length mismatch, bad enum */
  RH_INVALID_SETTINGS = 8,     /* PressureError::InvalidSettings */
  RH_INCOMPATIBLE_RHS = 9,     /* PressureError::IncompatibleRhs */
  RH_BREAKDOWN = 10,           /* PressureError::Breakdown */
  RH_ITERATION_LIMIT = 11,     /* PressureError::IterationLimit */
  RH_NONFINITE_INPUT = 12,     /* OperatorError::NonFiniteInput / SamplingError::NonFiniteField */
  RH_NONZERO_WALL_VELOCITY = 13, /* OperatorError::NonZeroWallVelocity */
  RH_ARITHMETIC_FAILURE = 14,  /* OperatorError/SamplingError::ArithmeticFailure */
  RH_INVALID_TIME_STEP = 15,   /* OperatorError::InvalidTimeStep */
  RH_INVALID_DENSITY = 16,     /* OperatorError::InvalidDensity */
  RH_TIME_RESOLUTION = 17,     /* SimulationError::TimeResolution */
  RH_DIVERGENCE_LIMIT = 18,    /* SimulationError::DivergenceLimit */
  RH_COURANT_LIMIT = 19,       /* SimulationError::CourantLimit */
  RH_GEOMETRY = 20,            /* GeometryError::* */
  RH_BUFFER_LIMIT = 21,        /* PressureError::BufferLimit / GeometryError::BufferLimit */
  RH_DEVICE_ERROR = 22,        /* HIP runtime failure (message in rh_last_error) */
  RH_REFUSED = 23,             /* ProjectionError::Refused (the kind is in rh_projection_report.refusal) */
  RH_INVALID_INITIAL_PRESSURE = 24, /* PressureError::InvalidInitialPressure */
  RH_INVALID_STATE = 25        /* Dense3dError::InvalidState (the reference's message in rh_last_error) */
} rh_status;

/* GridGeometry::new inputs (src/geometry.rs:114-118). */
typedef struct rh_grid {
  uint64_t counts[3];
  double spacing[3];
  double origin[3];
} rh_grid;

/* Storage precision of face and pressure arrays. */
typedef enum rh_precision { RH_F64 = 0, RH_F32 = 1 } rh_precision;

/* Pressure methods.  JACOBI_PCG is the reference's "jacobi-pcg-v1"
 * (src/pressure.rs:6-29), iterate-for-iterate.  RBSGS_PCG is PCG with a
 * symmetric red-black Gauss-Seidel preconditioner (replaces the
 * lexicographic "sgs-pcg-v1"; different preconditioner, same operator,
 * gauge and stopping rules).  MG_PCG and MIXED_MG are the multigrid solvers. */
typedef enum rh_method {
  RH_JACOBI_PCG = 0,
  RH_RBSGS_PCG = 1,
  RH_MG_PCG = 2,
  RH_MIXED_MG = 3,
  RH_SGS_PCG = 4   /* the reference's "sgs-pcg-v1" (src/pressure.rs:467-532), iterate-for-iterate; box only */
} rh_method;

/* PressureSettings (src/pressure.rs:33-39) + device polling cadence. */
typedef struct rh_pressure_settings {
  double relative_residual;
  double absolute_residual;
  double divergence_limit;
  int64_t max_iterations;
  int32_t poll_interval;   /* iterations between host polls; 0 = default (16) */
  int32_t reserved;
} rh_pressure_settings;

/* PressureReport (src/pressure.rs:43-49) + error payload. */
typedef struct rh_pressure_report {
  int64_t iterations;
  double true_residual_l2;
  double true_residual_max;
  double predicted_divergence_max;
  int64_t error_iteration;   /* Breakdown/IterationLimit/Cancelled iteration */
  double error_value;        /* IncompatibleRhs sum, IterationLimit residual_max */
  double threshold;          /* max(absolute, relative*||b||) actually used */
  double seconds;            /* device wall time of the solve */
} rh_pressure_report;

typedef int (*rh_cancel_fn)(void* ctx); /* nonzero = cancel */

/* ---------------------------------------------------------------- basics */
int32_t rh_abi_version(void);
const char* rh_last_error(void);
/* Number of visible devices and the gfx arch name of device 0 ("gfx950"). */
rh_status rh_device_info(int32_t* count, char* arch, size_t arch_len);
rh_status rh_device_alloc(size_t bytes, void** out);
rh_status rh_device_free(void* ptr);
rh_status rh_copy_to_device(void* dst, const void* src, size_t bytes);
rh_status rh_copy_to_host(void* dst, const void* src, size_t bytes);
rh_status rh_device_synchronize(void);
/* Device-to-device copy by a plain streaming kernel (also the known-size
 * reference for hardware traffic counters); bytes a multiple of 16. */
rh_status rh_copy_device(void* dst, const void* src, size_t bytes);
/* GridGeometry::new validation (src/geometry.rs:114-191); cells may be NULL. */
rh_status rh_grid_validate(const rh_grid* grid, uint64_t* cells);
/* BufferPlan::for_grid (src/geometry.rs:284-315): reference f32 payload. */
rh_status rh_buffer_plan(const rh_grid* grid, uint64_t* velocity_pair, uint64_t* tracer_pair,
                         uint64_t* pressure_workspace, uint64_t* total);

/* Kernel timing with HIP events recorded on each kernel's own stream
 * (diagnostics; no reference counterpart).  enable(1) clears the totals;
 * query() returns the launches and summed seconds of one kernel, by name
 * ("pcg_update", "mg_smooth", ...).  Launches enqueued after a solve has
 * converged (which return at once) are not counted. */
rh_status rh_profile_enable(int32_t on);
rh_status rh_profile_query(const char* kernel, int64_t* launches, double* seconds);
/* Time only `kernel`'s launches (NULL = every kernel); each timed launch
 * costs the stream an event pair. */
rh_status rh_profile_focus(const char* kernel);

/* ------------------------------------------------ operator kernels (device) */
/* PressureOperator::apply_full, box path (src/operator.rs:378-456). f64. */
rh_status rh_apply_full(const rh_grid* grid, double density, const double* d_x, double* d_y);
/* advect_velocity (src/advection.rs:28-68): faces in `precision` storage. */
rh_status rh_advect_velocity(const rh_grid* grid, rh_precision precision, const void* d_u,
                             const void* d_v, const void* d_w, double dt, void* d_out_u,
                             void* d_out_v, void* d_out_w);
/* advect_tracer (src/advection.rs:72-102), f32 storage. */
rh_status rh_advect_tracer(const rh_grid* grid, const float* d_old, const float* d_u,
                           const float* d_v, const float* d_w, double dt, float* d_out);
/* build_rhs / build_rhs_f64 (src/operator.rs:460-509, :672-708), validating
 * finiteness and zero wall faces (:850-875). Output f64. */
rh_status rh_build_rhs(const rh_grid* grid, rh_precision precision, const void* d_u,
                       const void* d_v, const void* d_w, double dt, double* d_rhs);
/* correct_candidate_in_place / correct_candidate_f64 (src/operator.rs:627-668,
 * :746-781). Pressure f64. */
rh_status rh_correct_velocity(const rh_grid* grid, rh_precision precision, double density,
                              const double* d_pressure, double dt, void* d_u, void* d_v,
                              void* d_w);
/* actual_divergence_max (src/pressure.rs:330-342 with src/operator.rs:513-568). */
rh_status rh_divergence_max(const rh_grid* grid, rh_precision precision, const void* d_u,
                            const void* d_v, const void* d_w, double* out_max);
/* rh_divergence_max without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  Same checks and result as rh_divergence_max, on the
 * null stream, with its reduction scratch (partials for kMaxBlocks = 16384
 * blocks, flags, result: 1.5 MiB) in a fixed device array that each GPU holds
 * until the process ends; every grid that passes the 32-bit indexing check
 * fits.  Calls no hipMalloc/hipFree.  Calls that overlap on one GPU race on
 * that GPU's array. */
rh_status rh_divergence_max_noalloc(const rh_grid* grid, rh_precision precision, const void* d_u,
                                    const void* d_v, const void* d_w, double* out_max);
/* velocity_rate (src/simulation.rs:1415-1429): sum_d max|u_d| / h_d. */
rh_status rh_velocity_rate(const rh_grid* grid, rh_precision precision, const void* d_u,
                           const void* d_v, const void* d_w, double* out_rate);
/* One preconditioner application z = M^-1 r on the gauge-eliminated system
 * (precondition_jacobi src/pressure.rs:534-575; RBSGS: red-black symmetric
 * Gauss-Seidel).  r[0] is zeroed as in the reference; returns dot(r, z). */
rh_status rh_precondition(const rh_grid* grid, double density, rh_method method, double* d_r,
                          double* d_z, double* out_rz);

/* -------------------------------------------- pressure workspace (device) */
/* PressureWorkspace on the device (src/pressure.rs:243-314): the solver
 * owns its Krylov arrays and reduction scratch; construction allocates,
 * solves do not. */
typedef struct rh_solver rh_solver;
typedef struct rh_solver_config {
  rh_method method;
  rh_precision precision;    /* Krylov vector storage (F64 = reference arithmetic) */
  int32_t mg_levels;         /* multigrid levels (0 = automatic) */
  int32_t mg_pre_sweeps;     /* 0 = default (Chebyshev 3, Jacobi 2); at most 8 */
  int32_t mg_post_sweeps;    /* 0 = default (Chebyshev 1, Jacobi 2); at most 8 */
  double mg_omega;           /* weighted-Jacobi smoother weight (0 = 6/7) */
  uint64_t memory_limit;     /* bytes; 0 = unlimited (PressureWorkspace limit) */
  uint32_t kernel_variant;   /* 0 = default kernels; RH_VARIANT_* bits select alternatives */
  uint32_t mg_smoother;      /* RH_SMOOTHER_* (0 = Chebyshev) */
  int32_t mg_coarse_pre_sweeps;   /* levels below the finest; 0 = the finest level's count */
  int32_t mg_coarse_post_sweeps;  /* levels below the finest; 0 = the finest level's count */
} rh_solver_config;

/* Multigrid smoothers.  Both sweep x <- x + tau_k (b - K x) / D; Jacobi uses
 * tau_k = mg_omega, Chebyshev the reciprocal roots of the Chebyshev
 * polynomial of the side's sweep count on the interval [lo, 2] of D^-1 K's
 * spectrum (lo = 0.3 for up to 2 sweeps, 0.25 beyond), the same memory
 * traffic per sweep with a smaller error-reduction factor. */
#define RH_SMOOTHER_CHEBYSHEV 0u
#define RH_SMOOTHER_JACOBI 1u

/* Kernel variants (same results; for tuning and for cross-checking). */
#define RH_VARIANT_SCALAR_MG 1u /* one cell per lane in every multigrid kernel */

rh_status rh_solver_create(const rh_grid* grid, const rh_solver_config* config,
                           rh_solver** out);
void rh_solver_destroy(rh_solver* solver);
uint64_t rh_solver_allocated_bytes(const rh_solver* solver);
/* solve_rhs (src/pressure.rs:365-393 -> solve_loaded :577-740).  d_rhs: N f64
 * (full, ungauged RHS).  d_pressure: N f64, receives the last iterate (also
 * on ITERATION_LIMIT), gauge cell 0 exactly 0. */
rh_status rh_solver_solve_rhs(rh_solver* solver, double density, const double* d_rhs,
                              double dt, const rh_pressure_settings* settings,
                              double* d_pressure, rh_pressure_report* report,
                              rh_cancel_fn cancel, void* cancel_ctx);
/* solve_rhs_with_initial_pressure (src/pressure.rs:413-429, feature
 * accepted-pressure-warm-start): the solve started from d_initial (N f64,
 * device; may be d_pressure itself), which must be finite (NONFINITE_INPUT)
 * and 0 at the gauge (INVALID_INITIAL_PRESSURE).  A seeded jacobi-pcg-v1 /
 * sgs-pcg-v1 solve replaces its recursive residual only when the true
 * residual without the gauge row exceeds the threshold (:689-707); the
 * multigrid methods start their defect correction from the seed. */
rh_status rh_solver_solve_rhs_initial(rh_solver* solver, double density, const double* d_rhs,
                                      const double* d_initial, double dt,
                                      const rh_pressure_settings* settings, double* d_pressure,
                                      rh_pressure_report* report, rh_cancel_fn cancel,
                                      void* cancel_ctx);

/* Solid obstacles for this solver: solid = one flag per cell (nonzero =
 * solid, host, len N) or NULL for the sealed box.  The discretisation of
 * StaticObstaclePressure (src/obstacle_pressure.rs:224-309) with open areas in
 * {0, A} and fluid volumes in {0, V}: a face is open iff both cells are fluid,
 * solid cells are not unknowns, and every connected fluid region has one
 * gauge, its lowest-index cell (p there is exactly 0; p of solid cells is
 * returned as 0), and its own compatibility check (src/obstacle_pressure.rs:
 * 310-323, :455-472).  More than 254 regions: UNSUPPORTED. */
rh_status rh_solver_set_solid(rh_solver* solver, const uint8_t* solid, size_t len);

/* ------------------------------------------------- provider (host memory) */
/* PressureRequest (src/solver_provider.rs:47-52). rhs: N f64, host. */
typedef struct rh_pressure_request {
  const rh_grid* grid;
  double density;
  double dt;
  const double* rhs;
  size_t rhs_len;
} rh_pressure_request;

/* SolverMetadata + TensorMetadata (src/solver_provider.rs:17-33). */
typedef struct rh_solver_metadata {
  char id[48];
  char version[16];
  char stage[16];
  uint64_t shape_xyz[3];
  char layout[32];
  char scalar[8];
  char units_in[16];
  char units_out[16];
  char domain[48];
} rh_solver_metadata;

typedef struct rh_provider_config {
  rh_solver_config solver;
  rh_pressure_settings settings;  /* the provider's own stopping rule */
  uint64_t shape_xyz[3];          /* cells the provider serves; {0,0,0} = any shape */
} rh_provider_config;

/* impl PressureProvider for the GPU solver (trait src/solver_provider.rs:67-76). */
typedef struct rh_provider rh_provider;
rh_status rh_provider_create(const rh_provider_config* config, rh_provider** out);
void rh_provider_destroy(rh_provider* provider);
/* metadata() (src/solver_provider.rs:67-68): per provider; shape_xyz is the
 * configured shape ({0,0,0} when the provider serves any shape). */
rh_status rh_provider_metadata(const rh_provider* provider, rh_solver_metadata* out);
/* supports(): RH_OK or RH_UNSUPPORTED with a static reason string (also for
 * a request whose grid differs from a configured shape). */
rh_status rh_provider_supports(const rh_provider* provider, const rh_pressure_request* request,
                               const char** reason);
/* predict(): writes all N cells of `pressure` (host), p[0] == 0.  Returns
 * RH_OK, RH_CANCELLED, RH_UNSUPPORTED or RH_EXECUTION (+ reason). */
rh_status rh_provider_predict(rh_provider* provider, const rh_pressure_request* request,
                              double* pressure, size_t pressure_len, rh_cancel_fn cancel,
                              void* cancel_ctx, const char** reason);

/* ProviderReport (src/solver_provider.rs:78-92); durations in seconds. */
typedef struct rh_provider_report {
  char selected[48];
  int32_t has_attempted;             /* a preferred provider was offered ... */
  rh_solver_metadata attempted;      /* ... and this is its metadata (Option<SolverMetadata>) */
  rh_status rejection;               /* RH_OK if none */
  int32_t has_candidate;
  rh_pressure_report candidate;
  rh_pressure_report accepted;
  double total, inference, qualification, fallback;
} rh_provider_report;

/* ProviderWorkspace (src/solver_provider.rs:94-327): input checks, NaN
 * pre-fill, predict, residual qualification (on the device), and the native
 * fallback -- the device "jacobi-pcg-v1" (or "sgs-pcg-v1", "rbsgs-pcg") -- when
 * the candidate is rejected. */
typedef struct rh_provider_workspace rh_provider_workspace;
rh_status rh_provider_workspace_create(const rh_grid* grid, uint64_t limit,
                                       rh_method fallback, rh_provider_workspace** out);
void rh_provider_workspace_destroy(rh_provider_workspace* ws);
rh_status rh_provider_workspace_solve(rh_provider_workspace* ws,
                                      const rh_pressure_request* request,
                                      const rh_pressure_settings* settings,
                                      rh_provider* preferred /* may be NULL */,
                                      rh_cancel_fn cancel, void* cancel_ctx,
                                      rh_provider_report* report);
/* pressure(): the last accepted pressure (host copy, N f64). */
rh_status rh_provider_workspace_pressure(const rh_provider_workspace* ws, double* out,
                                         size_t len);

/* ------------------------------- supplied-projection qualification (device) */
/* ProjectionWorkspace::qualify and its request (src/solver_provider/composition/
 * projection.rs:24-131, :238-542): a supplied (pressure, corrected velocity)
 * pair for a sealed-box projection is accepted only when the velocity is, bit
 * for bit, the input velocity corrected by the supplied pressure (the momentum
 * witness), the pressure passes the provider workspace's requalification and
 * the corrected faces pass the actual-divergence gate; otherwise the native
 * device solve answers.  Arrays are host memory; the work runs on the device. */
typedef enum rh_projection_domain {
  RH_DOMAIN_SEALED_BOX = 0, RH_DOMAIN_FREE_SURFACE = 1, RH_DOMAIN_EMBEDDED_OBSTACLE = 2,
  RH_DOMAIN_MOVING_WALL = 3
} rh_projection_domain;
/* ProjectionInput (:31-42); SessionIdentity = (run_id, source_revision). */
typedef struct rh_projection_input {
  const rh_grid* grid;
  double density;
  double interval_start;
  double dt;
  uint64_t run_id;
  const char* source_revision;      /* UTF-8, 1..4096 bytes, no control characters */
  uint64_t generation;
  const float* velocity[3];         /* provisional faces (host) */
  size_t velocity_len[3];
  rh_pressure_settings settings;
  double actual_divergence_limit;
  rh_projection_domain domain;
} rh_projection_input;
/* TensorMetadata and ProjectionRepresentation (:150-179). */
typedef struct rh_tensor_metadata {
  uint64_t shape_xyz[3];
  char layout[32];
  char scalar[8];
  char units[16];
} rh_tensor_metadata;
typedef struct rh_projection_representation {
  uint32_t version;
  char operator_id[64];
  rh_tensor_metadata pressure;
  rh_tensor_metadata velocity[3];
} rh_projection_representation;
/* SuppliedProjection (:187-193) with its ProjectionCandidateIdentity. */
typedef struct rh_projection_candidate {
  uint8_t input_id[32];
  rh_projection_representation representation;
  const char* algorithm;            /* 1..256 bytes */
  const char* version;              /* 1..256 bytes */
  const char* source_revision;      /* 1..4096 bytes */
  const double* pressure;
  size_t pressure_len;
  const float* velocity[3];
  size_t velocity_len[3];
} rh_projection_candidate;
/* ProjectionSelection, ProjectionRejection, ProjectionRefusal (:195-219). */
enum { RH_SELECT_SUPPLIED = 0, RH_SELECT_NATIVE_BASELINE = 1, RH_SELECT_NATIVE_FALLBACK = 2 };
enum { RH_REJECT_NONE = 0, RH_REJECT_FIELD = 1, RH_REJECT_MOMENTUM_WITNESS = 2, RH_REJECT_PRESSURE = 3,
       RH_REJECT_ACTUAL_DIVERGENCE = 4 };
enum { RH_REFUSE_NONE = 0, RH_REFUSE_DOMAIN = 1, RH_REFUSE_WORKSPACE_GRID = 2, RH_REFUSE_INPUT_BINDING = 3,
       RH_REFUSE_REPRESENTATION = 4, RH_REFUSE_IDENTITY = 5 };
/* ProjectionReport (:256-270). */
typedef struct rh_projection_report {
  uint8_t input_id[32];
  int32_t selection;
  int32_t refusal;                  /* set with RH_REFUSED */
  int32_t rejection;                /* RH_REJECT_* */
  rh_status rejection_status;       /* FIELD: the OperatorError; PRESSURE: the ProviderRejection */
  int32_t witness_axis;             /* MOMENTUM_WITNESS: first differing face (axis, index); */
  int32_t wall_axis;                /* FIELD + NONZERO_WALL_VELOCITY: the offending face */
  int64_t witness_index;
  int64_t wall_index;
  double rejected_divergence;       /* ACTUAL_DIVERGENCE: the maximum that failed */
  int32_t has_supplied_identity;
  int32_t has_discarded;
  rh_provider_report pressure;
  rh_provider_report discarded_candidate_pressure;  /* valid when has_discarded */
  double actual_divergence_max;
} rh_projection_report;
typedef struct rh_projection_workspace rh_projection_workspace;
/* ProjectionRequest::new (:58-131): validates the input and returns its
 * input id (SHA-256 of the contract, representation, grid, identity,
 * settings and every face's bits, in the reference's byte order). */
rh_status rh_projection_input_id(const rh_projection_input* input, uint8_t out[32]);
/* ProjectionRepresentation::sealed_v1 (:158-170). */
rh_status rh_projection_representation_of(const rh_grid* grid, rh_projection_representation* out);
/* ProjectionWorkspace::new (:283-335): limit = byte cap (16 N + 4 F + the
 * provider workspace), 0 = none; fallback = RH_JACOBI_PCG, RH_SGS_PCG or
 * RH_RBSGS_PCG. */
rh_status rh_projection_workspace_create(const rh_grid* grid, uint64_t limit, rh_method fallback,
                                         rh_projection_workspace** out);
void rh_projection_workspace_destroy(rh_projection_workspace* ws);
/* qualify (:350-479).  candidate may be NULL (native baseline).  On RH_OK the
 * qualified pressure and faces are read with rh_projection_result. */
rh_status rh_projection_qualify(rh_projection_workspace* ws, const rh_projection_input* input,
                                const rh_projection_candidate* candidate, rh_cancel_fn cancel,
                                void* cancel_ctx, rh_projection_report* report);
rh_status rh_projection_result(const rh_projection_workspace* ws, double* pressure, size_t pressure_len,
                               float* u, float* v, float* w);

/* --------------------------------------------- device-resident scene step */
/* A fixed-box carrier with accepted and candidate faces on the device
 * (src/simulation.rs accepted/candidate sets, commit swap :805-811). */
typedef struct rh_scene rh_scene;
typedef struct rh_scene_config {
  double density;
  rh_precision precision;           /* face storage */
  rh_solver_config solver;
  rh_pressure_settings pressure;
  double actual_divergence_limit;   /* SimulationConfig (src/simulation.rs:11-18) */
  double max_courant;               /* 0 disables the Courant gate */
} rh_scene_config;

/* StepReport (src/simulation.rs:44-53) + timings. */
typedef struct rh_step_report {
  double dt;
  rh_pressure_report pressure;
  double actual_divergence_max;
  double courant;
  double seconds;          /* device time of the whole step */
  double advect_seconds, rhs_seconds, solve_seconds, correct_seconds;
  double time;             /* scene time after the step (0 + sum of committed dt) */
  uint64_t generation;     /* committed steps after the step */
  double tracer_integral;  /* sum of the candidate tracer * cell volume (0 without a tracer) */
  double kinetic_energy;   /* 0.5 rho V sum of squared candidate faces */
} rh_step_report;

rh_status rh_scene_create(const rh_grid* grid, const rh_scene_config* config, rh_scene** out);
void rh_scene_destroy(rh_scene* scene);
/* Device bytes the scene holds: its arrays and solver, plus what its steps
 * added (the spare faces of the out-of-place correction, the accepted
 * pressure kept aside by warm steps). */
uint64_t rh_scene_allocated_bytes(const rh_scene* scene);
/* Upload/download the accepted faces (host arrays in the scene precision). */
rh_status rh_scene_set_velocity(rh_scene* scene, const void* u, const void* v, const void* w);
rh_status rh_scene_get_velocity(const rh_scene* scene, void* u, void* v, void* w);
/* Solid obstacles of the scene (see rh_solver_set_solid).  Obstacle faces
 * of the accepted velocity are set to 0 now and after every set_velocity,
 * advection writes 0 on them and the correction leaves them at 0. */
rh_status rh_scene_set_solid(rh_scene* scene, const uint8_t* solid, size_t len);
/* The scene's pressure (N f64, host): after an accepted step, that step's;
 * after a failed attempt whose solve ran, that attempt's solve output (as
 * the reference's, the view is meaningful only after an accepted step). */
rh_status rh_scene_get_pressure(const rh_scene* scene, double* out, size_t len);
/* Device pointers of the accepted faces (for zero-copy producers). */
rh_status rh_scene_device_velocity(rh_scene* scene, void** d_u, void** d_v, void** d_w);
/* One projection step with an explicit dt: advect -> build_rhs -> solve ->
 * correct -> divergence gate -> Courant gate -> tracer (when the scene has
 * one) -> commit.  On any failure the accepted faces (and tracer) are
 * unchanged (transactional, src/simulation.rs:391-392). */
rh_status rh_scene_step(rh_scene* scene, double dt, rh_step_report* report,
                        rh_cancel_fn cancel, void* cancel_ctx);
/* Simulation::step's time-step policy (src/simulation.rs:926-979, the fixed
 * box without sources): dt = min(requested, 0.5 max_courant / rate) with
 * rate = sum_d max|u_d| / h_d of the accepted faces (src/simulation.rs:
 * 1415-1429), then rh_scene_step(dt).  max_courant must lie in (0, 1]
 * (src/simulation.rs:293-295).  The rate of faces a step committed is kept
 * from that step's Courant gate, so no extra pass runs. */
rh_status rh_scene_step_requested(rh_scene* scene, double requested_dt, rh_step_report* report,
                                  rh_cancel_fn cancel, void* cancel_ctx);
/* Passive tracer of the scene (Simulation's accepted/candidate tracer,
 * src/simulation.rs:1290-1297): one f32 per cell (host), or NULL to drop
 * it.  With a tracer, every step advects it with the corrected candidate
 * velocity (advect_tracer, src/advection.rs:72-162, cells outside the
 * displaced support of its nonzero cells are exactly 0), rejects a value
 * outside [0, 1] or non-finite (src/simulation.rs:1322-1331) and commits it
 * with the faces. */
/* Pressure initialization of every later step's solve
 * (PressureInitializationPolicy, feature accepted-pressure-warm-start,
 * src/liquid_step.rs:800-850): COLD starts from p = 0 (the default);
 * ACCEPTED_PREVIOUS seeds the solve with the pressure of the last committed
 * step (solve_rhs_with_initial_pressure), cold when there is none (a new
 * velocity or obstacle set clears it).  Batched scenes: UNSUPPORTED. */
typedef enum rh_init_policy { RH_INIT_COLD = 0, RH_INIT_ACCEPTED_PREVIOUS = 1 } rh_init_policy;
rh_status rh_scene_set_initialization(rh_scene* scene, rh_init_policy policy);
rh_status rh_scene_set_tracer(rh_scene* scene, const float* tracer, size_t len);
rh_status rh_scene_get_tracer(const rh_scene* scene, float* out, size_t len);

/* ------------------------------------------------------ batched scenes */
/* Many device-resident scenes advanced together: each stage of the step is
 * one launch over every scene (no counterpart in the reference, which runs
 * scenes one at a time or as CPU BatchRun jobs, src/scheduling.rs:128-221;
 * per scene the step is rh_scene_step's, src/simulation.rs:988-1277).  The
 * solve enqueues max_cycles V-cycles per step and converged scenes' blocks
 * return at once, so one step is a fixed launch sequence that can be
 * captured once in a hipGraph and replayed. */
typedef struct rh_batch rh_batch;
typedef struct rh_batch_config {
  rh_scene_config scene;  /* shared: f32 faces, RH_MIXED_MG with f32 V-cycles */
  int32_t max_cycles;     /* V-cycles per step at most (0 = scene.pressure.max_iterations) */
  /* 0: direct launches, the host stops enqueuing cycles once every scene's
   *    solve is done (polled two cycles behind);
   * 1: one hipGraph replay per step holding max_cycles V-cycles;
   * 2: hipGraph segments per step: the cycles the last step needed, then
   *    one more at a time while some scene is still solving (one host round
   *    trip per check).  Every mode gives the same results. */
  int32_t use_graph;
} rh_batch_config;

rh_status rh_batch_create(const rh_grid* grids, size_t n, const rh_batch_config* config,
                          rh_batch** out);
void rh_batch_destroy(rh_batch* batch);
uint64_t rh_batch_allocated_bytes(const rh_batch* batch);
/* Borrowed handle of scene i (set_velocity / set_solid / get_* as usual). */
rh_status rh_batch_scene(rh_batch* batch, size_t i, rh_scene** out);
/* dt per scene (n entries), kept for every following step. */
rh_status rh_batch_set_dt(rh_batch* batch, const double* dt, size_t n);
/* Kernel variant (RH_VARIANT_* bits) of the batched launches. */
rh_status rh_batch_set_variant(rh_batch* batch, uint32_t kernel_variant);
/* Launch shape of the batched stages: kernel variant and the divisor of
 * each scene's one-scene block count (1 = the one-scene decomposition). */
typedef struct rh_batch_tuning {
  uint32_t kernel_variant;
  int32_t block_divisor;
  double seconds_per_step;  /* measured by rh_batch_autotune */
} rh_batch_tuning;
rh_status rh_batch_set_tuning(rh_batch* batch, const rh_batch_tuning* tuning);
/* Autotuner: times every candidate shape for `steps` steps on the batch's
 * own scenes (accepted faces saved before and restored after; pressure
 * buffers hold the last trial's values), applies the fastest and returns it
 * in `chosen`; `table` (table_len entries, may be NULL) receives every
 * candidate tried and `tried` their number. */
rh_status rh_batch_autotune(rh_batch* batch, int32_t steps, rh_batch_tuning* table,
                            size_t table_len, size_t* tried, rh_batch_tuning* chosen);
/* `steps` steps of every scene; reports/statuses (n entries each, may be
 * NULL) describe the last step.  A scene whose step fails keeps its
 * accepted faces (transactional, as rh_scene_step) and its status says why. */
rh_status rh_batch_step(rh_batch* batch, int32_t steps, rh_step_report* reports,
                        rh_status* statuses);

/* ------------------------------------------------------------ serving */
/* Continuous-batching server over device-resident scenes (no counterpart in
 * the reference).  Requests are admitted in arrival order while their
 * planned device bytes fit hbm_budget (and max_active) by a worker thread of
 * the server (scene creation, obstacles, uploads), off the ticks' path; each
 * tick runs one batched step (rh_batch) of every running request, retires
 * finished ones (results kept on the host) and lets admitted ones join; a
 * tick with nothing running admits on the caller's thread.  A request's
 * result does not depend on the mix.  One thread at a time may call a given
 * server's functions. */
typedef struct rh_server rh_server;
typedef struct rh_server_config {
  rh_scene_config scene;     /* f32 faces, RH_MIXED_MG with f32 V-cycles */
  uint64_t hbm_budget;       /* bytes of scene state held at once (0 = 80% of free) */
  int32_t max_active;        /* running requests at once (0 = memory-bound only) */
  int32_t block_divisor;     /* batched launch shape (rh_batch_tuning; 0 = 1) */
  uint32_t kernel_variant;
  int32_t reserved;
} rh_server_config;
typedef struct rh_request {
  rh_grid grid;
  const float* u;            /* initial faces, host (copied at submit) */
  const float* v;
  const float* w;
  const uint8_t* solid;      /* solid cell flags or NULL */
  double dt;
  int32_t steps;
  int32_t reserved;
} rh_request;
typedef struct rh_request_state {
  int32_t state;             /* 0 queued, 1 running, 2 done, 3 failed */
  rh_status status;
  int32_t steps_done;
  int32_t reserved;
  double submitted, admitted, finished;  /* seconds since server creation */
  uint64_t bytes;            /* planned device bytes while running */
  rh_step_report last;       /* report of the last step taken */
} rh_request_state;
typedef struct rh_tick_report {
  int32_t stepped;           /* requests that took a step this tick */
  int32_t admitted;          /* requests that joined the running set this tick */
  int32_t finished;          /* requests retired (done or failed) */
  int32_t active;            /* running after the tick */
  int32_t queued;
  int32_t reserved;
  int64_t cells;             /* cells stepped */
  uint64_t bytes_in_use;
  double seconds;            /* host wall time of the tick */
  double device_seconds;     /* device time of the batched step */
  double admit_seconds;      /* host time spent admitting requests during the tick */
  double build_seconds;      /* host time rebuilding the batch's launch tables */
  double wait_seconds;       /* host time waiting for the step after enqueuing it */
} rh_tick_report;
rh_status rh_server_create(const rh_server_config* config, rh_server** out);
void rh_server_destroy(rh_server* server);
rh_status rh_server_submit(rh_server* server, const rh_request* request, int64_t* id);
rh_status rh_server_tick(rh_server* server, rh_tick_report* report);
rh_status rh_server_state(rh_server* server, int64_t id, rh_request_state* out);
/* Final faces (host, scene precision f32) and pressure of a finished request. */
rh_status rh_server_result(rh_server* server, int64_t id, float* u, float* v, float* w,
                           double* pressure);
/* Planned device bytes of one request's scene (the admission figure). */
uint64_t rh_server_plan_bytes(const rh_grid* grid, const rh_scene_config* config);

/* ------------------------------------------------------- dense3d frames */
/* write_dense3d_frame (src/dense3d_sequence.rs:412-564): one accepted
 * transport state as one JSON line, every field number formatted on the
 * device exactly as the reference's `{v:.9e}` (f32) and `{v:.17e}` (f64). */
typedef struct rh_volume_stamp {
  uint64_t id;
  uint64_t version;
} rh_volume_stamp;
/* LiquidTransportView as the frame reads it: the grid's flat face and cell
 * arrays in device memory (carrier faces and tracer f32, liquid fraction and
 * pressure f64), the clocks and stamps, and whether the view is
 * transport-only (no pressure surface, pressure columns, reconstructed
 * surface or flat column MAC). */
typedef struct rh_dense3d_view {
  const float* velocity[3];
  const float* tracer;
  const double* fraction;
  const double* pressure;
  double carrier_time;
  double liquid_time;
  uint64_t carrier_generation;
  rh_volume_stamp carrier_stamp;
  rh_volume_stamp liquid_stamp;
  int32_t transport_only;
  int32_t reserved;
} rh_dense3d_view;
/* The LiquidStepReport fields a frame checks and prints (carrier.step: time,
 * dt, generation, pressure.iterations, pressure.true_residual_max,
 * actual_divergence_max; liquid: time, dt, stamp, flow and the volume
 * accounting). */
typedef struct rh_dense3d_report {
  double carrier_time;
  double carrier_dt;
  uint64_t carrier_generation;
  uint64_t pressure_iterations;
  double pressure_true_residual_max;
  double carrier_actual_divergence_max;
  double liquid_time;
  double liquid_dt;
  rh_volume_stamp liquid_stamp;
  rh_volume_stamp liquid_flow;
  double liquid_actual_divergence_max;
  double liquid_volume_before;
  double liquid_volume_after;
  double liquid_mass_after;
  double inward_boundary_volume;
  double outward_boundary_volume;
  double source_volume;
  double volume_balance_error;
  double volume_rounding_budget;
} rh_dense3d_report;
/* write_dense3d_frame (src/dense3d_sequence.rs:415-564) into device memory
 * d_out: the preflight checks in the reference's order (a failure is
 * RH_INVALID_STATE, message in rh_last_error, nothing written), then the
 * line, newline included; *bytes = its length.  report null: the
 * constructor frame (frame 0 at time 0).  A null d_out or a capacity below
 * the length returns RH_BUFFER_LIMIT with the length needed in *bytes. */
rh_status rh_dense3d_frame(const rh_grid* grid, const rh_dense3d_view* view, uint64_t frame,
                           const rh_dense3d_report* report, char* d_out, uint64_t capacity, uint64_t* bytes);
/* The frame's number format on the host: finite v with `digits` (10 or 18)
 * significant digits, NUL-terminated into out (>= 26 bytes); returns the
 * length, -1 for a non-finite v or another digit count. */
int32_t rh_format_sci(double v, int32_t digits, char* out);

#ifdef __cplusplus
}
#endif
#endif /* RHEON_GPU_H */
