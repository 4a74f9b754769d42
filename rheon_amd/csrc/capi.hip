// extern "C" entry points of include/rheon_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"
#include "handles.hpp"
#include "profiler.hpp"

using rh::DevGrid;

namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

rh_status grid_ok(const rh_grid* g) {
  if (!g) return RH_INVALID_ARGUMENT;
  rh_status s = rh::validate_grid(g, nullptr);
  if (s) return s;
  return rh::check_indexable(g);
}

size_t cells_of(const rh_grid& g) { return (size_t)g.counts[0] * g.counts[1] * g.counts[2]; }
size_t face_len(const rh_grid& g, int a) {
  size_t f = 1;
  for (int d = 0; d < 3; ++d) f *= g.counts[d] + (d == a ? 1 : 0);
  return f;
}
size_t prec_bytes(rh_precision p) { return p == RH_F32 ? 4 : 8; }

bool same_grid(const rh_grid& a, const rh_grid& b) {
  for (int d = 0; d < 3; ++d)
    if (a.counts[d] != b.counts[d] || a.spacing[d] != b.spacing[d] || a.origin[d] != b.origin[d])
      return false;
  return true;
}

void copy_str(char* dst, size_t n, const char* src) {
  std::strncpy(dst, src, n - 1);
  dst[n - 1] = '\0';
}

const char* method_id(rh_method m, rh_precision p) {
  switch (m) {
    case RH_JACOBI_PCG: return p == RH_F64 ? "jacobi-pcg-v1" : "jacobi-pcg-f32";
    case RH_RBSGS_PCG: return p == RH_F64 ? "rbsgs-pcg" : "rbsgs-pcg-f32";
    case RH_SGS_PCG: return p == RH_F64 ? "sgs-pcg-v1" : "sgs-pcg-f32";
    case RH_MG_PCG: return p == RH_F64 ? "mg-pcg" : "mg-pcg-f32";
    case RH_MIXED_MG: return p == RH_F32 ? "mixed-mg-dc" : "mg-f64-dc";
  }
  return "unknown";
}

template <typename F>
rh_status dispatch_precision(rh_precision p, F f) {
  if (p == RH_F64) return f((double*)nullptr);
  if (p == RH_F32) return f((float*)nullptr);
  return RH_INVALID_ARGUMENT;
}

}  // namespace

// --------------------------------------------------------------- basics

extern "C" int32_t rh_abi_version(void) { return RH_ABI_VERSION; }
extern "C" const char* rh_last_error(void) { return rh::last_error(); }

extern "C" rh_status rh_device_info(int32_t* count, char* arch, size_t arch_len) {
  int n = 0;
  RH_HIP(hipGetDeviceCount(&n));
  if (count) *count = n;
  if (arch && arch_len) {
    hipDeviceProp_t prop;
    RH_HIP(hipGetDeviceProperties(&prop, 0));
    copy_str(arch, arch_len, prop.gcnArchName);
  }
  return RH_OK;
}

extern "C" rh_status rh_device_alloc(size_t bytes, void** out) {
  if (!out) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  RH_HIP(hipMalloc(out, bytes ? bytes : 1));
  return RH_OK;
}
extern "C" rh_status rh_device_free(void* ptr) {
  if (ptr) RH_HIP(hipFree(ptr));
  return RH_OK;
}
extern "C" rh_status rh_copy_to_device(void* dst, const void* src, size_t bytes) {
  RH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return RH_OK;
}
extern "C" rh_status rh_copy_to_host(void* dst, const void* src, size_t bytes) {
  RH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return RH_OK;
}
extern "C" rh_status rh_device_synchronize(void) {
  RH_HIP(hipDeviceSynchronize());
  return RH_OK;
}

// Streaming copy: 16-byte vector loads and stores, grid-stride.
__global__ void __launch_bounds__(256) k_copy_bytes(uint4* __restrict__ dst,
                                                    const uint4* __restrict__ src, size_t n) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    dst[i] = src[i];
}

extern "C" rh_status rh_copy_device(void* dst, const void* src, size_t bytes) {
  if (!dst || !src || bytes % 16) return RH_INVALID_ARGUMENT;
  const size_t n = bytes / 16;
  if (n == 0) return RH_OK;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  k_copy_bytes<<<(unsigned)blocks, 256>>>((uint4*)dst, (const uint4*)src, n);
  RH_HIP(hipGetLastError());
  RH_HIP(hipDeviceSynchronize());
  return RH_OK;
}

extern "C" rh_status rh_grid_validate(const rh_grid* grid, uint64_t* cells) {
  size_t n = 0;
  rh_status s = rh::validate_grid(grid, &n);
  if (!s && cells) *cells = n;
  return s;
}

// BufferPlan::for_grid (src/geometry.rs:284-315).
extern "C" rh_status rh_buffer_plan(const rh_grid* grid, uint64_t* velocity_pair,
                                    uint64_t* tracer_pair, uint64_t* pressure_workspace,
                                    uint64_t* total) {
  size_t n = 0;
  rh_status s = rh::validate_grid(grid, &n);
  if (s) return s;
  const uint64_t faces = face_len(*grid, 0) + face_len(*grid, 1) + face_len(*grid, 2);
  const uint64_t v = faces * 8, t = (uint64_t)n * 8, p = (uint64_t)n * 48;
  if (velocity_pair) *velocity_pair = v;
  if (tracer_pair) *tracer_pair = t;
  if (pressure_workspace) *pressure_workspace = p;
  if (total) *total = v + t + p;
  return RH_OK;
}

// ------------------------------------------------------- operator kernels

extern "C" rh_status rh_apply_full(const rh_grid* grid, double density, const double* d_x,
                                   double* d_y) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_x || !d_y) return RH_INVALID_ARGUMENT;
  if ((s = rh::check_density(*grid, density))) return s;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, density);
  if ((s = sc.clear_flags())) return s;
  rh::launch_apply_full(g, d_x, d_y, sc.flags, sc.stream);
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  return rh::flags_to_status(f);
}

extern "C" rh_status rh_advect_velocity(const rh_grid* grid, rh_precision precision,
                                        const void* d_u, const void* d_v, const void* d_w,
                                        double dt, void* d_out_u, void* d_out_v, void* d_out_w) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_u || !d_v || !d_w || !d_out_u || !d_out_v || !d_out_w) return RH_INVALID_ARGUMENT;
  if (!std::isfinite(dt) || dt <= 0.0) return RH_INVALID_TIME_STEP;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, 1.0);
  if ((s = sc.clear_flags())) return s;
  s = dispatch_precision(precision, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::launch_validate<T>(g, d_u, d_v, d_w, 0, sc.flags, sc.stream);
    rh::launch_advect<T>(g, d_u, d_v, d_w, dt, d_out_u, d_out_v, d_out_w, sc.flags, sc.stream);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  return rh::flags_to_status(f);
}

extern "C" rh_status rh_advect_tracer(const rh_grid* grid, const float* d_old, const float* d_u,
                                      const float* d_v, const float* d_w, double dt,
                                      float* d_out) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_old || !d_u || !d_v || !d_w || !d_out) return RH_INVALID_ARGUMENT;
  if (!std::isfinite(dt) || dt <= 0.0) return RH_INVALID_TIME_STEP;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, 1.0);
  if ((s = sc.clear_flags())) return s;
  rh::launch_validate<float>(g, d_u, d_v, d_w, 0, sc.flags, sc.stream);
  rh::launch_advect_tracer(g, d_old, d_u, d_v, d_w, dt, d_out, sc.flags, sc.stream);
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  return rh::flags_to_status(f);
}

extern "C" rh_status rh_build_rhs(const rh_grid* grid, rh_precision precision, const void* d_u,
                                  const void* d_v, const void* d_w, double dt, double* d_rhs) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_u || !d_v || !d_w || !d_rhs) return RH_INVALID_ARGUMENT;
  if (!std::isfinite(dt) || dt <= 0.0) return RH_INVALID_TIME_STEP;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, 1.0);
  if ((s = sc.clear_flags())) return s;
  s = dispatch_precision(precision, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::launch_validate<T>(g, d_u, d_v, d_w, 1, sc.flags, sc.stream);
    rh::launch_rhs<T>(g, d_u, d_v, d_w, dt, d_rhs, sc.flags, sc.stream);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  return rh::flags_to_status(f);
}

extern "C" rh_status rh_correct_velocity(const rh_grid* grid, rh_precision precision,
                                         double density, const double* d_pressure, double dt,
                                         void* d_u, void* d_v, void* d_w) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_pressure || !d_u || !d_v || !d_w) return RH_INVALID_ARGUMENT;
  if (!std::isfinite(dt) || dt <= 0.0) return RH_INVALID_TIME_STEP;
  if ((s = rh::check_density(*grid, density))) return s;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, density);
  if ((s = sc.clear_flags())) return s;
  s = dispatch_precision(precision, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::launch_validate<T>(g, d_u, d_v, d_w, 1, sc.flags, sc.stream);
    return RH_OK;
  });
  if (s) return s;
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  if (f) return rh::flags_to_status(f);
  s = dispatch_precision(precision, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::launch_correct<T>(g, d_pressure, dt, d_u, d_v, d_w, sc.flags, sc.stream);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipGetLastError());
  if ((s = sc.read_flags(&f))) return s;
  return rh::flags_to_status(f);
}

extern "C" rh_status rh_divergence_max(const rh_grid* grid, rh_precision precision,
                                       const void* d_u, const void* d_v, const void* d_w,
                                       double* out_max) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_u || !d_v || !d_w || !out_max) return RH_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, 1.0);
  if ((s = sc.clear_flags())) return s;
  s = dispatch_precision(precision, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::launch_validate<T>(g, d_u, d_v, d_w, 1, sc.flags, sc.stream);
    rh::launch_divergence_max<T>(g, d_u, d_v, d_w, sc.reduce(0), sc.stream);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  if (f) return rh::flags_to_status(f);
  return sc.read_scalars(out_max, 1, 0);
}

extern "C" rh_status rh_velocity_rate(const rh_grid* grid, rh_precision precision,
                                      const void* d_u, const void* d_v, const void* d_w,
                                      double* out_rate) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_u || !d_v || !d_w || !out_rate) return RH_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, 1.0);
  s = dispatch_precision(precision, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::launch_velocity_rate<T>(g, d_u, d_v, d_w, sc.reduce(0), sc.stream);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipGetLastError());
  return sc.read_scalars(out_rate, 1, 0);
}

extern "C" rh_status rh_precondition(const rh_grid* grid, double density, rh_method method,
                                     double* d_r, double* d_z, double* out_rz) {
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (!d_r || !d_z) return RH_INVALID_ARGUMENT;
  if (method != RH_JACOBI_PCG && method != RH_RBSGS_PCG && method != RH_SGS_PCG) return RH_UNSUPPORTED;
  if ((s = rh::check_density(*grid, density))) return s;
  std::unique_lock<std::mutex> lock;
  rh::Scratch& sc = rh::global_scratch(lock);
  if ((s = sc.init())) return s;
  const DevGrid g = rh::make_dev_grid(*grid, density);
  std::vector<rh::SgsTile> tiles;
  std::vector<int> wave_start;
  rh::SgsTile* d_tiles = nullptr;
  if (method == RH_SGS_PCG) {
    rh::sgs_plan(g, tiles, wave_start);
    RH_HIP(hipMalloc(&d_tiles, sizeof(rh::SgsTile) * tiles.size()));
    RH_HIP(hipMemcpy(d_tiles, tiles.data(), sizeof(rh::SgsTile) * tiles.size(), hipMemcpyHostToDevice));
  }
  RH_HIP(hipMemsetAsync(sc.ctl, 0, sizeof(rh::PcgControl), sc.stream));
  rh::launch_precondition(g, method, d_r, d_z, sc.ctl, sc.pdd, sc.pmx, sc.stream, d_tiles, wave_start.data(),
                          (int)wave_start.size() - 1);
  const hipError_t le = hipGetLastError();
  rh::PcgControl c;
  const hipError_t ce = hipMemcpyAsync(sc.host, sc.ctl, sizeof(c), hipMemcpyDeviceToHost, sc.stream);
  const hipError_t se = hipStreamSynchronize(sc.stream);
  if (d_tiles) (void)hipFree(d_tiles);
  RH_HIP(le);
  RH_HIP(ce);
  RH_HIP(se);
  std::memcpy(&c, sc.host, sizeof(c));
  if (c.done && c.status == RH_BREAKDOWN) return RH_BREAKDOWN;
  if (out_rz) *out_rz = c.rz[1];
  return RH_OK;
}

// ------------------------------------------------------- device solver


extern "C" rh_status rh_solver_create(const rh_grid* grid, const rh_solver_config* config,
                                      rh_solver** out) {
  if (!out || !config) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  rh_status s = grid_ok(grid);
  if (s) return s;
  auto* solver = new (std::nothrow) rh_solver();
  if (!solver) return RH_DEVICE_ERROR;
  solver->impl.reset(rh::make_solver(config->method));
  if (!solver->impl) {
    delete solver;
    return RH_UNSUPPORTED;
  }
  if ((s = solver->impl->create(*grid, *config))) {
    delete solver;
    return s;
  }
  *out = solver;
  return RH_OK;
}

extern "C" void rh_solver_destroy(rh_solver* solver) { delete solver; }

extern "C" uint64_t rh_solver_allocated_bytes(const rh_solver* solver) {
  return solver ? solver->impl->allocated_bytes() : 0;
}

extern "C" rh_status rh_solver_solve_rhs(rh_solver* solver, double density, const double* d_rhs,
                                         double dt, const rh_pressure_settings* settings,
                                         double* d_pressure, rh_pressure_report* report,
                                         rh_cancel_fn cancel, void* cancel_ctx) {
  if (!solver || !settings) return RH_INVALID_ARGUMENT;
  return solver->impl->solve(density, d_rhs, dt, *settings, d_pressure, report, cancel, cancel_ctx);
}

extern "C" rh_status rh_solver_solve_rhs_initial(rh_solver* solver, double density, const double* d_rhs,
                                                 const double* d_initial, double dt,
                                                 const rh_pressure_settings* settings, double* d_pressure,
                                                 rh_pressure_report* report, rh_cancel_fn cancel,
                                                 void* cancel_ctx) {
  if (!solver || !settings || !d_initial) return RH_INVALID_ARGUMENT;
  solver->impl->set_seed(d_initial);
  const rh_status s = solver->impl->solve(density, d_rhs, dt, *settings, d_pressure, report, cancel, cancel_ctx);
  solver->impl->set_seed(nullptr);
  return s;
}

extern "C" rh_status rh_solver_set_solid(rh_solver* solver, const uint8_t* solid, size_t len) {
  if (!solver) return RH_INVALID_ARGUMENT;
  if (solid && len != cells_of(solver->impl->grid())) return RH_INVALID_ARGUMENT;
  return solver->impl->set_solid(solid);
}

// ------------------------------------------------------------- provider


extern "C" rh_status rh_provider_create(const rh_provider_config* config, rh_provider** out) {
  if (!config || !out) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  auto* p = new (std::nothrow) rh_provider();
  if (!p) return RH_DEVICE_ERROR;
  p->config = *config;
  *out = p;
  return RH_OK;
}

extern "C" void rh_provider_destroy(rh_provider* provider) { delete provider; }

extern "C" rh_status rh_provider_metadata(const rh_provider* provider, rh_solver_metadata* out) {
  if (!provider || !out) return RH_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  copy_str(out->id, sizeof(out->id),
           provider->supplied ? kSuppliedPressureId
                              : method_id(provider->config.solver.method, provider->config.solver.precision));
  copy_str(out->version, sizeof(out->version), "1");
  copy_str(out->stage, sizeof(out->stage), "pressure");
  for (int d = 0; d < 3; ++d) out->shape_xyz[d] = provider->config.shape_xyz[d];
  copy_str(out->layout, sizeof(out->layout), "x-fastest-cell-centered");
  copy_str(out->scalar, sizeof(out->scalar), "f64");
  copy_str(out->units_in, sizeof(out->units_in), "m^3/s^2");
  copy_str(out->units_out, sizeof(out->units_out), "Pa");
  copy_str(out->domain, sizeof(out->domain),
           provider->supplied ? "sealed box supplied-array witness only" : "sealed-constant-density-mac-box");
  return RH_OK;
}

static rh_status provider_supports(const rh_provider* provider, const rh_pressure_request* req,
                                   const char** reason) {
  auto why = [&](const char* r) {
    if (reason) *reason = r;
    return RH_UNSUPPORTED;
  };
  if (reason) *reason = nullptr;
  if (!provider || !req || !req->grid) return RH_INVALID_ARGUMENT;
  if (provider->supplied) return RH_OK;  // SuppliedPressure::supports (projection.rs:581-583)
  if (rh::validate_grid(req->grid, nullptr)) return why("invalid grid geometry");
  if (rh::check_indexable(req->grid)) return why("grid exceeds 32-bit device indexing");
  if (rh::check_density(*req->grid, req->density)) return why("density must be finite and positive");
  if (!std::isfinite(req->dt) || req->dt <= 0.0) return why("time step must be finite and positive");
  if (req->rhs_len != cells_of(*req->grid)) return why("rhs length does not match the grid");
  const uint64_t* shape = provider->config.shape_xyz;
  if ((shape[0] || shape[1] || shape[2]) &&
      (shape[0] != req->grid->counts[0] || shape[1] != req->grid->counts[1] || shape[2] != req->grid->counts[2]))
    return why("grid shape differs from the provider's shape");
  const rh_method m = provider->config.solver.method;
  if (m != RH_JACOBI_PCG && m != RH_RBSGS_PCG && m != RH_SGS_PCG && m != RH_MIXED_MG)
    return why("method not available in this build");
  return RH_OK;
}

extern "C" rh_status rh_provider_supports(const rh_provider* provider,
                                          const rh_pressure_request* request, const char** reason) {
  return provider_supports(provider, request, reason);
}

extern "C" rh_status rh_provider_predict(rh_provider* provider, const rh_pressure_request* req,
                                         double* pressure, size_t pressure_len, rh_cancel_fn cancel,
                                         void* cancel_ctx, const char** reason) {
  if (reason) *reason = nullptr;
  rh_status s = provider_supports(provider, req, reason);
  if (s) return s;
  if (!pressure || pressure_len != req->rhs_len || !req->rhs) return RH_INVALID_ARGUMENT;
  const size_t n = req->rhs_len;
  if (provider->supplied) {  // SuppliedPressure::predict: the supplied values (projection.rs:584-592)
    if (provider->supplied_len != n) return RH_INVALID_ARGUMENT;
    std::memcpy(pressure, provider->supplied, n * sizeof(double));
    return RH_OK;
  }
  auto exec = [&](const char* r) {
    if (reason) *reason = r;
    return RH_EXECUTION;
  };
  if (!provider->solver || !same_grid(provider->grid, *req->grid)) {
    provider->solver.reset();
    rh_solver* raw = nullptr;
    if (rh_solver_create(req->grid, &provider->config.solver, &raw)) return exec("device solver creation failed");
    provider->solver.reset(raw);
    provider->grid = *req->grid;
  }
  if (provider->cap < n) {
    if (provider->d_rhs) (void)hipFree(provider->d_rhs);
    if (provider->d_p) (void)hipFree(provider->d_p);
    provider->d_rhs = provider->d_p = nullptr;
    provider->cap = 0;
    if (hipMalloc(&provider->d_rhs, n * 8) != hipSuccess || hipMalloc(&provider->d_p, n * 8) != hipSuccess)
      return exec("device allocation failed");
    provider->cap = n;
  }
  if (hipMemcpy(provider->d_rhs, req->rhs, n * 8, hipMemcpyHostToDevice) != hipSuccess)
    return exec("host-to-device copy failed");
  rh_pressure_report rep;
  s = rh_solver_solve_rhs(provider->solver.get(), req->density, provider->d_rhs, req->dt,
                          &provider->config.settings, provider->d_p, &rep, cancel, cancel_ctx);
  if (s == RH_CANCELLED) return RH_CANCELLED;
  if (s != RH_OK) return exec("device pressure solve did not converge");
  if (hipMemcpy(pressure, provider->d_p, n * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return exec("device-to-host copy failed");
  return RH_OK;
}

// ---------------------------------------------------- provider workspace


extern "C" rh_status rh_provider_workspace_create(const rh_grid* grid, uint64_t limit,
                                                  rh_method fallback,
                                                  rh_provider_workspace** out) {
  if (!out) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  rh_status s = grid_ok(grid);
  if (s) return s;
  if (fallback != RH_JACOBI_PCG && fallback != RH_RBSGS_PCG && fallback != RH_SGS_PCG) return RH_UNSUPPORTED;
  const size_t n = cells_of(*grid);
  // ProviderWorkspace::new (src/solver_provider.rs:105-144): 16 N extra plus
  // the 48 N reference workspace must fit the limit.
  if (limit && (uint64_t)n * 64 > limit) return RH_BUFFER_LIMIT;
  auto ws = std::unique_ptr<rh_provider_workspace>(new (std::nothrow) rh_provider_workspace());
  if (!ws) return RH_DEVICE_ERROR;
  ws->grid = *grid;
  ws->fallback = fallback;
  rh_solver_config cfg{};
  cfg.method = fallback;
  cfg.precision = RH_F64;
  rh_solver* raw = nullptr;
  if ((s = rh_solver_create(grid, &cfg, &raw))) return s;
  ws->solver.reset(raw);
  RH_HIP(hipMalloc(&ws->d_rhs, n * 8));
  RH_HIP(hipMalloc(&ws->d_p, n * 8));
  try {
    ws->pressure.assign(n, 0.0);
  } catch (...) {
    return RH_BUFFER_LIMIT;
  }
  *out = ws.release();
  return RH_OK;
}

extern "C" void rh_provider_workspace_destroy(rh_provider_workspace* ws) { delete ws; }

extern "C" rh_status rh_provider_workspace_pressure(const rh_provider_workspace* ws, double* out,
                                                    size_t len) {
  if (!ws || !out || len != ws->pressure.size()) return RH_INVALID_ARGUMENT;
  std::memcpy(out, ws->pressure.data(), len * 8);
  return RH_OK;
}

// ProviderWorkspace::solve_observed (src/solver_provider.rs:160-327).
extern "C" rh_status rh_provider_workspace_solve(rh_provider_workspace* ws,
                                                 const rh_pressure_request* req,
                                                 const rh_pressure_settings* settings,
                                                 rh_provider* preferred, rh_cancel_fn cancel,
                                                 void* cancel_ctx, rh_provider_report* report) {
  const auto start = Clock::now();
  if (!ws || !req || !req->grid || !settings || !report) return RH_INVALID_ARGUMENT;
  std::memset(report, 0, sizeof(*report));
  auto cancelled = [&]() { return cancel && cancel(cancel_ctx); };
  if (!same_grid(*req->grid, ws->grid)) return RH_INVALID_ARGUMENT;  // LengthMismatch
  rh_status s = rh::check_density(ws->grid, req->density);
  if (s) return s;
  const size_t n = ws->pressure.size();
  if (!req->rhs || req->rhs_len != n) return RH_INVALID_ARGUMENT;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(req->rhs[i])) return RH_NONFINITE_INPUT;
  if (!std::isfinite(req->dt) || req->dt <= 0.0 || !std::isfinite(settings->relative_residual) ||
      settings->relative_residual < 0.0 || !std::isfinite(settings->absolute_residual) ||
      settings->absolute_residual < 0.0 || !std::isfinite(settings->divergence_limit) ||
      settings->divergence_limit < 0.0)
    return RH_INVALID_SETTINGS;
  rh::Scratch& sc = ws->solver->impl->scratch();
  const DevGrid g = rh::make_dev_grid(ws->grid, req->density);
  // Compatibility sums and threshold on the device (:186-206).
  RH_HIP(hipMemcpyAsync(ws->d_rhs, req->rhs, n * 8, hipMemcpyHostToDevice, sc.stream));
  rh::PcgControl h{};
  h.rel = settings->relative_residual;
  h.abs_tol = settings->absolute_residual;
  std::memcpy(sc.host, &h, sizeof(h));
  RH_HIP(hipMemcpyAsync(sc.ctl, sc.host, sizeof(h), hipMemcpyHostToDevice, sc.stream));
  rh::launch_rhs_sums(g, ws->d_rhs, sc.ctl, sc.pdd, sc.pmx, sc.stream);
  RH_HIP(hipMemcpyAsync(sc.host, sc.ctl, sizeof(h), hipMemcpyDeviceToHost, sc.stream));
  RH_HIP(hipStreamSynchronize(sc.stream));
  std::memcpy(&h, sc.host, sizeof(h));
  if (h.done) {
    report->accepted.error_value = h.error_value;
    return (rh_status)h.status;
  }
  const double threshold = h.threshold;
  if (cancelled()) return RH_CANCELLED;
  copy_str(report->selected, sizeof(report->selected), method_id(ws->fallback, RH_F64));
  if (preferred) {
    report->has_attempted = 1;
    rh_solver_metadata& meta = report->attempted;
    rh_provider_metadata(preferred, &meta);
    const char* reason = nullptr;
    rh_status attempt = rh_provider_supports(preferred, req, &reason);
    if (attempt == RH_OK) {
      std::fill(ws->pressure.begin(), ws->pressure.end(), std::numeric_limits<double>::quiet_NaN());
      const auto t0 = Clock::now();
      attempt = rh_provider_predict(preferred, req, ws->pressure.data(), n, cancel, cancel_ctx, &reason);
      report->inference = seconds_since(t0);
    }
    if (cancelled() || attempt == RH_CANCELLED) return RH_CANCELLED;
    const auto t1 = Clock::now();
    rh_status qualified = attempt;
    if (qualified == RH_OK) {
      for (size_t i = 0; i < n && qualified == RH_OK; ++i)
        if (!std::isfinite(ws->pressure[i])) qualified = RH_NONFINITE_OUTPUT;
      if (qualified == RH_OK && ws->pressure[0] != 0.0) qualified = RH_NONZERO_GAUGE;
    }
    if (qualified == RH_OK) {
      RH_HIP(hipMemcpyAsync(ws->d_p, ws->pressure.data(), n * 8, hipMemcpyHostToDevice, sc.stream));
      rh::launch_residual_report(g, ws->d_p, ws->d_rhs, sc.reduce(0), sc.stream);
      RH_HIP(hipGetLastError());
      double out[3];
      if ((s = sc.read_scalars(out, 3, 0))) return s;
      if (out[2] != 0.0) {
        qualified = RH_NONFINITE_OUTPUT;
      } else {
        const double l2 = std::sqrt(out[0]);
        const double maximum = out[1];
        const double divergence = maximum * req->dt / g.volume;
        report->has_candidate = 1;
        report->candidate.iterations = 0;
        report->candidate.true_residual_l2 = l2;
        report->candidate.true_residual_max = maximum;
        report->candidate.predicted_divergence_max = divergence;
        report->candidate.threshold = threshold;
        if (!std::isfinite(l2) || !std::isfinite(divergence) || l2 > threshold ||
            divergence > settings->divergence_limit)
          qualified = RH_QUALITY_GATE;
      }
    }
    report->qualification = seconds_since(t1);
    if (cancelled()) return RH_CANCELLED;
    if (qualified == RH_OK) {
      copy_str(report->selected, sizeof(report->selected), meta.id);
      report->accepted = report->candidate;
      report->total = seconds_since(start);
      return RH_OK;
    }
    report->rejection = qualified;
  }
  // Native fallback: the device jacobi-pcg-v1 (:310-318).
  const auto t2 = Clock::now();
  s = rh_solver_solve_rhs(ws->solver.get(), req->density, ws->d_rhs, req->dt, settings, ws->d_p,
                          &report->accepted, cancel, cancel_ctx);
  if (s != RH_OK) return s;
  RH_HIP(hipMemcpy(ws->pressure.data(), ws->d_p, n * 8, hipMemcpyDeviceToHost));
  report->fallback = seconds_since(t2);
  report->total = seconds_since(start);
  return RH_OK;
}

// ----------------------------------------------------------------- scene


extern "C" rh_status rh_scene_create(const rh_grid* grid, const rh_scene_config* config,
                                     rh_scene** out) {
  if (!config || !out) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  rh_status s = grid_ok(grid);
  if (s) return s;
  if ((s = rh::check_density(*grid, config->density))) return s;
  if (config->precision != RH_F64 && config->precision != RH_F32) return RH_INVALID_ARGUMENT;
  auto sc = std::unique_ptr<rh_scene>(new (std::nothrow) rh_scene());
  if (!sc) return RH_DEVICE_ERROR;
  sc->grid = *grid;
  sc->config = *config;
  sc->n = cells_of(*grid);
  const size_t fb = prec_bytes(config->precision);
  size_t total = 0;
  auto round = [](size_t b) { return (b + 255) / 256 * 256; };
  for (int a = 0; a < 3; ++a) {
    sc->flen[a] = face_len(*grid, a);
    total += 2 * round(sc->flen[a] * fb);
  }
  total += 2 * round(sc->n * 8);
  RH_HIP(hipMalloc(&sc->arena, total));
  RH_HIP(hipMemset(sc->arena, 0, total));
  char* p = (char*)sc->arena;
  for (int a = 0; a < 3; ++a) {
    sc->acc[a] = p;
    p += round(sc->flen[a] * fb);
    sc->cand[a] = p;
    p += round(sc->flen[a] * fb);
  }
  sc->rhs = (double*)p;
  p += round(sc->n * 8);
  sc->p = (double*)p;
  rh_solver* raw = nullptr;
  if ((s = rh_solver_create(grid, &config->solver, &raw))) return s;
  sc->solver.reset(raw);
  for (auto& e : sc->ev) RH_HIP(hipEventCreate(&e));
  sc->bytes = total + raw->impl->allocated_bytes();
  if (config->solver.memory_limit && sc->bytes > config->solver.memory_limit) return RH_BUFFER_LIMIT;
  *out = sc.release();
  return RH_OK;
}

extern "C" void rh_scene_destroy(rh_scene* scene) { delete scene; }

extern "C" uint64_t rh_scene_allocated_bytes(const rh_scene* scene) {
  return scene ? scene->bytes : 0;
}

// Obstacle faces of the accepted velocity are closed (exactly 0).
static rh_status zero_blocked(rh_scene* sc) {
  const DevGrid g = sc->solver->impl->dev_grid(sc->config.density);
  if (!g.mask) return RH_OK;
  hipStream_t st = sc->solver->impl->scratch().stream;
  rh_status s = dispatch_precision(sc->config.precision, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::launch_zero_blocked<T>(g, sc->acc[0], sc->acc[1], sc->acc[2], st);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipGetLastError());
  RH_HIP(hipStreamSynchronize(st));
  return RH_OK;
}

extern "C" rh_status rh_scene_set_velocity(rh_scene* sc, const void* u, const void* v,
                                           const void* w) {
  if (!sc || !u || !v || !w) return RH_INVALID_ARGUMENT;
  const void* src[3] = {u, v, w};
  const size_t fb = prec_bytes(sc->config.precision);
  sc->acc_checked = false;
  sc->rate_known = false;
  sc->p_state = 0;  // a new state: no accepted pressure yet
  for (int a = 0; a < 3; ++a)
    RH_HIP(hipMemcpy(sc->acc[a], src[a], sc->flen[a] * fb, hipMemcpyHostToDevice));
  return zero_blocked(sc);
}

extern "C" rh_status rh_scene_set_solid(rh_scene* sc, const uint8_t* solid, size_t len) {
  if (!sc) return RH_INVALID_ARGUMENT;
  if (solid && len != sc->n) return RH_INVALID_ARGUMENT;
  sc->acc_checked = false;
  sc->rate_known = false;
  sc->p_state = 0;
  rh_status s = sc->solver->impl->set_solid(solid);
  if (s) return s;
  return zero_blocked(sc);
}

extern "C" rh_status rh_scene_get_velocity(const rh_scene* sc, void* u, void* v, void* w) {
  if (!sc || !u || !v || !w) return RH_INVALID_ARGUMENT;
  void* dst[3] = {u, v, w};
  const size_t fb = prec_bytes(sc->config.precision);
  for (int a = 0; a < 3; ++a)
    RH_HIP(hipMemcpy(dst[a], sc->acc[a], sc->flen[a] * fb, hipMemcpyDeviceToHost));
  return RH_OK;
}

extern "C" rh_status rh_scene_get_pressure(const rh_scene* sc, double* out, size_t len) {
  if (!sc || !out || len != sc->n) return RH_INVALID_ARGUMENT;
  RH_HIP(hipMemcpy(out, sc->p, len * 8, hipMemcpyDeviceToHost));
  return RH_OK;
}

extern "C" rh_status rh_scene_device_velocity(rh_scene* sc, void** d_u, void** d_v, void** d_w) {
  if (!sc || !d_u || !d_v || !d_w) return RH_INVALID_ARGUMENT;
  sc->external = true;
  sc->rate_known = false;
  *d_u = sc->acc[0];
  *d_v = sc->acc[1];
  *d_w = sc->acc[2];
  return RH_OK;
}

// Simulation::prepare_impl carrier sequence (src/simulation.rs:988-1331):
// advect -> build_rhs -> pressure solve -> correct_candidate_in_place ->
// actual divergence gate -> Courant gate -> tracer (advect_tracer with the
// candidate velocity, range check, integral) -> commit swap (:805-811).
static rh_status scene_step(rh_scene* sc, double dt, rh_step_report* report, rh_cancel_fn cancel,
                            void* cancel_ctx);

extern "C" rh_status rh_scene_step(rh_scene* sc, double dt, rh_step_report* report,
                                   rh_cancel_fn cancel, void* cancel_ctx) {
  if (!sc) return RH_INVALID_ARGUMENT;
  sc->p_written = false;
  const rh_status s = scene_step(sc, dt, report, cancel, cancel_ctx);
  if (s == RH_OK)
    sc->p_state = 1;  // p holds the accepted pressure
  else if (sc->p_state == 1 && sc->p_written)
    sc->p_state = 0;  // the failed attempt's solve overwrote it
  return s;
}

extern "C" rh_status rh_scene_set_initialization(rh_scene* sc, rh_init_policy policy) {
  if (!sc) return RH_INVALID_ARGUMENT;
  if (policy != RH_INIT_COLD && policy != RH_INIT_ACCEPTED_PREVIOUS) return RH_INVALID_ARGUMENT;
  if (sc->batched && policy != RH_INIT_COLD) return RH_UNSUPPORTED;
  sc->init_policy = policy;
  return RH_OK;
}

static rh_status scene_step(rh_scene* sc, double dt, rh_step_report* report, rh_cancel_fn cancel,
                            void* cancel_ctx) {
  rh_step_report local;
  rh_step_report* rep = report ? report : &local;
  std::memset(rep, 0, sizeof(*rep));
  rep->dt = dt;
  rep->time = sc->time;
  rep->generation = sc->generation;
  if (!std::isfinite(dt) || dt <= 0.0) return RH_INVALID_TIME_STEP;
  if (cancel && cancel(cancel_ctx)) return RH_CANCELLED;
  rh::Scratch& scr = sc->solver->impl->scratch();
  hipStream_t st = scr.stream;
  const DevGrid g = sc->solver->impl->dev_grid(sc->config.density);
  const rh_precision prec = sc->config.precision;
  rh_status s = scr.clear_flags();
  if (s) return s;
  RH_HIP(hipEventRecord(sc->ev[0], st));
  s = dispatch_precision(prec, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    if (!sc->acc_checked || sc->external)
      rh::launch_validate<T>(g, sc->acc[0], sc->acc[1], sc->acc[2], 0, scr.flags, st);
    rh::prof::Scope ps(rh::prof::kAdvect, st, -1);
    rh::launch_advect<T>(g, sc->acc[0], sc->acc[1], sc->acc[2], dt, sc->cand[0], sc->cand[1],
                         sc->cand[2], scr.flags, st);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipEventRecord(sc->ev[1], st));
  const bool mg_method = sc->config.solver.method == RH_MIXED_MG;
  s = dispatch_precision(prec, [&](auto* tag) -> rh_status {
    using T = std::remove_pointer_t<decltype(tag)>;
    rh::prof::Scope ps(rh::prof::kRhs, st, -1);
    // a multigrid solve takes the RHS's compatibility partials from this pass
    const rh::Reduce sums = scr.reduce(0);
    rh::launch_rhs<T>(g, sc->cand[0], sc->cand[1], sc->cand[2], dt, sc->rhs, scr.flags, st, true,
                      mg_method ? &sums : nullptr);
    return RH_OK;
  });
  if (s) return s;
  RH_HIP(hipEventRecord(sc->ev[2], st));
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  // Faces this scene validated or committed itself advect to finite
  // candidates with zero walls, so their check (flags[0]) is read with the
  // step's final read-back instead of here, where it would hold the solve
  // back by a host round trip; a failed check still decides the step's
  // status before anything the solve reports.  Other faces are checked
  // here, before the solve, as in the reference.
  const bool defer = sc->acc_checked && !sc->external;
  if (!defer) {
    if ((s = scr.read_flags(&f))) return s;
    if (f) return rh::flags_to_status(f);
  }
  // AcceptedPreviousV1: seed with the last committed step's pressure, kept
  // aside first so that a failed attempt does not lose it.
  const double* seed = nullptr;
  if (sc->init_policy == RH_INIT_ACCEPTED_PREVIOUS && sc->p_state != 0) {
    if (!sc->p_acc) {
      RH_HIP(hipMalloc(&sc->p_acc, sc->n * sizeof(double)));
      sc->bytes += sc->n * sizeof(double);
    }
    if (sc->p_state == 1) {
      RH_HIP(hipMemcpyAsync(sc->p_acc, sc->p, sc->n * sizeof(double), hipMemcpyDeviceToDevice, st));
      sc->p_state = 2;
    }
    seed = sc->p_acc;
  }
  sc->p_written = true;
  // correct_candidate_in_place and the gates' sums in one pass, out of
  // place into the spare faces (which become the candidate on success),
  // sums into result slots 8..13 and flags into flags[1]; a multigrid solve
  // enqueues it behind its own kernels, before its read-back, which then
  // brings these results back too
  if (!sc->spare_base) {
    const size_t es = prec_bytes(prec);
    RH_HIP(hipMalloc(&sc->spare_base, es * (sc->flen[0] + sc->flen[1] + sc->flen[2])));
    sc->bytes += es * (sc->flen[0] + sc->flen[1] + sc->flen[2]);
    sc->spare[0] = sc->spare_base;
    sc->spare[1] = (char*)sc->spare_base + es * sc->flen[0];
    sc->spare[2] = (char*)sc->spare_base + es * (sc->flen[0] + sc->flen[1]);
  }
  constexpr int kSlot = 8;
  bool corrected = false;
  auto enqueue_correct = [&]() {
    (void)hipEventRecord(sc->ev[3], st);
    (void)dispatch_precision(prec, [&](auto* tag) -> rh_status {
      using T = std::remove_pointer_t<decltype(tag)>;
      rh::prof::Scope ps(rh::prof::kCorrect, st, -1);
      rh::Reduce red = scr.reduce(kSlot);
      red.flags = scr.flags + 1;  // the correction's own word: flags[0] may hold the deferred check
      rh::launch_correct_divergence_rate<T>(g, sc->p, dt, sc->cand, sc->spare, red, st);
      return RH_OK;
    });
    (void)hipEventRecord(sc->ev[4], st);
    corrected = true;
  };
  if (mg_method) {
    sc->solver->impl->rhs_sums_published(sc->rhs);
    sc->solver->impl->set_solve_tail(enqueue_correct);
  }
  if (seed)
    s = rh_solver_solve_rhs_initial(sc->solver.get(), sc->config.density, sc->rhs, seed, dt, &sc->config.pressure,
                                    sc->p, &rep->pressure, cancel, cancel_ctx);
  else
    s = rh_solver_solve_rhs(sc->solver.get(), sc->config.density, sc->rhs, dt, &sc->config.pressure, sc->p,
                            &rep->pressure, cancel, cancel_ctx);
  sc->solver->impl->set_solve_tail(nullptr);  // a solve that returned before taking it leaves it set
  if (s) {
    if (defer && scr.read_flags(&f) == RH_OK && f) {  // the candidate's check comes first
      std::memset(&rep->pressure, 0, sizeof(rep->pressure));
      return rh::flags_to_status(f);
    }
    return s;
  }
  // the candidate's deferred check, the correction's flags and [0]
  // max|div u|, [1] rate, [2..4] component maxima, [5] sum of squared faces:
  // from the solve's read-back, or one read-back of their own
  unsigned fw[2];
  double slots[kSlot + 6];
  const double* sv = slots + kSlot;
  if (corrected) {
    RH_HIP(hipGetLastError());
    scr.last_readback(fw, slots + kSlot, kSlot, 6);
  } else {
    enqueue_correct();
    RH_HIP(hipGetLastError());
    if ((s = scr.read_flags_scalars(fw, slots, kSlot + 6))) return s;
  }
  for (int a = 0; a < 3; ++a) std::swap(sc->cand[a], sc->spare[a]);
  if (fw[0]) {
    std::memset(&rep->pressure, 0, sizeof(rep->pressure));
    return rh::flags_to_status(fw[0]);
  }
  if (fw[1]) return rh::flags_to_status(fw[1]);
  if (rh::prof::enabled()) rh::prof::settle(INT32_MAX);
  float ms[4] = {0, 0, 0, 0}, total = 0;
  for (int i = 0; i < 4; ++i) RH_HIP(hipEventElapsedTime(&ms[i], sc->ev[i], sc->ev[i + 1]));
  RH_HIP(hipEventElapsedTime(&total, sc->ev[0], sc->ev[4]));
  rep->advect_seconds = ms[0] * 1e-3;
  rep->rhs_seconds = ms[1] * 1e-3;
  rep->solve_seconds = ms[2] * 1e-3;
  rep->correct_seconds = ms[3] * 1e-3;
  rep->seconds = total * 1e-3;
  rep->actual_divergence_max = sv[0];
  rep->courant = dt * sv[1];
  if (sv[0] > sc->config.actual_divergence_limit) return RH_DIVERGENCE_LIMIT;
  if (!std::isfinite(sv[1]) || !std::isfinite(rep->courant)) return RH_ARITHMETIC_FAILURE;
  if (sc->config.max_courant > 0.0 && rep->courant > sc->config.max_courant) return RH_COURANT_LIMIT;
  if (cancel && cancel(cancel_ctx)) return RH_CANCELLED;  // StepStage::BeforeTracer
  const double volume = g.volume;
  double tracer_sum = 0.0;
  if (sc->tracer[0]) {
    // TracerSupport (src/advection.rs:119-153): nonzero bounding box of the
    // accepted tracer grown by the largest RK2 displacement plus two cells;
    // the whole box when a displaced coordinate would not be finite.
    rh::launch_tracer_bounds(g, sc->tracer[0], scr.counters, scr.flags, st);
    unsigned b[6];
    RH_HIP(hipMemcpyAsync(scr.host, scr.counters, sizeof(b), hipMemcpyDeviceToHost, st));
    RH_HIP(hipStreamSynchronize(st));
    std::memcpy(b, scr.host, sizeof(b));
    if ((s = scr.read_flags(&f))) return s;
    if (f) return rh::flags_to_status(f);
    const int n[3] = {g.nx, g.ny, g.nz};
    rh::TracerBox box{{0, 0, 0}, {n[0], n[1], n[2]}};
    bool dense = false;
    double margin[3];
    for (int d = 0; d < 3; ++d) {
      const double displacement = dt * sv[2 + d];
      const double world = std::fmax(std::fabs(g.origin[d]), std::fabs(g.upper[d]));
      if (!std::isfinite(world + displacement)) dense = true;
      margin[d] = std::ceil(displacement / g.h[d]) + 2.0;
    }
    if (!dense) {
      const bool empty = b[3] == 0 && b[4] == 0 && b[5] == 0;
      for (int d = 0; d < 3; ++d) {
        if (empty) {
          box.lo[d] = n[d];
          box.hi[d] = 0;
          continue;
        }
        const double lo = (double)b[d] - margin[d], hi = (double)b[3 + d] + margin[d];
        box.lo[d] = lo <= 0.0 ? 0 : (int)lo;
        box.hi[d] = hi >= (double)n[d] ? n[d] : (int)hi;
      }
    }
    rh::launch_advect_tracer(g, sc->tracer[0], (const float*)sc->cand[0], (const float*)sc->cand[1],
                             (const float*)sc->cand[2], dt, sc->tracer[1], scr.flags, st, &box);
    rh::launch_tracer_sum(g, sc->tracer[1], scr.reduce(8), st);
    RH_HIP(hipGetLastError());
    if ((s = scr.read_flags(&f))) return s;
    if (f) return rh::flags_to_status(f);
    if ((s = scr.read_scalars(&tracer_sum, 1, 8))) return s;
  }
  rep->tracer_integral = tracer_sum * volume;
  rep->kinetic_energy = 0.5 * sc->config.density * volume * sv[5];
  if (!std::isfinite(rep->tracer_integral) || !std::isfinite(rep->kinetic_energy)) return RH_ARITHMETIC_FAILURE;
  const double next_time = sc->time + dt;
  if (cancel && cancel(cancel_ctx)) return RH_CANCELLED;  // StepStage::BeforeCommit
  for (int a = 0; a < 3; ++a) std::swap(sc->acc[a], sc->cand[a]);
  if (sc->tracer[0]) std::swap(sc->tracer[0], sc->tracer[1]);
  sc->acc_checked = true;
  sc->acc_rate = sv[1];
  sc->rate_known = !sc->external;
  sc->time = next_time;
  sc->generation += 1;
  rep->time = sc->time;
  rep->generation = sc->generation;
  return RH_OK;
}

// Simulation::step's dt (src/simulation.rs:926-979) for the fixed box with
// no source, forces or emitter: the acceleration rate is 0, so
// dt = min(requested, 0.5 max_courant / rate) when rate > 0.
extern "C" rh_status rh_scene_step_requested(rh_scene* sc, double requested_dt, rh_step_report* report,
                                             rh_cancel_fn cancel, void* cancel_ctx) {
  if (!sc) return RH_INVALID_ARGUMENT;
  if (report) std::memset(report, 0, sizeof(*report));
  const double c = sc->config.max_courant;
  if (!std::isfinite(c) || c <= 0.0 || c > 1.0) return RH_INVALID_SETTINGS;
  if (!std::isfinite(requested_dt) || requested_dt <= 0.0) return RH_INVALID_TIME_STEP;
  double rate = sc->acc_rate;
  if (!sc->rate_known || sc->external) {
    rh::Scratch& scr = sc->solver->impl->scratch();
    const DevGrid g = sc->solver->impl->dev_grid(sc->config.density);
    rh_status s = dispatch_precision(sc->config.precision, [&](auto* tag) -> rh_status {
      using T = std::remove_pointer_t<decltype(tag)>;
      rh::launch_velocity_rate<T>(g, sc->acc[0], sc->acc[1], sc->acc[2], scr.reduce(16), scr.stream);
      return RH_OK;
    });
    if (s) return s;
    RH_HIP(hipGetLastError());
    if ((s = scr.read_scalars(&rate, 1, 16))) return s;
  }
  if (!std::isfinite(rate)) return RH_ARITHMETIC_FAILURE;
  double dt = requested_dt;
  if (rate > 0.0) dt = std::fmin(dt, 0.5 * c / rate);
  const double next_time = sc->time + dt;
  if (dt <= 0.0 || !std::isfinite(next_time) || next_time <= sc->time) return RH_TIME_RESOLUTION;
  return rh_scene_step(sc, dt, report, cancel, cancel_ctx);
}

extern "C" rh_status rh_scene_set_tracer(rh_scene* sc, const float* tracer, size_t len) {
  if (!sc) return RH_INVALID_ARGUMENT;
  if (!tracer) {
    if (sc->tracer[0]) (void)hipFree(sc->tracer[0]);
    sc->tracer[0] = sc->tracer[1] = nullptr;
    return RH_OK;
  }
  if (len != sc->n) return RH_INVALID_ARGUMENT;
  if (sc->batched) {
    rh::set_last_error("scenes stepped by a batch carry no tracer", "rh_scene_set_tracer");
    return RH_UNSUPPORTED;
  }
  for (size_t i = 0; i < len; ++i)
    if (!std::isfinite(tracer[i])) return RH_NONFINITE_INPUT;
  if (!sc->tracer[0]) {
    const size_t b = (sc->n * sizeof(float) + 255) / 256 * 256;
    void* p = nullptr;
    RH_HIP(hipMalloc(&p, 2 * b));
    sc->tracer[0] = (float*)p;
    sc->tracer[1] = (float*)((char*)p + b);
  }
  RH_HIP(hipMemcpy(sc->tracer[0], tracer, len * sizeof(float), hipMemcpyHostToDevice));
  return RH_OK;
}

extern "C" rh_status rh_scene_get_tracer(const rh_scene* sc, float* out, size_t len) {
  if (!sc || !out || len != sc->n) return RH_INVALID_ARGUMENT;
  if (!sc->tracer[0]) return RH_INVALID_ARGUMENT;
  RH_HIP(hipMemcpy(out, sc->tracer[0], len * sizeof(float), hipMemcpyDeviceToHost));
  return RH_OK;
}
