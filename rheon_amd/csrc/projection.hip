// Qualification of supplied sealed-box projections on the device
// (src/solver_provider/composition/projection.rs).
//
// A caller (a learned or external producer, or this library's own solvers)
// supplies a pressure and the velocity it claims to be the provisional
// velocity corrected by that pressure.  The input is bound by a SHA-256
// input id; the supplied velocity must equal the correction of the
// provisional faces by the supplied pressure bit for bit (the momentum
// witness, computed with the device's reference-order correction kernel);
// the pressure goes through the provider workspace's requalification
// (finite, zero gauge, true residual and predicted divergence) as a
// supplied-array provider; the corrected faces must pass the actual
// divergence gate.  Any failure falls back to the native device solve from
// the same immutable input.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "common.hpp"
#include "handles.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rh {
namespace {

// ------------------------------------------------------------------ SHA-256
// FIPS 180-4 SHA-256 (the digest the reference's input id uses, through the
// sha2 crate, projection.rs:87-126).
class Sha256 {
 public:
  Sha256() {
    static const uint32_t h0[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    std::memcpy(h_, h0, sizeof(h_));
  }
  void update(const void* data, size_t n) {
    const unsigned char* p = (const unsigned char*)data;
    total_ += n;
    while (n) {
      const size_t take = std::min(n, sizeof(buf_) - fill_);
      std::memcpy(buf_ + fill_, p, take);
      fill_ += take;
      p += take;
      n -= take;
      if (fill_ == sizeof(buf_)) {
        block(buf_);
        fill_ = 0;
      }
    }
  }
  void u64(uint64_t v) {
    unsigned char b[8];
    for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(v >> (8 * i));
    update(b, 8);
  }
  void u32(uint32_t v) {
    unsigned char b[4];
    for (int i = 0; i < 4; ++i) b[i] = (unsigned char)(v >> (8 * i));
    update(b, 4);
  }
  void f64(double v) {
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    u64(bits);
  }
  // hash_text (projection.rs:143-146): byte length as u64, then the bytes
  void text(const char* s) {
    const size_t n = std::strlen(s);
    u64(n);
    update(s, n);
  }
  void finish(uint8_t out[32]) {
    const uint64_t bits = total_ * 8;
    const unsigned char one = 0x80;
    update(&one, 1);
    const unsigned char zero = 0;
    while (fill_ != 56) update(&zero, 1);
    unsigned char len[8];
    for (int i = 0; i < 8; ++i) len[i] = (unsigned char)(bits >> (56 - 8 * i));
    update(len, 8);
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(h_[i] >> (24 - 8 * j));
  }

 private:
  static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
  void block(const unsigned char* m) {
    static const uint32_t k[64] = {
        0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
        0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
        0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
        0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
        0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
        0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
        0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
        0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
      w[i] = ((uint32_t)m[4 * i] << 24) | ((uint32_t)m[4 * i + 1] << 16) | ((uint32_t)m[4 * i + 2] << 8) | m[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
      const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
      const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
      w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
    for (int i = 0; i < 64; ++i) {
      const uint32_t t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + k[i] + w[i];
      const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
      h = g;
      g = f;
      f = e;
      e = d + t1;
      d = c;
      c = b;
      b = a;
      a = t1 + t2;
    }
    h_[0] += a;
    h_[1] += b;
    h_[2] += c;
    h_[3] += d;
    h_[4] += e;
    h_[5] += f;
    h_[6] += g;
    h_[7] += h;
  }
  uint32_t h_[8];
  unsigned char buf_[64];
  size_t fill_ = 0;
  uint64_t total_ = 0;
};

constexpr const char* kContract = "sealed-mac-projection-si-f32-faces-f64-cell-zero-gauge-v1";
constexpr const char* kOperator = "sealed-constant-density-mac-cell-zero-gauge-v1";

// valid_text (projection.rs:140-142): non-empty, at most `limit` bytes, no
// Unicode control character (U+0000-U+001F, U+007F-U+009F; the latter are
// the UTF-8 pairs C2 80 .. C2 9F).
bool valid_text(const char* s, size_t limit) {
  if (!s) return false;
  const size_t n = std::strlen(s);
  if (n == 0 || n > limit) return false;
  for (size_t i = 0; i < n; ++i) {
    const unsigned char c = (unsigned char)s[i];
    if (c < 0x20 || c == 0x7f) return false;
    if (c == 0xc2 && i + 1 < n && (unsigned char)s[i + 1] >= 0x80 && (unsigned char)s[i + 1] <= 0x9f) return false;
  }
  return true;
}

size_t face_len_of(const rh_grid& g, int a) {
  size_t f = 1;
  for (int d = 0; d < 3; ++d) f *= g.counts[d] + (d == a ? 1 : 0);
  return f;
}

void copy_text(char* dst, size_t n, const char* src) {
  std::strncpy(dst, src, n - 1);
  dst[n - 1] = '\0';
}

void representation_of(const rh_grid& g, rh_projection_representation* r) {
  std::memset(r, 0, sizeof(*r));
  r->version = 1;
  copy_text(r->operator_id, sizeof(r->operator_id), kOperator);
  for (int d = 0; d < 3; ++d) r->pressure.shape_xyz[d] = g.counts[d];
  copy_text(r->pressure.layout, sizeof(r->pressure.layout), "x-fastest-cell-centered");
  copy_text(r->pressure.scalar, sizeof(r->pressure.scalar), "f64");
  copy_text(r->pressure.units, sizeof(r->pressure.units), "Pa");
  static const char* layouts[3] = {"x-fastest-x-face", "x-fastest-y-face", "x-fastest-z-face"};
  for (int a = 0; a < 3; ++a) {
    for (int d = 0; d < 3; ++d) r->velocity[a].shape_xyz[d] = g.counts[d] + (d == a ? 1 : 0);
    copy_text(r->velocity[a].layout, sizeof(r->velocity[a].layout), layouts[a]);
    copy_text(r->velocity[a].scalar, sizeof(r->velocity[a].scalar), "f32");
    copy_text(r->velocity[a].units, sizeof(r->velocity[a].units), "m/s");
  }
}

bool same_metadata(const rh_tensor_metadata& a, const rh_tensor_metadata& b) {
  for (int d = 0; d < 3; ++d)
    if (a.shape_xyz[d] != b.shape_xyz[d]) return false;
  return std::strncmp(a.layout, b.layout, sizeof(a.layout)) == 0 &&
         std::strncmp(a.scalar, b.scalar, sizeof(a.scalar)) == 0 &&
         std::strncmp(a.units, b.units, sizeof(a.units)) == 0;
}
bool same_representation(const rh_projection_representation& a, const rh_projection_representation& b) {
  if (a.version != b.version || std::strncmp(a.operator_id, b.operator_id, sizeof(a.operator_id)) != 0) return false;
  if (!same_metadata(a.pressure, b.pressure)) return false;
  for (int d = 0; d < 3; ++d)
    if (!same_metadata(a.velocity[d], b.velocity[d])) return false;
  return true;
}

// PressureOperator::validate_velocity (src/operator.rs:850-875): per axis
// the length, then finiteness, then every wall face exactly 0 in x-fastest
// order (the first nonzero one is reported).
rh_status validate_velocity_host(const rh_grid& g, const float* const v[3], const size_t len[3], int* wall_axis,
                                 int64_t* wall_index) {
  for (int a = 0; a < 3; ++a) {
    if (!v[a] || len[a] != face_len_of(g, a)) return RH_INVALID_ARGUMENT;
    for (size_t i = 0; i < len[a]; ++i)
      if (!std::isfinite(v[a][i])) return RH_NONFINITE_INPUT;
    const uint64_t fx = g.counts[0] + (a == 0), fy = g.counts[1] + (a == 1), fz = g.counts[2] + (a == 2);
    for (uint64_t k = 0; k < fz; ++k)
      for (uint64_t j = 0; j < fy; ++j)
        for (uint64_t i = 0; i < fx; ++i) {
          const uint64_t p[3] = {i, j, k};
          if (p[a] == 0 || p[a] == g.counts[a]) {
            const uint64_t idx = i + fx * (j + fy * k);
            if (v[a][idx] != 0.0f) {
              if (wall_axis) *wall_axis = a;
              if (wall_index) *wall_index = (int64_t)idx;
              return RH_NONZERO_WALL_VELOCITY;
            }
          }
        }
  }
  return RH_OK;
}

// ProjectionRequest::new (projection.rs:58-131): checks, then the input id.
rh_status request_id(const rh_projection_input* in, uint8_t out[32]) {
  if (!in || !in->grid) return RH_INVALID_ARGUMENT;
  if (in->domain != RH_DOMAIN_SEALED_BOX) return RH_REFUSED;
  if (!valid_text(in->source_revision, 4096)) return RH_REFUSED;
  const double end = in->interval_start + in->dt;
  const double limits[4] = {in->settings.relative_residual, in->settings.absolute_residual,
                            in->settings.divergence_limit, in->actual_divergence_limit};
  bool bad = !std::isfinite(in->interval_start) || in->interval_start < 0.0 || !std::isfinite(in->dt) ||
             in->dt <= 0.0 || !std::isfinite(end) || end <= in->interval_start || in->settings.max_iterations == 0;
  for (double v : limits) bad = bad || !std::isfinite(v) || v < 0.0;
  if (bad) return RH_INVALID_SETTINGS;
  const rh_grid& g = *in->grid;
  rh_status s = validate_grid(&g, nullptr);
  if (s) return s;
  if ((s = check_indexable(&g))) return s;
  if ((s = check_density(g, in->density))) return s;
  if ((s = validate_velocity_host(g, in->velocity, in->velocity_len, nullptr, nullptr))) return s;
  Sha256 h;
  h.text(kContract);
  rh_projection_representation r;
  representation_of(g, &r);
  h.u32(r.version);
  h.text(r.operator_id);
  const rh_tensor_metadata* meta[4] = {&r.pressure, &r.velocity[0], &r.velocity[1], &r.velocity[2]};
  for (const rh_tensor_metadata* m : meta) {
    for (int d = 0; d < 3; ++d) h.u64(m->shape_xyz[d]);
    h.text(m->layout);
    h.text(m->scalar);
    h.text(m->units);
  }
  for (int d = 0; d < 3; ++d) h.u64(g.counts[d]);
  for (int d = 0; d < 3; ++d) h.f64(g.spacing[d]);
  for (int d = 0; d < 3; ++d) h.f64(g.origin[d]);
  h.u64(in->run_id);
  h.text(in->source_revision);
  h.u64(in->generation);
  const double values[7] = {in->density, in->interval_start, in->dt, in->settings.relative_residual,
                            in->settings.absolute_residual, in->settings.divergence_limit,
                            in->actual_divergence_limit};
  for (double v : values) h.f64(v);
  h.u64((uint64_t)in->settings.max_iterations);
  for (int a = 0; a < 3; ++a) {
    h.u64(in->velocity_len[a]);
    h.update(in->velocity[a], in->velocity_len[a] * sizeof(float));  // f32 bits, little-endian
  }
  h.finish(out);
  return RH_OK;
}

// First index per axis at which two face arrays differ bitwise.
__global__ void k_first_mismatch(const unsigned* __restrict__ a, const unsigned* __restrict__ b, unsigned n,
                                 unsigned* __restrict__ first) {
  unsigned mine = 0xffffffffu;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (a[i] != b[i]) {
      mine = i;
      break;
    }
  if (mine != 0xffffffffu) atomicMin(first, mine);
}
__global__ void k_fill_words(unsigned* p, int n, unsigned v) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

}  // namespace
}  // namespace rh

struct rh_projection_workspace {
  rh_grid grid{};
  rh_method fallback = RH_JACOBI_PCG;
  rh_provider_workspace* pressure = nullptr;  // ProviderWorkspace of the qualification
  rh::Scratch scratch;
  void* arena = nullptr;
  float* in[3] = {nullptr, nullptr, nullptr};    // provisional faces
  float* out[3] = {nullptr, nullptr, nullptr};   // rebuilt (corrected) faces
  float* cand[3] = {nullptr, nullptr, nullptr};  // supplied faces
  double* cand_p = nullptr;                       // supplied pressure
  size_t n = 0, flen[3] = {0, 0, 0};
  std::vector<double> rhs;  // host RHS for the provider request
  bool have_result = false;
  ~rh_projection_workspace() {
    if (pressure) rh_provider_workspace_destroy(pressure);
    scratch.release();
    if (arena) (void)hipFree(arena);
  }
};

namespace {

// Corrects the provisional faces by the device pressure `p` into ws->out and
// returns the correction's status (operator.rs:627-668).
rh_status rebuild(rh_projection_workspace* ws, const rh::DevGrid& g, const double* p, double dt) {
  rh::Scratch& sc = ws->scratch;
  for (int a = 0; a < 3; ++a)
    RH_HIP(hipMemcpyAsync(ws->out[a], ws->in[a], ws->flen[a] * 4, hipMemcpyDeviceToDevice, sc.stream));
  rh_status s = sc.clear_flags();
  if (s) return s;
  rh::launch_correct<float>(g, p, dt, ws->out[0], ws->out[1], ws->out[2], sc.flags, sc.stream);
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  return rh::flags_to_status(f);
}

rh_status divergence_of_out(rh_projection_workspace* ws, const rh::DevGrid& g, double* out) {
  rh::Scratch& sc = ws->scratch;
  rh::launch_divergence_max<float>(g, ws->out[0], ws->out[1], ws->out[2], sc.reduce(0), sc.stream);
  RH_HIP(hipGetLastError());
  return sc.read_scalars(out, 1, 0);
}

}  // namespace

extern "C" rh_status rh_projection_input_id(const rh_projection_input* input, uint8_t out[32]) {
  if (!out) return RH_INVALID_ARGUMENT;
  return rh::request_id(input, out);
}

extern "C" rh_status rh_projection_representation_of(const rh_grid* grid, rh_projection_representation* out) {
  if (!grid || !out) return RH_INVALID_ARGUMENT;
  rh_status s = rh::validate_grid(grid, nullptr);
  if (s) return s;
  rh::representation_of(*grid, out);
  return RH_OK;
}

extern "C" rh_status rh_projection_workspace_create(const rh_grid* grid, uint64_t limit, rh_method fallback,
                                                    rh_projection_workspace** out) {
  if (!grid || !out) return RH_INVALID_ARGUMENT;
  *out = nullptr;
  rh_status s = rh::validate_grid(grid, nullptr);
  if (s) return s;
  if ((s = rh::check_indexable(grid))) return s;
  const size_t n = (size_t)grid->counts[0] * grid->counts[1] * grid->counts[2];
  size_t faces = 0;
  for (int a = 0; a < 3; ++a) faces += rh::face_len_of(*grid, a);
  // ProjectionWorkspace::new (projection.rs:283-335): RHS and divergence
  // cells plus three f32 face arrays, then the provider workspace within
  // what remains of the limit
  const uint64_t extra = 16ull * n + 4ull * faces;
  if (limit && extra > limit) return RH_BUFFER_LIMIT;
  auto ws = std::unique_ptr<rh_projection_workspace>(new (std::nothrow) rh_projection_workspace());
  if (!ws) return RH_DEVICE_ERROR;
  ws->grid = *grid;
  ws->fallback = fallback;
  const uint64_t remaining = limit ? std::max<uint64_t>(limit - extra, 1) : 0;  // 0 = no limit
  if ((s = rh_provider_workspace_create(grid, remaining, fallback, &ws->pressure))) return s;
  if ((s = ws->scratch.init())) return s;
  ws->n = n;
  size_t total = 0;
  auto round = [](size_t b) { return (b + 255) / 256 * 256; };
  for (int a = 0; a < 3; ++a) {
    ws->flen[a] = rh::face_len_of(*grid, a);
    total += 3 * round(ws->flen[a] * 4);
  }
  total += round(n * 8);
  RH_HIP(hipMalloc(&ws->arena, total));
  char* p = (char*)ws->arena;
  for (int a = 0; a < 3; ++a) {
    ws->in[a] = (float*)p;
    p += round(ws->flen[a] * 4);
    ws->out[a] = (float*)p;
    p += round(ws->flen[a] * 4);
    ws->cand[a] = (float*)p;
    p += round(ws->flen[a] * 4);
  }
  ws->cand_p = (double*)p;
  try {
    ws->rhs.assign(n, 0.0);
  } catch (...) {
    return RH_BUFFER_LIMIT;
  }
  *out = ws.release();
  return RH_OK;
}

extern "C" void rh_projection_workspace_destroy(rh_projection_workspace* ws) { delete ws; }

// ProjectionWorkspace::qualify (projection.rs:350-479).
extern "C" rh_status rh_projection_qualify(rh_projection_workspace* ws, const rh_projection_input* in,
                                           const rh_projection_candidate* cand, rh_cancel_fn cancel,
                                           void* cancel_ctx, rh_projection_report* rep) {
  using namespace rh;
  if (!ws || !in || !rep) return RH_INVALID_ARGUMENT;
  std::memset(rep, 0, sizeof(*rep));
  ws->have_result = false;
  auto cancelled = [&]() { return cancel && cancel(cancel_ctx); };
  // the request (ProjectionRequest::new)
  rh_status s = request_id(in, rep->input_id);
  if (s == RH_REFUSED) {
    rep->refusal = in->domain != RH_DOMAIN_SEALED_BOX ? RH_REFUSE_DOMAIN : RH_REFUSE_IDENTITY;
    return s;
  }
  if (s) return s;
  const rh_grid& g0 = *in->grid;
  for (int d = 0; d < 3; ++d)
    if (g0.counts[d] != ws->grid.counts[d] || g0.spacing[d] != ws->grid.spacing[d] ||
        g0.origin[d] != ws->grid.origin[d]) {
      rep->refusal = RH_REFUSE_WORKSPACE_GRID;
      return RH_REFUSED;
    }
  if (cand) {
    if (std::memcmp(cand->input_id, rep->input_id, 32) != 0) {
      rep->refusal = RH_REFUSE_INPUT_BINDING;
      return RH_REFUSED;
    }
    rh_projection_representation want;
    representation_of(ws->grid, &want);
    if (!same_representation(cand->representation, want)) {
      rep->refusal = RH_REFUSE_REPRESENTATION;
      return RH_REFUSED;
    }
    if (!valid_text(cand->algorithm, 256) || !valid_text(cand->version, 256) ||
        !valid_text(cand->source_revision, 4096)) {
      rep->refusal = RH_REFUSE_IDENTITY;
      return RH_REFUSED;
    }
    rep->has_supplied_identity = 1;
  }
  if (cancelled()) return RH_CANCELLED;  // BeforeQualification
  Scratch& sc = ws->scratch;
  const DevGrid g = make_dev_grid(ws->grid, in->density);
  for (int a = 0; a < 3; ++a)
    RH_HIP(hipMemcpyAsync(ws->in[a], in->velocity[a], ws->flen[a] * 4, hipMemcpyHostToDevice, sc.stream));
  // RHS of the provisional faces (operator.rs:460-509), on the device
  double* d_rhs = ws->pressure->d_rhs;
  if ((s = sc.clear_flags())) return s;
  launch_rhs<float>(g, ws->in[0], ws->in[1], ws->in[2], in->dt, d_rhs, sc.flags, sc.stream, true);
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  if ((s = sc.read_flags(&f))) return s;
  if (f) return flags_to_status(f);
  RH_HIP(hipMemcpy(ws->rhs.data(), d_rhs, ws->n * 8, hipMemcpyDeviceToHost));
  // check_candidate (projection.rs:507-534)
  if (cand) {
    auto field = [&](rh_status why) {
      rep->rejection = RH_REJECT_FIELD;
      rep->rejection_status = why;
    };
    bool ok = true;
    if (!cand->pressure || cand->pressure_len != ws->n) {
      field(RH_INVALID_ARGUMENT);  // OperatorError::LengthMismatch
      ok = false;
    } else {
      for (size_t i = 0; i < ws->n && ok; ++i)
        if (!std::isfinite(cand->pressure[i])) {
          field(RH_NONFINITE_INPUT);
          ok = false;
        }
    }
    if (ok) {
      int wa = 0;
      int64_t wi = 0;
      const rh_status v = validate_velocity_host(ws->grid, cand->velocity, cand->velocity_len, &wa, &wi);
      if (v) {
        field(v);
        if (v == RH_NONZERO_WALL_VELOCITY) {
          rep->wall_axis = wa;
          rep->wall_index = wi;
        }
        ok = false;
      }
    }
    if (ok) {
      RH_HIP(hipMemcpyAsync(ws->cand_p, cand->pressure, ws->n * 8, hipMemcpyHostToDevice, sc.stream));
      const rh_status r = rebuild(ws, g, ws->cand_p, in->dt);
      if (r == RH_DEVICE_ERROR) return r;
      if (r) {
        field(r);
        ok = false;
      }
    }
    if (ok) {
      // the momentum witness: rebuilt faces == supplied faces, bit for bit
      for (int a = 0; a < 3; ++a)
        RH_HIP(hipMemcpyAsync(ws->cand[a], cand->velocity[a], ws->flen[a] * 4, hipMemcpyHostToDevice, sc.stream));
      unsigned* first = sc.counters;
      k_fill_words<<<1, 64, 0, sc.stream>>>(first, 3, 0xffffffffu);
      for (int a = 0; a < 3; ++a) {
        const unsigned n = (unsigned)ws->flen[a];
        const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
        k_first_mismatch<<<blocks, 256, 0, sc.stream>>>((const unsigned*)ws->out[a], (const unsigned*)ws->cand[a],
                                                         n, first + a);
      }
      RH_HIP(hipGetLastError());
      unsigned idx[3];
      RH_HIP(hipMemcpyAsync(sc.host, first, sizeof(idx), hipMemcpyDeviceToHost, sc.stream));
      RH_HIP(hipStreamSynchronize(sc.stream));
      std::memcpy(idx, sc.host, sizeof(idx));
      for (int a = 0; a < 3 && ok; ++a)
        if (idx[a] != 0xffffffffu) {
          rep->rejection = RH_REJECT_MOMENTUM_WITNESS;
          rep->witness_axis = a;
          rep->witness_index = idx[a];
          ok = false;
        }
    }
  }
  if (cancelled()) return RH_CANCELLED;  // BeforePressure
  rh_pressure_request preq{&ws->grid, in->density, in->dt, ws->rhs.data(), ws->n};
  rh_provider supplied;
  const bool offer = cand && rep->rejection == RH_REJECT_NONE;
  if (offer) {
    supplied.supplied = cand->pressure;
    supplied.supplied_len = ws->n;
    for (int d = 0; d < 3; ++d) supplied.config.shape_xyz[d] = ws->grid.counts[d];
  }
  if ((s = rh_provider_workspace_solve(ws->pressure, &preq, &in->settings, offer ? &supplied : nullptr, cancel,
                                       cancel_ctx, &rep->pressure)))
    return s;
  if (cancelled()) return RH_CANCELLED;  // AfterPressure
  bool supplied_accepted = std::strncmp(rep->pressure.selected, kSuppliedPressureId, sizeof(rep->pressure.selected)) == 0;
  if (rep->pressure.rejection != RH_OK) {
    rep->rejection = RH_REJECT_PRESSURE;
    rep->rejection_status = rep->pressure.rejection;
  }
  if ((s = rebuild(ws, g, ws->pressure->d_p, in->dt))) return s;
  double divergence = 0.0;
  if ((s = divergence_of_out(ws, g, &divergence))) return s;
  if (supplied_accepted && divergence > in->actual_divergence_limit) {
    rep->rejection = RH_REJECT_ACTUAL_DIVERGENCE;
    rep->rejected_divergence = divergence;
    rep->discarded_candidate_pressure = rep->pressure;
    rep->has_discarded = 1;
    if ((s = rh_provider_workspace_solve(ws->pressure, &preq, &in->settings, nullptr, cancel, cancel_ctx,
                                         &rep->pressure)))
      return s;
    if (cancelled()) return RH_CANCELLED;  // AfterPressure
    supplied_accepted = false;
    if ((s = rebuild(ws, g, ws->pressure->d_p, in->dt))) return s;
    if ((s = divergence_of_out(ws, g, &divergence))) return s;
  }
  rep->actual_divergence_max = divergence;
  if (divergence > in->actual_divergence_limit) return RH_DIVERGENCE_LIMIT;
  if (cancelled()) return RH_CANCELLED;  // BeforeReturn
  rep->selection = supplied_accepted ? RH_SELECT_SUPPLIED : (cand ? RH_SELECT_NATIVE_FALLBACK : RH_SELECT_NATIVE_BASELINE);
  ws->have_result = true;
  return RH_OK;
}

extern "C" rh_status rh_projection_result(const rh_projection_workspace* ws, double* pressure, size_t pressure_len,
                                          float* u, float* v, float* w) {
  if (!ws || !ws->have_result) return RH_INVALID_ARGUMENT;
  if (pressure) {
    if (pressure_len != ws->n) return RH_INVALID_ARGUMENT;
    std::memcpy(pressure, ws->pressure->pressure.data(), ws->n * 8);
  }
  float* dst[3] = {u, v, w};
  for (int a = 0; a < 3; ++a)
    if (dst[a]) RH_HIP(hipMemcpy(dst[a], ws->out[a], ws->flen[a] * 4, hipMemcpyDeviceToHost));
  return RH_OK;
}
