// Dense3d frames on the device: write_dense3d_frame (src/dense3d_sequence.rs:
// 412-564) for a device-resident accepted transport view.
//
// The frame is one JSON line: a header (frame, time, dt, stamps), six field
// arrays (velocity_x/y/z and tracer as f32 with 10 significant digits,
// fraction and pressure as f64 with 18) and the diagnostics object or null.
// The host checks the view's clock, stamps and report (the reference's
// preflight, in its order) and formats the header and diagnostics (a few
// numbers) with the same formatter; the field arrays, which are the frame's
// bytes, are formatted on the device:
//
//   k_d3_count  per chunk of kChunk values, the bytes of its numbers and
//               their separating commas; on the same reads, the frame's
//               field checks (non-finite velocity, tracer or pressure,
//               fraction outside [0, 1]);
//   k_d3_scan   chunk offsets of every field (one workgroup per field);
//   k_d3_write  per chunk: the numbers formatted into LDS at their offsets,
//               then copied out with 16-byte stores (head and tail bytes
//               singly), so the line is written once, in order, unaligned.
//
// Every number is dense3d.hpp's exact formatter: the bytes equal the
// reference's `{v:.9e}` / `{v:.17e}` output.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "dense3d.hpp"
#include "profiler.hpp"
#include "runtime.hpp"

namespace rh {

namespace {

constexpr int kD3Threads = 256;
constexpr int kD3PerThread = 4;
constexpr int kChunk = kD3Threads * kD3PerThread;  // values per workgroup
constexpr unsigned kBadVelocity = 1u, kBadCell = 2u;

template <typename T>
__device__ __forceinline__ double widen(T v) {
  return (double)v;
}

// The frame's field checks (src/dense3d_sequence.rs:431-450), made on the
// values the count pass reads anyway: velocity and tracer finite, pressure
// finite, fraction finite and in [0, 1].  kFraction selects the last.
template <typename T, int N, bool kFraction>
__global__ void __launch_bounds__(kD3Threads)
k_d3_count(const T* __restrict__ vals, long long n, unsigned* __restrict__ chunk_bytes, unsigned bad_flag,
           unsigned* flags) {
  __shared__ unsigned red[kD3Threads / 64];
  const long long base = (long long)blockIdx.x * kChunk;
  unsigned mine = 0;
  bool bad = false;
#pragma unroll 1
  for (int q = 0; q < kD3PerThread; ++q) {
    const long long i = base + q * kD3Threads + threadIdx.x;
    if (i >= n) continue;
    const double v = widen(vals[i]);
    const bool ok = kFraction ? (v >= 0.0 && v <= 1.0) : (bool)isfinite(v);
    if (!ok) {
      bad = true;
      continue;
    }
    mine += (unsigned)(d3::sci_length(d3::decompose<N>(v), N) + (i > 0 ? 1 : 0));
  }
  if (bad) atomicOr(flags, bad_flag);
  for (int m = 32; m >= 1; m >>= 1) mine += __shfl_xor(mine, m, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int k = 0; k < kD3Threads / 64; ++k) t += red[k];
    chunk_bytes[blockIdx.x] = t;
  }
}

// Exclusive scans of the fields' chunk byte counts into 64-bit offsets,
// one workgroup per field: counts in[c0 .. c0 + nb) -> out[o0 .. o0 + nb),
// out[o0 + nb] = totals[q] = the field's total.
struct ScanJobs {
  long long c0[6], o0[6];
  int nb[6], q[6];
};
__global__ void __launch_bounds__(1024) k_d3_scan(const unsigned* __restrict__ counts, ScanJobs jobs,
                                                  unsigned long long* __restrict__ offs,
                                                  unsigned long long* __restrict__ totals) {
  __shared__ unsigned long long wsum[16];
  const int f = blockIdx.x;
  const unsigned* in = counts + jobs.c0[f];
  unsigned long long* out = offs + jobs.o0[f];
  const int nb = jobs.nb[f];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int per = (nb + 1023) / 1024;
  const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  unsigned long long s = 0;
  for (int b = b0; b < b1; ++b) s += in[b];
  // inclusive scan over the workgroup: within each wave, then the wave sums
  unsigned long long incl = s;
  for (int m = 1; m < 64; m <<= 1) {
    const unsigned long long y = __shfl_up(incl, m, 64);
    if (lane >= m) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  unsigned long long run = incl - s;
  for (int k = 0; k < wave; ++k) run += wsum[k];
  if (t == 1023) out[nb] = totals[jobs.q[f]] = run + s;
  for (int b = b0; b < b1; ++b) {
    out[b] = run;
    run += in[b];
  }
}

// Formats chunk blockIdx.x into out + base + chunk offset.  Value q of
// thread t is chunk value q * kD3Threads + t, so a wave's LDS writes fall on
// consecutive numbers (spread over the banks); each value's offset comes
// from an exclusive scan of the chunk's lengths in value order.
template <typename T, int N>
__global__ void __launch_bounds__(kD3Threads)
k_d3_write(const T* __restrict__ vals, long long n, const unsigned long long* __restrict__ chunk_off,
           unsigned long long base, char* __restrict__ out) {
  constexpr int kMaxChars = (N > 10 ? d3::kMaxF64Chars : d3::kMaxF32Chars) + 1;
  __shared__ __align__(16) char buf[kChunk * kMaxChars + 32];
  __shared__ unsigned long long sD[kChunk];
  __shared__ short sE[kChunk];
  __shared__ unsigned voff[kChunk];  // lengths, then offsets within the chunk (bit 31: negative)
  __shared__ unsigned wsum[kD3Threads / 64];
  const long long first = (long long)blockIdx.x * kChunk;
  const unsigned long long start = base + chunk_off[blockIdx.x];
  const int lead = (int)(start & 15u);  // buf offset with start's 16-byte phase
#pragma unroll 1
  for (int q = 0; q < kD3PerThread; ++q) {
    const int v = q * kD3Threads + threadIdx.x;
    const long long i = first + v;
    unsigned len = 0;
    if (i < n) {
      const d3::Sci sc = d3::decompose<N>(widen(vals[i]));
      sD[v] = sc.D;
      sE[v] = (short)sc.E;
      len = (unsigned)(d3::sci_length(sc, N) + (i > 0 ? 1 : 0)) | (sc.neg ? 0x80000000u : 0u);
    }
    voff[v] = len;
  }
  __syncthreads();
  // exclusive scan of the lengths: thread t owns entries t*kD3PerThread ..
  // t*kD3PerThread + kD3PerThread - 1
  unsigned mine = 0;
  unsigned loc[kD3PerThread], neg[kD3PerThread];
#pragma unroll
  for (int q = 0; q < kD3PerThread; ++q) {
    const unsigned w = voff[threadIdx.x * kD3PerThread + q];
    loc[q] = mine;
    neg[q] = w & 0x80000000u;
    mine += w & 0x7fffffffu;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned incl = mine;
  for (int m = 1; m < 64; m <<= 1) {
    const unsigned y = __shfl_up(incl, m, 64);
    if (lane >= m) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  unsigned off = incl - mine;
  for (int k = 0; k < wave; ++k) off += wsum[k];
  unsigned total = 0;
  for (int k = 0; k < kD3Threads / 64; ++k) total += wsum[k];
#pragma unroll
  for (int q = 0; q < kD3PerThread; ++q) voff[threadIdx.x * kD3PerThread + q] = (off + loc[q]) | neg[q];
  __syncthreads();
#pragma unroll 1
  for (int q = 0; q < kD3PerThread; ++q) {
    const int v = q * kD3Threads + threadIdx.x;
    const long long i = first + v;
    if (i >= n) break;
    const unsigned w = voff[v];
    char* p = buf + lead + (w & 0x7fffffffu);
    if (i > 0) *p++ = ',';
    const d3::Sci sc{sD[v], (int)sE[v], (w & 0x80000000u) != 0};
    d3::sci_write(sc, N, p);
  }
  __syncthreads();
  // copy [start, start + total) out: 16-byte stores on the aligned body
  const unsigned long long end = start + total;
  const unsigned long long a0 = (start + 15u) & ~15ull, a1 = end & ~15ull;
  if (a0 >= a1) {
    for (unsigned k = threadIdx.x; k < total; k += kD3Threads) out[start + k] = buf[lead + k];
    return;
  }
  for (unsigned k = threadIdx.x; k < (unsigned)(a0 - start); k += kD3Threads) out[start + k] = buf[lead + k];
  for (unsigned k = threadIdx.x; k < (unsigned)(end - a1); k += kD3Threads) out[a1 + k] = buf[lead + (a1 - start) + k];
  const unsigned nv = (unsigned)((a1 - a0) >> 4);
  const uint4* src = reinterpret_cast<const uint4*>(buf + lead + (a0 - start));
  uint4* dst = reinterpret_cast<uint4*>(out + a0);
  for (unsigned k = threadIdx.x; k < nv; k += kD3Threads) dst[k] = src[k];
}

std::string sci_host(double v, int digits) {
  char b[32];
  const int n = digits == 10 ? d3::format_sci<10>(v, b) : d3::format_sci<18>(v, b);
  return std::string(b, b + n);
}

rh_status invalid_state(const char* msg) {
  set_last_error(msg, "write_dense3d_frame");
  return RH_INVALID_STATE;
}

struct Field {
  const char* name;
  const void* data;
  long long n;
  bool f64;
};

}  // namespace

}  // namespace rh

using namespace rh;

// The view's clock and stamps and the report (src/dense3d_sequence.rs:
// 452-495), after the field checks as in the reference.
static rh_status header_checks(const rh_dense3d_view* view, uint64_t frame, const rh_dense3d_report* report) {
  if (!(std::isfinite(view->carrier_time) && view->carrier_time == view->liquid_time &&
        view->carrier_stamp.version == view->carrier_generation))
    return invalid_state("state clock/stamp");
  if (report) {
    const rh_dense3d_report& r = *report;
    if (!(r.carrier_time == view->carrier_time && r.liquid_time == r.carrier_time && r.carrier_dt == r.liquid_dt &&
          r.carrier_dt > 0.0 && r.carrier_generation == view->carrier_generation &&
          r.liquid_stamp.id == view->liquid_stamp.id && r.liquid_stamp.version == view->liquid_stamp.version &&
          r.liquid_flow.id == view->carrier_stamp.id && r.liquid_flow.version == view->carrier_stamp.version))
      return invalid_state("report continuity");
    const double vals[12] = {r.carrier_dt, r.pressure_true_residual_max, r.carrier_actual_divergence_max,
                             r.liquid_actual_divergence_max, r.liquid_volume_before, r.liquid_volume_after,
                             r.liquid_mass_after, r.inward_boundary_volume, r.outward_boundary_volume,
                             r.source_volume, r.volume_balance_error, r.volume_rounding_budget};
    for (double v : vals)
      if (!std::isfinite(v)) return invalid_state("nonfinite report");
  } else if (!(frame == 0 && view->carrier_time == 0.0 && view->carrier_generation == 0)) {
    return invalid_state("initial frame");
  }
  return RH_OK;
}

extern "C" int32_t rh_format_sci(double v, int32_t digits, char* out) {
  if (!out || (digits != 10 && digits != 18) || !std::isfinite(v)) return -1;
  const std::string s = sci_host(v, digits);
  std::memcpy(out, s.data(), s.size());
  out[s.size()] = 0;
  return (int32_t)s.size();
}

extern "C" rh_status rh_dense3d_frame(const rh_grid* grid, const rh_dense3d_view* view, uint64_t frame,
                                      const rh_dense3d_report* report, char* d_out, uint64_t capacity,
                                      uint64_t* bytes) {
  if (!grid || !view || !bytes) return RH_INVALID_ARGUMENT;
  *bytes = 0;
  size_t cells = 0;
  rh_status st = validate_grid(grid, &cells);
  if (st) return st;
  for (int a = 0; a < 3; ++a)
    if (!view->velocity[a]) return RH_INVALID_ARGUMENT;
  if (!view->tracer || !view->fraction || !view->pressure) return RH_INVALID_ARGUMENT;
  long long fl[3];
  for (int a = 0; a < 3; ++a) {
    fl[a] = 1;
    for (int d = 0; d < 3; ++d) fl[a] *= (long long)grid->counts[d] + (d == a ? 1 : 0);
  }
  const long long nc = (long long)cells;
  // preflight in the reference's order (src/dense3d_sequence.rs:422-495)
  if (!view->transport_only) return invalid_state("transport-only view required");
  std::unique_lock<std::mutex> lock;
  Scratch& scr = global_scratch(lock);
  if ((st = scr.init())) return st;
  hipStream_t s = scr.stream;
  RH_HIP(hipMemsetAsync(scr.flags, 0, sizeof(unsigned), s));
  // the text around the field arrays (src/dense3d_sequence.rs:497-563)
  const std::string head = "{\"frame\":" + std::to_string(frame) + ",\"time_s\":" + sci_host(view->carrier_time, 18) +
                           ",\"dt_s\":" + sci_host(report ? report->carrier_dt : 0.0, 18) +
                           ",\"carrier_stamp\":{\"id\":\"" + std::to_string(view->carrier_stamp.id) +
                           "\",\"version\":\"" + std::to_string(view->carrier_stamp.version) +
                           "\"},\"liquid_stamp\":{\"id\":\"" + std::to_string(view->liquid_stamp.id) +
                           "\",\"version\":\"" + std::to_string(view->liquid_stamp.version) + "\"},\"fields\":{";
  std::string diag;
  if (report) {
    const rh_dense3d_report& r = *report;
    diag = "{\"pressure_iterations\":" + std::to_string(r.pressure_iterations) +
           ",\"pressure_residual_m3_s2\":" + sci_host(r.pressure_true_residual_max, 18) +
           ",\"carrier_divergence_s_inv\":" + sci_host(r.carrier_actual_divergence_max, 18) +
           ",\"liquid_divergence_s_inv\":" + sci_host(r.liquid_actual_divergence_max, 18) +
           ",\"volume_before_m3\":" + sci_host(r.liquid_volume_before, 18) +
           ",\"volume_after_m3\":" + sci_host(r.liquid_volume_after, 18) +
           ",\"mass_after_kg\":" + sci_host(r.liquid_mass_after, 18) +
           ",\"inward_m3\":" + sci_host(r.inward_boundary_volume, 18) +
           ",\"outward_m3\":" + sci_host(r.outward_boundary_volume, 18) +
           ",\"source_m3\":" + sci_host(r.source_volume, 18) + ",\"balance_m3\":" + sci_host(r.volume_balance_error, 18) +
           ",\"rounding_budget_m3\":" + sci_host(r.volume_rounding_budget, 18) + "}";
  } else {
    diag = "null";
  }
  const std::string tail = "},\"diagnostics\":" + diag + "}\n";
  const Field fields[6] = {{"velocity_x", view->velocity[0], fl[0], false},
                           {"velocity_y", view->velocity[1], fl[1], false},
                           {"velocity_z", view->velocity[2], fl[2], false},
                           {"tracer", view->tracer, nc, false},
                           {"fraction", view->fraction, nc, true},
                           {"pressure", view->pressure, nc, true}};
  // chunk byte counts and offsets per field, in one device allocation
  long long nchunks[6], coff[6], total_chunks = 0;
  for (int q = 0; q < 6; ++q) {
    nchunks[q] = (fields[q].n + kChunk - 1) / kChunk;
    coff[q] = total_chunks;
    total_chunks += nchunks[q];
  }
  // counts, then (256-aligned) offsets with one total per field; the
  // allocation is kept for later frames (grown when a larger grid comes),
  // guarded by the global scratch's lock like the rest of this call
  static void* work = nullptr;
  static size_t work_bytes = 0;
  const size_t counts_bytes = sizeof(unsigned) * (size_t)total_chunks;
  const size_t offs_bytes = sizeof(unsigned long long) * (size_t)(total_chunks + 12);
  const size_t need = ((counts_bytes + 255) / 256) * 256 + offs_bytes;
  if (work_bytes < need) {
    if (work) RH_HIP(hipFree(work));
    work = nullptr;
    work_bytes = 0;
    RH_HIP(hipMalloc(&work, need));
    work_bytes = need;
  }
  unsigned* counts = (unsigned*)work;
  unsigned long long* offs = (unsigned long long*)((char*)work + ((counts_bytes + 255) / 256) * 256);
  unsigned long long* totals = offs + total_chunks + 6;  // [6]: each field's total (0 when empty)
  RH_HIP(hipMemsetAsync(totals, 0, 6 * sizeof(unsigned long long), s));
  for (int q = 0; q < 6; ++q) {
    if (!nchunks[q]) continue;
    prof::Scope ps(prof::kDense3dCount, s, -1);
    const unsigned flag = q < 3 ? kBadVelocity : kBadCell;
    if (q == 4)
      launch(k_d3_count<double, 18, true>, dim3((unsigned)nchunks[q]), dim3(kD3Threads), 0, s,
             (const double*)fields[q].data, fields[q].n, counts + coff[q], flag, scr.flags);
    else if (fields[q].f64)
      launch(k_d3_count<double, 18, false>, dim3((unsigned)nchunks[q]), dim3(kD3Threads), 0, s,
             (const double*)fields[q].data, fields[q].n, counts + coff[q], flag, scr.flags);
    else
      launch(k_d3_count<float, 10, false>, dim3((unsigned)nchunks[q]), dim3(kD3Threads), 0, s,
             (const float*)fields[q].data, fields[q].n, counts + coff[q], flag, scr.flags);
  }
  ScanJobs jobs{};
  int njobs = 0;
  for (int q = 0; q < 6; ++q)
    if (nchunks[q]) {
      jobs.c0[njobs] = coff[q];
      jobs.o0[njobs] = coff[q] + q;
      jobs.nb[njobs] = (int)nchunks[q];
      jobs.q[njobs] = q;
      ++njobs;
    }
  if (njobs) k_d3_scan<<<njobs, 1024, 0, s>>>(counts, jobs, offs, totals);
  RH_HIP(hipGetLastError());
  // the six field totals, gathered by the scan, in one copy to the pinned page
  unsigned long long* tot_h = (unsigned long long*)((char*)scr.host + 2048);
  RH_HIP(hipMemcpyAsync(tot_h, totals, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  unsigned f = 0;
  if ((st = scr.read_flags(&f))) return st;  // synchronises the stream
  unsigned long long tot[6];
  for (int q = 0; q < 6; ++q) tot[q] = tot_h[q];
  if (f & kBadVelocity) return invalid_state("nonfinite velocity");
  if (f & kBadCell) return invalid_state("nonfinite or invalid cell field");
  if ((st = header_checks(view, frame, report))) return st;
  // layout: head, then per field "name":[ values ] and the separator
  std::vector<std::string> pre(6), post(6);
  for (int q = 0; q < 6; ++q) {
    pre[q] = std::string("\"") + fields[q].name + "\":[";
    post[q] = q < 4 ? "]," : (q == 4 ? "]," : "]");
  }
  unsigned long long at = head.size();
  unsigned long long field_at[6];
  for (int q = 0; q < 6; ++q) {
    at += pre[q].size();
    field_at[q] = at;
    at += tot[q] + post[q].size();
  }
  at += tail.size();
  *bytes = at;
  if (!d_out || capacity < at) {
    set_last_error("output buffer smaller than the frame (the length needed is in *bytes)", "rh_dense3d_frame");
    return RH_BUFFER_LIMIT;
  }
  // text pieces (offset, text), each copied in one transfer: from the pinned
  // page when they fit there (asynchronous), else from host memory
  std::vector<std::pair<unsigned long long, std::string>> pieces;
  pieces.push_back({0, head});
  for (int q = 0; q < 6; ++q) {
    pieces.push_back({field_at[q] - pre[q].size(), pre[q]});
    pieces.push_back({field_at[q] + tot[q], post[q]});
  }
  pieces.push_back({at - tail.size(), tail});
  size_t text_len = 0;
  for (auto& pc : pieces) text_len += pc.second.size();
  char* stage = (char*)scr.host + 2304;  // the page's last 1792 bytes
  const bool pinned = text_len <= 4096 - 2304;
  for (auto& pc : pieces) {
    const char* src = pc.second.data();
    if (pinned) {
      std::memcpy(stage, src, pc.second.size());
      src = stage;
      stage += pc.second.size();
    }
    RH_HIP(hipMemcpyAsync(d_out + pc.first, src, pc.second.size(), hipMemcpyHostToDevice, s));
  }
  for (int q = 0; q < 6; ++q) {
    if (!nchunks[q]) continue;
    prof::Scope ps(prof::kDense3dWrite, s, -1);
    if (fields[q].f64)
      launch(k_d3_write<double, 18>, dim3((unsigned)nchunks[q]), dim3(kD3Threads), 0, s,
             (const double*)fields[q].data, fields[q].n, (const unsigned long long*)(offs + coff[q] + q),
             (unsigned long long)field_at[q], d_out);
    else
      launch(k_d3_write<float, 10>, dim3((unsigned)nchunks[q]), dim3(kD3Threads), 0, s,
             (const float*)fields[q].data, fields[q].n, (const unsigned long long*)(offs + coff[q] + q),
             (unsigned long long)field_at[q], d_out);
  }
  RH_HIP(hipGetLastError());
  RH_HIP(hipStreamSynchronize(s));
  if (prof::enabled()) prof::settle(INT32_MAX);
  return RH_OK;
}
