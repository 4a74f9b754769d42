// An entry point that allocates nothing: rh_divergence_max_noalloc, the
// divergence gate of rh_divergence_max on a fixed device scratch array.
//
// Callers that cannot call hipMalloc (tight serving loops, harnesses that
// forbid allocation) get the same validation and maximum as
// rh_divergence_max, with its reduction partials, error flags and result in
// `scratch`: a file-scope device array of kNoAllocScratchBytes that every GPU
// the library runs on holds until the process ends.  The array's name is a
// host-side stand-in in host code; the device address of each GPU's copy is
// taken once per device with hipGetSymbolAddress and only that address is
// used.  Calls that overlap on one GPU share (and race on) that GPU's array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rh {
namespace {

// Reduction partials for up to kMaxBlocks blocks (four double-double sums
// and four maxima per block, the layout of block_publish), then the error
// flag word and the result scalars.
constexpr size_t kNoAllocPddBytes = sizeof(DD) * 4 * kMaxBlocks;
constexpr size_t kNoAllocPmxBytes = sizeof(double) * 4 * kMaxBlocks;
constexpr size_t kNoAllocTailBytes = 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocPddBytes + kNoAllocPmxBytes + kNoAllocTailBytes;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (taken on first use).
rh_status scratch_ptr(void** out) {
  int dev = 0;
  RH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_last_error("device index beyond the no-allocation table", "rh_divergence_max_noalloc");
    return RH_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_last_error(e != hipSuccess ? hipGetErrorString(e) : "null symbol address", "hipGetSymbolAddress");
      return RH_DEVICE_ERROR;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return RH_OK;
}

// Zeroes the flag word and result slots (a kernel, not a memset, so the
// sequence holds under stream capture as well).
__global__ void k_clear_words(unsigned* words, int n) {
  const int i = threadIdx.x;
  if (i < n) words[i] = 0u;
}

}  // namespace
}  // namespace rh

// rh_divergence_max (src/pressure.rs:330-342 with src/operator.rs:513-568)
// without allocation; same checks, same result.
extern "C" rh_status rh_divergence_max_noalloc(const rh_grid* grid, rh_precision precision, const void* d_u,
                                               const void* d_v, const void* d_w, double* out_max) {
  using namespace rh;
  if (!grid || !d_u || !d_v || !d_w || !out_max) return RH_INVALID_ARGUMENT;
  rh_status s = validate_grid(grid, nullptr);
  if (s) return s;
  if ((s = check_indexable(grid))) return s;
  if (precision != RH_F32 && precision != RH_F64) return RH_INVALID_ARGUMENT;
  const DevGrid g = make_dev_grid(*grid, 1.0);
  // Launches use at most kMaxBlocks blocks (a larger grid gives each block a
  // longer run of rows), so the partials of every grid that passes the
  // indexing check above fit the scratch array; the check is kept explicit.
  if (grid_blocks(g.ny * g.nz) > kMaxBlocks) {
    set_last_error("reduction needs more blocks than the no-allocation scratch holds", "rh_divergence_max_noalloc");
    return RH_BUFFER_LIMIT;
  }
  void* base = nullptr;
  if ((s = scratch_ptr(&base))) return s;
  char* p = (char*)base;
  Reduce red;
  red.pdd = (DD*)p;
  red.pmx = (double*)(p + kNoAllocPddBytes);
  unsigned* tail = (unsigned*)(p + kNoAllocPddBytes + kNoAllocPmxBytes);
  red.flags = tail;
  red.out = (double*)(tail + 16);  // 64 bytes in: one flag word, then result doubles
  const hipStream_t st = nullptr;   // the null stream: no stream object to create
  k_clear_words<<<1, 64, 0, st>>>(tail, (int)(kNoAllocTailBytes / sizeof(unsigned)));
  if (precision == RH_F32) {
    launch_validate<float>(g, d_u, d_v, d_w, 1, red.flags, st);
    launch_divergence_max<float>(g, d_u, d_v, d_w, red, st);
  } else {
    launch_validate<double>(g, d_u, d_v, d_w, 1, red.flags, st);
    launch_divergence_max<double>(g, d_u, d_v, d_w, red, st);
  }
  RH_HIP(hipGetLastError());
  unsigned f = 0;
  double mx = 0.0;
  RH_HIP(hipMemcpy(&f, red.flags, sizeof(f), hipMemcpyDeviceToHost));
  RH_HIP(hipMemcpy(&mx, red.out, sizeof(mx), hipMemcpyDeviceToHost));
  if (f) return flags_to_status(f);
  *out_max = mx;
  return RH_OK;
}
