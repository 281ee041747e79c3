// ajet_probe_at: the probes of ajet_set_probes + ajet_probe for callers that cannot allocate
// device memory.  The point list and the values go through one file-scope device array of fixed
// size (each GPU the library runs on holds its own copy until the process ends); the call makes no
// hipMalloc, hipFree or other allocating call.  Calls that overlap on one GPU share that array
// and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ajet_internal.h"
#include "ajet_kernels.h"

namespace {

constexpr size_t kPointBytes = sizeof(int32_t) * 2 * AJET_PROBE_AT_MAX;            // 32 KiB
constexpr size_t kScratchBytes = kPointBytes + sizeof(double) * 5 * AJET_PROBE_AT_MAX;  // + 160 KiB
constexpr int kMaxDevices = 64;

}  // namespace

// Points (int32 pairs) at offset 0, values (5 doubles per point) at kPointBytes.  Only the
// device address from hipGetSymbolAddress is ever passed to a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with it current.
static hipError_t scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = p;
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

extern "C" int ajet_probe_at(ajet_ctx* ctx, const int32_t* points, int32_t n, double* values) {
  using namespace ajet;
  if (!ctx || n < 0 || (n > 0 && (!points || !values))) return AJET_EINVAL;
  ProbeView v{};
  if (!ctx_probe_view(ctx, &v)) return AJET_EINVAL;
  if (n > AJET_PROBE_AT_MAX) {
    ctx_set_error(ctx, "ajet_probe_at: " + std::to_string(n) + " points, at most " +
                           std::to_string((int)AJET_PROBE_AT_MAX));
    return AJET_EINVAL;
  }
  for (int32_t k = 0; k < n; ++k)
    if (points[2 * k] < 0 || points[2 * k] >= v.g.nz || points[2 * k + 1] < 0 || points[2 * k + 1] >= v.g.nr) {
      ctx_set_error(ctx, "ajet_probe_at: point outside the grid");
      return AJET_EINVAL;
    }
  if (n == 0) return AJET_OK;
  if (v.device < 0 || v.device >= kMaxDevices) {
    ctx_set_error(ctx, "ajet_probe_at: device index out of range");
    return AJET_EINVAL;
  }
  hipError_t e = hipSetDevice(v.device);
  void* base = nullptr;
  if (e == hipSuccess) e = scratch_address(v.device, &base);
  if (e != hipSuccess) {
    ctx_set_error(ctx, std::string("ajet_probe_at: ") + hipGetErrorString(e));
    return (int)e;
  }
  int* pts = reinterpret_cast<int*>(base);
  double* out = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kPointBytes);
  // every point and value slot used is written before it is read: nothing needs clearing
  e = hipMemcpyAsync(pts, points, sizeof(int32_t) * 2 * (size_t)n, hipMemcpyHostToDevice, v.stream);
  if (e == hipSuccess) e = launch_probe(v.q, v.g, pts, n, out, v.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(values, out, sizeof(double) * 5 * (size_t)n, hipMemcpyDeviceToHost, v.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
  if (e != hipSuccess) {
    ctx_set_error(ctx, std::string("ajet_probe_at: ") + hipGetErrorString(e));
    return (int)e;
  }
  return AJET_OK;
}
