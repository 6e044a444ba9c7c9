// kNN that allocates nothing: fuel_amd_knn_device_noalloc.  Its candidate
// slots and list counts live in `scratch`, a file-scope device array that
// every GPU the library runs on holds until the process ends.  Host code
// never uses the array's name as an address: the address of each device's copy
// comes from hipGetSymbolAddress, taken once per device on that device's first
// call (with the device current) and kept.
#include <hip/hip_runtime.h>

#include <mutex>

#include "api_internal.h"
#include "kernels.h"

namespace fa {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (the device is current).
int32_t scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(FUEL_AMD_ERR_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(FUEL_AMD_ERR_DEVICE, "scratch array has no device address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FUEL_AMD_OK;
}

}  // namespace
}  // namespace fa

using namespace fa;

extern "C" int32_t fuel_amd_knn_device_noalloc(fuel_amd_index* ix, const void* d_query, int64_t m, int64_t ldq,
                                               int32_t k, int32_t exclude_self, int64_t* d_out_idx, void* d_out_dist,
                                               void* stream) {
  if (!ix) return fail(FUEL_AMD_ERR_INVALID, "index must not be NULL");
  if (m < 0 || k < 0) return fail(FUEL_AMD_ERR_INVALID, "m and k must be non-negative");
  if (exclude_self < 0 || exclude_self > 2) return fail(FUEL_AMD_ERR_INVALID, "bad exclude_self");
  if (ldq < ix->d) return fail(FUEL_AMD_ERR_INVALID, "ldq must be >= d");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  void* p = nullptr;
  rc = scratch_address(ix->device, &p);
  if (rc) return rc;
  return knn_run_scratch(ix, d_query, m, ldq, k, exclude_self, d_out_idx, d_out_dist,
                         static_cast<hipStream_t>(stream), p, kScratchBytes);
}
