// jsi_noalloc.hip - jsi_min_distance_noalloc: the minimum-distance query of
// jsi_min_distance for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Every device buffer of the query (cell centres, tile boxes, candidate
// records, selection state, refinement maps, the FP64 copy of a surface that
// has none yet) is carved from `scratch`, one file-scope device array per GPU
// of kScratchBytes.  The host never names the array itself in a kernel
// argument or copy: its device address is looked up once per device with
// hipGetSymbolAddress (the first call on that device, with the device
// current) and that pointer is handed to the kernels.  Calls on one GPU share
// its array, so calls that overlap on one GPU race; the array keeps the last
// call's contents, and every buffer that must start cleared is cleared by a
// fill kernel at the start of its use.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/jsi_mi355x.h"
#include "jsi_runtime.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(256) << 20;     // 256 MiB per GPU
constexpr unsigned long long kScratchCand = 1ull << 20;  // candidate records per screening pass (16 MiB)
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this GPU's scratch (looked up once per device).
static int scratch_ptr(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return JSI_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return JSI_ERR_DEVICE;
    }
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return JSI_OK;
}

extern "C" int jsi_min_distance_noalloc(jsi_ctx* ctx, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                                        const jsi_params* params, jsi_result* result, jsi_pair* pairs_out,
                                        int pairs_cap) {
  if (!ctx) return JSI_ERR_INVALID;
  if (hipSetDevice(ctx->device) != hipSuccess) return JSI_ERR_DEVICE;
  void* p = nullptr;
  const int rc = scratch_ptr(ctx->device, &p);
  if (rc != JSI_OK) return rc;
  return min_distance_in_scratch(ctx, h1, r1, h2, r2, params, result, pairs_out, pairs_cap, p, kScratchBytes,
                                 kScratchCand);
}
