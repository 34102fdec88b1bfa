// jsi_sort.hip - per-row ordering of the intersection pairs when rows are long.
//
// The overlap screen writes each row's pairs (row = slot * cells1 + i) at the
// row's offset in arrival order; intersect_collect_pairs' order
// (intersection.cu:298-307) needs them by j within the row.  Short rows are
// sorted in place by k_row_sort; long rows (the first round of a densely
// overlapping pair can hold hundreds of pairs per row) go through a
// segmented radix sort of the 20-bit cell ids, which is stable and
// deterministic, so the result is the same order.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "jsi_internal.hpp"

namespace jsi {

namespace {

__global__ void k_pairs_to_keys(const IPair* p, unsigned long long n, uint32_t* keys) {
  const unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) keys[e] = (uint32_t)p[e].j;
}

// Every pair of a row shares (slot, i): only j moves.
__global__ void k_keys_to_pairs(const uint32_t* keys, unsigned long long n, IPair* p) {
  const unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e].j = (int32_t)keys[e];
}

}  // namespace

size_t segmented_sort_tmp_bytes(unsigned long long n, unsigned long long nrows) {
  size_t bytes = 0;
  const unsigned long long* off = nullptr;
  (void)rocprim::segmented_radix_sort_keys(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (unsigned)n, (unsigned)nrows, off, off + 1, 0, 20);
  return bytes + 2 * sizeof(uint32_t) * (size_t)n + 512;
}

hipError_t segmented_sort_pairs(IPair* pairs, unsigned long long n, const unsigned long long* rowoff,
                                unsigned long long nrows, void* tmp, size_t tmp_bytes, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (n > 0xffffffffull || nrows > 0xffffffffull) return hipErrorInvalidValue;
  uint32_t* ka = reinterpret_cast<uint32_t*>(tmp);
  uint32_t* kb = ka + n;
  void* rt = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(kb + n) + 255) & ~uintptr_t(255));
  size_t rt_bytes = tmp_bytes - (size_t)(reinterpret_cast<char*>(rt) - reinterpret_cast<char*>(tmp));
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_pairs_to_keys, dim3(blocks), dim3(256), 0, s, pairs, n, ka);
  hipError_t e = rocprim::segmented_radix_sort_keys(rt, rt_bytes, ka, kb, (unsigned)n, (unsigned)nrows, rowoff,
                                                    rowoff + 1, 0, 20, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_keys_to_pairs, dim3(blocks), dim3(256), 0, s, kb, n, pairs);
  return hipGetLastError();
}

}  // namespace jsi
