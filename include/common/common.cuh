// common/common.cuh - the platform layer of the JSI_CAD_NEXT VERSION5 API
// (JSI_CAD_NEXT_VERSION5/src/common/common.cuh:1-35), served by
// libjsi_mi355x.so on an MI355X.  Same include guard, names and signatures as
// the reference header, so drivers that include it compile unchanged.
//   * cadMalloc / cadFree / cadMemcpy / cadMemset work on HBM (hipMalloc and
//     HIP copies; host or device pointers are accepted where the reference's
//     host fallback accepted host memory); 0 = ok, -1 = error.
//   * cuda_streams / current_stream_id are kept for source compatibility; the
//     library runs each query on its own HIP stream.
//   * compute_K and PrefixSumAdd run on the GPU; their pointers may be host
//     or device memory.
#ifndef _COMMON_UTILS_HPP
#define _COMMON_UTILS_HPP

#include <string>
#include <cstdio>

#include <cstddef>
#include <cstdint>

using CadStream = int;
enum class CadMemcpyKind { HostToDevice, DeviceToHost, DeviceToDevice };

inline constexpr int num_cuda_streams = 4;
extern int current_stream_id;
extern CadStream cuda_streams[num_cuda_streams];

// common.cu:25-33: 256-byte-aligned bump allocation from the workspace that
// init_cad made (device memory here); nullptr when it is exhausted.
void* allocate_from_workspace(size_t size);

void init_cad();

void deinit_cad();

int cadMalloc(void** ptr, size_t size);
int cadFree(void* ptr);
int cadMemcpy(void* dst, const void* src, size_t size, CadMemcpyKind kind);
int cadMemset(void* ptr, int value, size_t size);
bool cadIsAclEnabled();

// common.cu:137-144: inclusive prefix sum of size int32 values; returns the
// last sum (0 when size is 0).
int PrefixSumAdd(const int32_t* d_input, int32_t* d_output, size_t size);

// common.cu:149-172: d_deriv[3*i + {0,1,2}] = max |x|,|y|,|z| of Suu, Suv,
// Svv at point i of an 18-value nuv x nuv grid (evalWithCuda layout), and
// *d_out = (Muu/(n-1)^2 + 2*Muv/(n-1)^2 + Mvv/(n-1)^2) / 8 over the grid.
void compute_K(float* d_evalresults,  // evaluation results
               float* d_deriv,        // per point: max |second derivative| of uu, uv, vv
               float* d_out, int n);  // K
#endif
