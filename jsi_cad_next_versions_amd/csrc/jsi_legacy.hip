// jsi_legacy.hip - the VERSION5 C++ query API (include/jsi_cad.hpp) on top of
// the C ABI.  One process-wide context plays the role of init_cad's
// workspace; the workspace itself is a bump allocator over one device
// allocation, as allocate_from_workspace (common.cu:25-33) is over host memory.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "../../include/common/common.cuh"
#include "../../include/jsi_cad.hpp"
#include "../../include/jsi_mi355x.h"

// common.cu:21-22: the reference's stream bookkeeping; kept for drivers that
// name it.  Every query runs on the context's own HIP stream.
int current_stream_id = 0;
CadStream cuda_streams[num_cuda_streams] = {0, 1, 2, 3};

namespace {

jsi_ctx* g_ctx = nullptr;
void* g_ws = nullptr;
size_t g_ws_size = 4ull << 30;  // common.cu:18
std::atomic<size_t> g_ws_off(0);
// Refinement nuv of the intersection path: a process-wide static that halves
// across calls and is never reset (intersection.cu:36,370).
int g_isect_nuv = 16;

bool env_on(const char* name, bool dflt) {
  const char* v = std::getenv(name);
  if (!v) return dflt;
  return std::strcmp(v, "0") != 0;
}

void ensure_ctx() {
  if (g_ctx) return;
  int rc = jsi_ctx_create(0, 0, &g_ctx);
  if (rc != JSI_OK) {
    printf("jsi: context creation failed: %s\n", jsi_last_error());
    std::exit(1);
  }
}

int register_task(const EvalAndConstructTask& t) {
  int h = -1;
  int rc = jsi_surface_register_device(g_ctx, t.d_ivec, t.d_data, t.d_ndata, &h);
  if (rc != JSI_OK) {
    printf("ERROR! jsi surface: %s\n", jsi_last_error());
    std::exit(-1);
  }
  return h;
}

jsi_region region_of(const EvalAndConstructTask& t) {
  jsi_region r;
  r.ustart = t.ustart;
  r.vstart = t.vstart;
  r.ustop = t.ustop;
  r.vstop = t.vstop;
  r.nuv = t.nuv;
  r.side = t.eval_task_side;
  return r;
}

void print_timing(const jsi_result& res) {
  if (!std::getenv("CAD_TIMING_BREAKDOWN")) return;
  const double kernel = res.ms_eval + res.ms_screen;
  const double orch = res.ms_total - kernel > 0 ? res.ms_total - kernel : 0.0;
  printf("CAD_TIMING end_to_end_ms=%.6f kernel_ms=%.6f evaluation_ms=%.6f screening_ms=%.6f orchestration_ms=%.6f\n",
         res.ms_total, kernel, res.ms_eval, res.ms_screen, orch);
}

}  // namespace

void init_cad() {
  ensure_ctx();
  if (const char* s = std::getenv("JSI_WORKSPACE_BYTES")) g_ws_size = std::strtoull(s, nullptr, 10);
  if (!g_ws && hipMalloc(&g_ws, g_ws_size) != hipSuccess) {
    printf("alloc workspace failed\n");
    std::exit(1);
  }
  g_ws_off.store(0);
}

void deinit_cad() {
  if (g_ws) (void)hipFree(g_ws);
  g_ws = nullptr;
  g_ws_off.store(0);
  if (g_ctx) jsi_ctx_destroy(g_ctx);
  g_ctx = nullptr;
}

void* allocate_from_workspace(size_t size) {
  if (size % 256 != 0) size += 256 - size % 256;
  const size_t off = g_ws_off.fetch_add(size);
  if (!g_ws || off + size > g_ws_size) {
    printf("Error: Out of workspace memory!\n");
    return nullptr;
  }
  return static_cast<char*>(g_ws) + off;
}

float minimum_distance(const EvalAndConstructTask& t1, const EvalAndConstructTask& t2) {
  ensure_ctx();
  const int h1 = register_task(t1), h2 = register_task(t2);
  jsi_params p;
  jsi_params_default(&p);
  jsi_region r1 = region_of(t1), r2 = region_of(t2);
  jsi_result res;
  const int rc = jsi_min_distance(g_ctx, h1, &r1, h2, &r2, &p, &res, nullptr, 0);
  jsi_surface_release(g_ctx, h1);
  jsi_surface_release(g_ctx, h2);
  if (rc != JSI_OK) {
    printf("ERROR! minimum_distance: %s\n", jsi_last_error());
    std::exit(-1);
  }
  for (int r = 0; r < res.rounds; ++r) {
    printf("Round: %d (pair refine, aligned_slots=%lld)\n", r + 1, (long long)res.round[r].slots);
    // sqrt of the round's k-th d2, widened from float (distance_analysis.cu:337,384)
    printf("[current_min_dist] %lf\n", (double)sqrtf((float)res.round[r].kth_d2));
    printf("  pairs to refine: %lld\n", (long long)res.round[r].npairs);
  }
  print_timing(res);
  return (float)res.distance;
}

void intersection(EvalAndConstructTask t1, EvalAndConstructTask t2) {
  ensure_ctx();
  const int h1 = register_task(t1), h2 = register_task(t2);
  jsi_params p;
  jsi_params_default(&p);
  p.use_compute_k = env_on("CAD_INTERSECT_USE_K", true) ? 1 : 0;
  p.isect_refine_nuv = g_isect_nuv;
  jsi_region r1 = region_of(t1), r2 = region_of(t2);
  jsi_result res;
  const int rc = jsi_intersect(g_ctx, h1, &r1, h2, &r2, &p, &res, nullptr, 0);
  jsi_surface_release(g_ctx, h1);
  jsi_surface_release(g_ctx, h2);
  if (rc != JSI_OK) {
    printf("ERROR! intersection: %s\n", jsi_last_error());
    std::exit(-1);
  }
  g_isect_nuv = res.isect_refine_nuv;
  if (p.use_compute_k) {
    // the first two evaluations of a call are the initial regions of side 1
    // and side 2 (intersection.cu:187-205)
    for (int side = 1; side <= 2; ++side) {
      const float raw = (float)res.isect_k_raw[side - 1];
      const float scale = (float)res.isect_k_scale[side - 1];
      float pad = raw * scale * scale;
      if (!std::isfinite(pad) || pad < 0.0f) pad = 0.0f;
      printf("Compute_K initial side%d raw=%g scale=%g padding=%g\n", side, raw, scale, pad);
    }
  }
  printf("Compute_K AABB padding: %s\n", p.use_compute_k ? "on" : "off");
  for (int r = 0; r < res.rounds; ++r) {
    printf("Round: %d (pair refine, aligned_slots=%lld)\n", r + 1, (long long)res.round[r].slots);
    printf("Round %d: overlap pairs = %lld\n", r + 1, (long long)res.round[r].npairs);
  }
  fflush(stdout);
  print_timing(res);
}

void evalWithCuda(int* ivec, float* data, int ndata, float ustep, float vstep, float ustart, float vstart, int nuv,
                  float* results) {
  ensure_ctx();
  int hiv[7];
  if (hipMemcpy(hiv, ivec, sizeof(hiv), hipMemcpyDefault) != hipSuccess) {
    printf("evalWithCuda: cannot read ivec\n");
    std::exit(1);
  }
  std::vector<float> hdata((size_t)ndata);
  if (hipMemcpy(hdata.data(), data, sizeof(float) * (size_t)ndata, hipMemcpyDefault) != hipSuccess) {
    printf("evalWithCuda: cannot read data\n");
    std::exit(1);
  }
  int h = -1;
  if (jsi_surface_upload(g_ctx, hiv, hdata.data(), ndata, &h) != JSI_OK) {
    printf("evalWithCuda: %s\n", jsi_last_error());
    std::exit(1);
  }
  std::vector<float> out((size_t)nuv * nuv * 18);
  if (jsi_eval_grid_f32(g_ctx, h, ustep, vstep, ustart, vstart, nuv, out.data()) != JSI_OK) {
    printf("evalWithCuda: %s\n", jsi_last_error());
    std::exit(1);
  }
  jsi_surface_release(g_ctx, h);
  (void)hipMemcpy(results, out.data(), sizeof(float) * out.size(), hipMemcpyDefault);
}

int cadMalloc(void** ptr, size_t size) {
  if (!ptr) return -1;
  if (size == 0) {
    *ptr = nullptr;
    return 0;
  }
  return hipMalloc(ptr, size) == hipSuccess ? 0 : -1;
}

int cadFree(void* ptr) {
  if (!ptr) return 0;
  return hipFree(ptr) == hipSuccess ? 0 : -1;
}

int cadMemcpy(void* dst, const void* src, size_t size, CadMemcpyKind kind) {
  if (size == 0) return 0;
  hipMemcpyKind k = hipMemcpyDefault;
  if (kind == CadMemcpyKind::HostToDevice) k = hipMemcpyHostToDevice;
  if (kind == CadMemcpyKind::DeviceToHost) k = hipMemcpyDeviceToHost;
  if (kind == CadMemcpyKind::DeviceToDevice) k = hipMemcpyDeviceToDevice;
  return hipMemcpy(dst, src, size, k) == hipSuccess ? 0 : -1;
}

int cadMemset(void* ptr, int value, size_t size) {
  if (size == 0) return 0;
  return hipMemset(ptr, value, size) == hipSuccess ? 0 : -1;
}

bool cadIsAclEnabled() { return false; }

// ---------------------------------------------------------------------------
// compute_K and PrefixSumAdd (common.cu:137-172) on the GPU.  Their pointers
// may be host or device memory (the reference passes host vectors); inputs are
// staged into device buffers and results copied back with hipMemcpyDefault.
// ---------------------------------------------------------------------------
namespace {

// max(a, b) as CadMax (common.cu:132-135): a > b ? a : b
__device__ __forceinline__ float cad_max(float a, float b) { return a > b ? a : b; }

// Per point: the largest |coordinate| of Suu, Suv, Svv (values 9-17 of the 18),
// written to deriv[3*i + {0,1,2}] and max-reduced into m[0..2] as float bits
// (non-negative floats order as their bit patterns).
__global__ __launch_bounds__(256) void k_legacy_second_deriv(const float* __restrict__ ev, long long npts,
                                                           float* __restrict__ deriv, unsigned* m) {
  __shared__ unsigned red[3][256];
  unsigned best[3] = {0u, 0u, 0u};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npts;
       i += (long long)gridDim.x * blockDim.x) {
    const float* p = ev + i * 18;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float v = cad_max(fabsf(p[9 + 3 * q]), cad_max(fabsf(p[10 + 3 * q]), fabsf(p[11 + 3 * q])));
      deriv[i * 3 + q] = v;
      best[q] = max(best[q], __float_as_uint(v));
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = best[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = max(red[q][threadIdx.x], red[q][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) atomicMax(&m[q], red[q][0]);
  }
}

// K = (Muu/(s*s) + 2*Muv/(s*s) + Mvv/(s*s)) / 8 with s = n - 1 (1 when n <= 1),
// the maxima starting at -FLT_MAX as in the reference (so an empty grid gives
// its value too).
__global__ void k_legacy_k(const unsigned* m, long long npts, int n, float* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float mx[3];
  for (int q = 0; q < 3; ++q) mx[q] = npts > 0 ? __uint_as_float(m[q]) : -3.402823466e+38f;
  const int side = n > 1 ? (n - 1) : 1;
  const float s2 = (float)(side * side);
  *out = (mx[0] / s2 + 2.0f * mx[1] / s2 + mx[2] / s2) / 8.0f;
}

// Inclusive scan of n int32 values by one workgroup: each thread scans a
// contiguous chunk, the chunk totals are scanned in LDS, then each chunk adds
// its offset.  Sums wrap modulo 2^32 (the reference's int accumulator).
__global__ __launch_bounds__(1024) void k_legacy_prefix(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                       long long n) {
  __shared__ unsigned tot[1024];
  const long long chunk = (n + blockDim.x - 1) / blockDim.x;
  const long long b = (long long)threadIdx.x * chunk, e = b + chunk < n ? b + chunk : n;
  unsigned acc = 0u;
  for (long long i = b; i < e; ++i) acc += (unsigned)in[i];
  tot[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned w = 1; w < blockDim.x; w <<= 1) {  // Hillis-Steele scan of the chunk totals
    const unsigned v = threadIdx.x >= w ? tot[threadIdx.x - w] : 0u;
    __syncthreads();
    tot[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned run = threadIdx.x > 0 ? tot[threadIdx.x - 1] : 0u;
  for (long long i = b; i < e; ++i) {
    run += (unsigned)in[i];
    out[i] = (int32_t)run;
  }
}

template <class X>
struct DevTmp {  // scoped device allocation of the legacy helpers
  X* p = nullptr;
  explicit DevTmp(size_t n) {
    if (n && hipMalloc(&p, n * sizeof(X)) != hipSuccess) p = nullptr;
  }
  ~DevTmp() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

void compute_K(float* d_evalresults, float* d_deriv, float* d_out, int n) {
  ensure_ctx();
  const long long npts = n > 0 ? (long long)n * n : 0;
  DevTmp<float> ev(npts * 18), der(npts * 3), k(1);
  DevTmp<unsigned> m(3);
  if ((npts && (!ev.p || !der.p)) || !k.p || !m.p) {
    printf("compute_K: device allocation failed\n");
    std::exit(1);
  }
  bool ok = hipMemset(m.p, 0, 3 * sizeof(unsigned)) == hipSuccess;
  if (npts) {
    ok = ok && hipMemcpy(ev.p, d_evalresults, npts * 18 * sizeof(float), hipMemcpyDefault) == hipSuccess;
    const long long blocks = std::min<long long>((npts + 255) / 256, 4096);
    hipLaunchKernelGGL(k_legacy_second_deriv, dim3((unsigned)blocks), dim3(256), 0, 0, ev.p, npts, der.p, m.p);
    ok = ok && hipGetLastError() == hipSuccess;
  }
  hipLaunchKernelGGL(k_legacy_k, dim3(1), dim3(64), 0, 0, m.p, npts, n, k.p);
  ok = ok && hipGetLastError() == hipSuccess;
  if (npts) ok = ok && hipMemcpy(d_deriv, der.p, npts * 3 * sizeof(float), hipMemcpyDefault) == hipSuccess;
  ok = ok && hipMemcpy(d_out, k.p, sizeof(float), hipMemcpyDefault) == hipSuccess;
  if (!ok) {
    printf("compute_K: HIP error\n");
    std::exit(1);
  }
}

int PrefixSumAdd(const int32_t* d_input, int32_t* d_output, size_t size) {
  if (size == 0) return 0;
  ensure_ctx();
  DevTmp<int32_t> in(size), out(size);
  if (!in.p || !out.p) {
    printf("PrefixSumAdd: device allocation failed\n");
    std::exit(1);
  }
  bool ok = hipMemcpy(in.p, d_input, size * sizeof(int32_t), hipMemcpyDefault) == hipSuccess;
  hipLaunchKernelGGL(k_legacy_prefix, dim3(1), dim3(1024), 0, 0, in.p, out.p, (long long)size);
  ok = ok && hipGetLastError() == hipSuccess;
  ok = ok && hipMemcpy(d_output, out.p, size * sizeof(int32_t), hipMemcpyDefault) == hipSuccess;
  int32_t last = 0;
  ok = ok && hipMemcpy(&last, out.p + (size - 1), sizeof(int32_t), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    printf("PrefixSumAdd: HIP error\n");
    std::exit(1);
  }
  return last;
}
