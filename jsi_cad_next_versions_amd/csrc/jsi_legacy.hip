// jsi_legacy.hip - the VERSION5 C++ query API (include/jsi_cad.hpp) on top of
// the C ABI.  One process-wide context plays the role of init_cad's
// workspace; the workspace itself is a bump allocator over one device
// allocation, as allocate_from_workspace (common.cu:25-33) is over host memory.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/jsi_cad.hpp"
#include "../../include/jsi_mi355x.h"

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
