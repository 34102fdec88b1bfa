// jsi_opabi.hip - the hotspot operator ABI of include/jsi_ascend_hotspots.h on
// the MI355X screens.  Semantics: jsi_ascend_hotspots.cpp:33-113 (reference).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/jsi_ascend_hotspots.h"
#include "../../include/jsi_mi355x.h"
#include "jsi_internal.hpp"
#include "jsi_runtime.hpp"

using namespace jsi;

namespace {

std::mutex g_op_mu;
jsi_ctx* g_op = nullptr;

int op_ctx() {
  if (g_op) return JSI_OK;
  return jsi_ctx_create(0, 0, &g_op);
}

int ceil_sqrt(long long n) {
  long long w = (long long)std::sqrt((double)n);
  while (w * w < n) ++w;
  while (w > 1 && (w - 1) * (w - 1) >= n) --w;
  return (int)w;
}

#define OCK(expr)                              \
  do {                                         \
    if ((expr) != hipSuccess) return -1;       \
  } while (0)
#define ORC(expr)                              \
  do {                                         \
    if ((expr) != JSI_OK) return -1;           \
  } while (0)

// Uploads one side and lays it out on a padded w x w grid.
int stage_side(jsi_ctx* c, const char* tag, const float* ctr, const float* rad, long long n, int w, float4** C,
               float4** R) {
  float *raw_c, *raw_r = nullptr;
  std::string t(tag);
  ORC(c->buf((t + ".rc").c_str(), (size_t)3 * n, &raw_c));
  ORC(c->h2d(raw_c, ctr, sizeof(float) * 3 * (size_t)n));
  if (rad) {
    ORC(c->buf((t + ".rr").c_str(), (size_t)3 * n, &raw_r));
    ORC(c->h2d(raw_r, rad, sizeof(float) * 3 * (size_t)n));
  }
  ORC(c->buf((t + ".C").c_str(), (size_t)w * w, C));
  if (rad) ORC(c->buf((t + ".R").c_str(), (size_t)w * w, R));
  OCK(launch_pack_cells(raw_c, raw_r, n, w, *C, rad ? *R : nullptr, c->stream));
  return JSI_OK;
}

}  // namespace

extern "C" int jsi_ascend_distance_topk_mask(const float* center1, size_t cnt1, const float* center2, size_t cnt2,
                                             int topk, int* mask1, int* mask2, float* kth_distance) {
  if (!center1 || !center2 || !mask1 || !mask2 || !kth_distance || cnt1 == 0 || cnt2 == 0 || topk <= 0) return -1;
  if (cnt1 > (1u << 20) || cnt2 > (1u << 20) || topk > (1 << 21)) return -1;  // 20-bit cell ids
  std::lock_guard<std::mutex> lock(g_op_mu);
  if (op_ctx() != JSI_OK) return -1;
  jsi_ctx* c = g_op;
  if (hipSetDevice(c->device) != hipSuccess) return -1;
  const int w1 = ceil_sqrt((long long)cnt1), w2 = ceil_sqrt((long long)cnt2);
  float4 *C1, *C2, *unused = nullptr;
  ORC(stage_side(c, "op1", center1, nullptr, (long long)cnt1, w1, &C1, &unused));
  ORC(stage_side(c, "op2", center2, nullptr, (long long)cnt2, w2, &C2, &unused));
  int* su;
  ORC(c->buf("op.su", 2, &su));
  OCK(launch_fill_i32(su, 0, 2, c->stream));
  const unsigned long long total = (unsigned long long)cnt1 * cnt2;
  const int kk = (int)std::min<unsigned long long>((unsigned long long)topk, total);
  Cand* rec;
  long long n = 0, got = 0;
  ORC(c->dist_topk<float>(1, su, su, w1, 1, C1, w2, 1, C2, kk, 1, &rec, &n, &got));
  if (got < kk) return -1;
  Cand last;
  ORC(c->d2h(&last, rec + (kk - 1), sizeof(Cand)));
  float d2;
  const uint32_t bits = (uint32_t)last.hi;
  std::memcpy(&d2, &bits, 4);
  const float K = std::sqrt(d2);  // k-th smallest distance (sqrt is monotone)
  // largest d2 whose rounded square root is <= K: the mask test dist <= K is d2 <= B
  float B = d2;
  for (int it = 0; it < 8; ++it) {
    const float nb = std::nextafter(B, INFINITY);
    if (std::sqrt(nb) <= K)
      B = nb;
    else
      break;
  }
  *kth_distance = K;
  // culled list of tile pairs that can hold a pair with d2 <= B, then masks
  const int tw1 = (w1 + kTile - 1) / kTile, tw2 = (w2 + kTile - 1) / kTile;
  float4 *lo1, *hi1, *lo2, *hi2;
  float* T;
  uint64_t* list;
  unsigned* cnt;
  int *m1, *m2;
  ORC(c->buf("op.lo1", (size_t)tw1 * tw1, &lo1));
  ORC(c->buf("op.hi1", (size_t)tw1 * tw1, &hi1));
  ORC(c->buf("op.lo2", (size_t)tw2 * tw2, &lo2));
  ORC(c->buf("op.hi2", (size_t)tw2 * tw2, &hi2));
  ORC(c->buf("op.T", 4, &T));
  ORC(c->buf("op.list", (size_t)tw1 * tw1 * tw2 * tw2, &list));
  ORC(c->buf("op.cnt", 4, &cnt));
  ORC(c->buf("op.m1", cnt1, &m1));
  ORC(c->buf("op.m2", cnt2, &m2));
  OCK(launch_tilebox<float>(C1, nullptr, 1, w1, 0, lo1, hi1, c->stream));
  OCK(launch_tilebox<float>(C2, nullptr, 1, w2, 0, lo2, hi2, c->stream));
  ORC(c->h2d(T, &B, sizeof(float)));
  OCK(launch_fill_u32(cnt, 0u, 1, c->stream));
  OCK(launch_dist_list<float>(1, su, su, w1, w2, lo1, hi1, lo2, hi2, T, 1, list, cnt, c->stream));
  OCK(launch_fill_i32(m1, 0, cnt1, c->stream));
  OCK(launch_fill_i32(m2, 0, cnt2, c->stream));
  OCK(launch_pair_mask(0, list, nullptr, cnt, w1, w2, C1, nullptr, C2, nullptr, T, (long long)cnt1, (long long)cnt2,
                       m1, m2, c->grid, c->stream));
  ORC(c->d2h(mask1, m1, sizeof(int) * cnt1));
  ORC(c->d2h(mask2, m2, sizeof(int) * cnt2));
  return 0;
}

extern "C" int jsi_ascend_intersection_mask(const float* center1, const float* radius1, int cnt1,
                                            const float* center2, const float* radius2, int cnt2, int* mask1,
                                            int* mask2) {
  if (!center1 || !radius1 || !center2 || !radius2 || !mask1 || !mask2 || cnt1 <= 0 || cnt2 <= 0) return -1;
  if (cnt1 > (1 << 20) || cnt2 > (1 << 20)) return -1;
  std::lock_guard<std::mutex> lock(g_op_mu);
  if (op_ctx() != JSI_OK) return -1;
  jsi_ctx* c = g_op;
  if (hipSetDevice(c->device) != hipSuccess) return -1;
  const int w1 = ceil_sqrt(cnt1), w2 = ceil_sqrt(cnt2);
  float4 *C1, *R1, *C2, *R2;
  ORC(stage_side(c, "op1", center1, radius1, cnt1, w1, &C1, &R1));
  ORC(stage_side(c, "op2", center2, radius2, cnt2, w2, &C2, &R2));
  int* su;
  ORC(c->buf("op.su", 2, &su));
  OCK(launch_fill_i32(su, 0, 2, c->stream));
  const int tw1 = (w1 + kTile - 1) / kTile, tw2 = (w2 + kTile - 1) / kTile;
  const unsigned long long ntp = (unsigned long long)tw1 * tw1 * tw2 * tw2;
  float4 *lo1, *hi1, *lo2, *hi2;
  uint64_t* list;
  unsigned long long* cnt;
  int *m1, *m2;
  ORC(c->buf("op.lo1", (size_t)tw1 * tw1, &lo1));
  ORC(c->buf("op.hi1", (size_t)tw1 * tw1, &hi1));
  ORC(c->buf("op.lo2", (size_t)tw2 * tw2, &lo2));
  ORC(c->buf("op.hi2", (size_t)tw2 * tw2, &hi2));
  ORC(c->buf("op.list", ntp, &list));
  ORC(c->buf("op.cnt64", 2, &cnt));
  ORC(c->buf("op.m1", (size_t)cnt1, &m1));
  ORC(c->buf("op.m2", (size_t)cnt2, &m2));
  OCK(launch_tilebox<float>(C1, R1, 1, w1, 1, lo1, hi1, c->stream));
  OCK(launch_tilebox<float>(C2, R2, 1, w2, 1, lo2, hi2, c->stream));
  OCK(launch_fill_u32(reinterpret_cast<unsigned*>(cnt), 0u, 2, c->stream));
  OCK(launch_isect_list<float>(1, su, su, w1, w2, lo1, hi1, lo2, hi2, 1, list, cnt, ntp, c->stream));
  OCK(launch_fill_i32(m1, 0, (unsigned long long)cnt1, c->stream));
  OCK(launch_fill_i32(m2, 0, (unsigned long long)cnt2, c->stream));
  OCK(launch_pair_mask(1, list, cnt, nullptr, w1, w2, C1, R1, C2, R2, nullptr, cnt1, cnt2, m1, m2, c->grid,
                       c->stream));
  ORC(c->d2h(mask1, m1, sizeof(int) * (size_t)cnt1));
  ORC(c->d2h(mask2, m2, sizeof(int) * (size_t)cnt2));
  return 0;
}
