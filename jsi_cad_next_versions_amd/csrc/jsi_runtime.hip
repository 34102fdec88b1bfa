// jsi_runtime.hip - host runtime of libjsi_mi355x.so: contexts, device
// buffers, resident surfaces, the distance / intersection round loops and the
// C ABI of include/jsi_mi355x.h.  All compute runs in jsi_kernels.hip; the
// host only sequences kernels and reads back counts and results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jsi_mi355x.h"
#include "jsi_internal.hpp"
#include "jsi_runtime.hpp"

using namespace jsi;

namespace {

thread_local char g_err[512] = "";

void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define HCK(expr)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return JSI_ERR_DEVICE;                                                               \
    }                                                                                      \
  } while (0)

#define RCK(expr)             \
  do {                        \
    int r_ = (expr);          \
    if (r_ != JSI_OK) return r_; \
  } while (0)

unsigned long long next_pow2(unsigned long long x) {
  unsigned long long p = 1;
  while (p < x) p <<= 1;
  return p;
}

constexpr unsigned kSmallCap = 16384;               // records sorted after selection
constexpr long long kSlotListMin = 1 << 18;  // intersection rounds with this many slots cull them first
constexpr size_t kSlotMaskBytes = size_t(256) << 20;  // row hit masks kept between the slot passes, at most


}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
int jsi_ctx::buf_raw(const char* name, size_t bytes, void** out, int fill_int) {
  if (scratch_.base) {
    DevBuf& b = scratch_.bufs[name];
    if (b.bytes < bytes) {
      const size_t off = (scratch_.off + 255) & ~(size_t)255;
      const size_t want = std::max(bytes, (size_t)256);
      if (off + want > scratch_.cap) {
        set_err("scratch of %zu bytes exceeded: buffer %s needs %zu bytes at offset %zu (this query is larger "
                "than the allocation-free entry point supports)", scratch_.cap, name, want, off);
        return JSI_ERR_NOMEM;
      }
      b.p = scratch_.base + off;
      b.bytes = want;
      scratch_.off = off + want;
      if (fill_int != 0) {
        hipError_t f = launch_fill_i32(static_cast<int*>(b.p), fill_int, want / sizeof(int), stream);
        if (f != hipSuccess) {
          set_err("fill %s: %s", name, hipGetErrorString(f));
          return JSI_ERR_DEVICE;
        }
      }
    }
    *out = b.p;
    return JSI_OK;
  }
  DevBuf& b = bufs[name];
  if (b.bytes < bytes) {
    if (b.p) {
      HCK(hipStreamSynchronize(stream));
      HCK(hipFree(b.p));
      pool->release(b.bytes);
      b.p = nullptr;
      b.bytes = 0;
    }
    size_t want = std::max(bytes + bytes / 2, (size_t)4096);
    if (!pool->reserve(want)) {
      want = std::max(bytes, (size_t)256);
      if (!pool->reserve(want)) {
        set_err("device memory cap exceeded: buffer %s needs %zu bytes, %zu of %zu in use", name, want,
                pool->used.load(), pool->cap);
        return JSI_ERR_NOMEM;
      }
    }
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      pool->release(want);
      set_err("hipMalloc(%zu) for %s failed: %s", want, name, hipGetErrorString(e));
      b.p = nullptr;
      return JSI_ERR_NOMEM;
    }
    b.bytes = want;
    if (fill_int != 0) {
      hipError_t f = launch_fill_i32(static_cast<int*>(b.p), fill_int, want / sizeof(int), stream);
      if (f != hipSuccess) {
        set_err("fill %s: %s", name, hipGetErrorString(f));
        return JSI_ERR_DEVICE;
      }
    }
  }
  *out = b.p;
  return JSI_OK;
}

int jsi_ctx::pinned(size_t bytes, void** out) {
  if (bytes > pinned_bytes_) {
    if (pinned_) HCK(hipHostFree(pinned_));
    pinned_ = nullptr;
    pinned_bytes_ = 0;
    const size_t want = std::max<size_t>(bytes, (size_t)1 << 20);
    HCK(hipHostMalloc(&pinned_, want, hipHostMallocDefault));
    pinned_bytes_ = want;
  }
  *out = pinned_;
  return JSI_OK;
}

int jsi_ctx::d2h(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return JSI_OK;
  if (pin_small_ && bytes <= kPinSmall && dst != pin_small_) {
    HCK(hipMemcpyAsync(pin_small_, src, bytes, hipMemcpyDeviceToHost, stream));
    HCK(hipStreamSynchronize(stream));
    std::memcpy(dst, pin_small_, bytes);
    return JSI_OK;
  }
  HCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
  HCK(hipStreamSynchronize(stream));
  return JSI_OK;
}

int jsi_ctx::d2h_many(std::initializer_list<Copy> copies) {
  size_t off = 0;
  for (const Copy& c : copies) off += (c.bytes + 15) & ~(size_t)15;
  if (!pin_small_ || off > kPinSmall) {  // one by one
    for (const Copy& c : copies) RCK(d2h(c.dst, c.src, c.bytes));
    return JSI_OK;
  }
  off = 0;
  for (const Copy& c : copies) {
    if (c.bytes) HCK(hipMemcpyAsync(static_cast<char*>(pin_small_) + off, c.src, c.bytes, hipMemcpyDeviceToHost, stream));
    off += (c.bytes + 15) & ~(size_t)15;
  }
  HCK(hipStreamSynchronize(stream));
  off = 0;
  for (const Copy& c : copies) {
    if (c.bytes) std::memcpy(c.dst, static_cast<const char*>(pin_small_) + off, c.bytes);
    off += (c.bytes + 15) & ~(size_t)15;
  }
  return JSI_OK;
}

int jsi_ctx::h2d(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return JSI_OK;
  if (pin_small_ && bytes <= kPinSmall && src != pin_small_) {
    std::memcpy(pin_small_, src, bytes);
    HCK(hipMemcpyAsync(dst, pin_small_, bytes, hipMemcpyHostToDevice, stream));
    HCK(hipStreamSynchronize(stream));
    return JSI_OK;
  }
  HCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
  HCK(hipStreamSynchronize(stream));
  return JSI_OK;
}

template <class T>
int jsi_ctx::surf_dev(int h, SurfDev<T>* out) {
  std::lock_guard<std::mutex> lock(surfaces->mu);
  auto& tab = surfaces->recs;
  if (h < 0 || h >= (int)tab.size() || !tab[h].live) {
    set_err("invalid surface handle %d", h);
    return JSI_ERR_INVALID;
  }
  SurfaceRec& s = tab[h];
  const T* base;
  if (std::is_same<T, float>::value) {
    base = reinterpret_cast<const T*>(s.d_data);
  } else if (!s.d_data64 && scratch_.base) {
    // no resident FP64 copy and no allocation allowed: promote into the scratch
    // (on this stream, ahead of the kernels that read it)
    double* p64 = nullptr;
    const std::string nm = "s64." + std::to_string(h);
    {
      void* v = nullptr;
      RCK(buf_raw(nm.c_str(), sizeof(double) * (size_t)s.ndata, &v, 0));
      p64 = static_cast<double*>(v);
      HCK(launch_promote<double>(s.d_data, p64, s.ndata, stream));
    }
    base = reinterpret_cast<const T*>(p64);
  } else {
    if (!s.d_data64) {
      // promoted once, synchronously, so every stream sees the finished copy
      if (!pool->reserve(sizeof(double) * (size_t)s.ndata)) {
        set_err("device memory cap exceeded promoting surface %d to FP64", h);
        return JSI_ERR_NOMEM;
      }
      HCK(hipMalloc(&s.d_data64, sizeof(double) * (size_t)s.ndata));
      HCK(launch_promote<double>(s.d_data, s.d_data64, s.ndata, stream));
      HCK(hipStreamSynchronize(stream));
    }
    base = reinterpret_cast<const T*>(s.d_data64);
  }
  const int* iv = s.ivec;
  SurfDev<T> d;
  d.degu = iv[1];
  d.ncpu = iv[2];
  d.nku = iv[3];
  d.degv = iv[4];
  d.ncpv = iv[5];
  d.nkv = iv[6];
  d.rational = (iv[0] & 2) != 0;
  d.zero = (((iv[0] & 1) == 0 && (iv[0] & 12) != 0) || d.degu >= 25 || d.degv >= 25) ? 1 : 0;
  d.U = base;
  d.V = base + d.nku;
  d.CP = d.V + d.nkv;
  d.W = d.rational ? d.CP + 3 * d.ncpu * d.ncpv : nullptr;
  *out = d;
  return JSI_OK;
}

// Evaluates `nunits` regions (device array `reg`) of surface h on an nuv x nuv
// grid.  mode: kEvalFull18 / kEvalPoint / kEvalPointK (see jsi_internal.hpp).
template <class T>
int jsi_ctx::eval_units(int h, const vec4_t<T>* reg, int nunits, int nuv, int mode, int steps_given, T* out,
                         unsigned long long* kmax, const char* tag) {
  (void)tag;
  SurfDev<T> sd;
  RCK(surf_dev<T>(h, &sd));
  // sd.zero (periodic / unsupported degree): all-zero output, as EvaluateOneUV (evaluation.cu:271-272)
  if (mode == kEvalPointK) HCK(launch_fill_u32(reinterpret_cast<unsigned*>(kmax), 0u, (size_t)nunits * 6, stream));
  if (prof) HCK(hipEventRecord(ev[0], stream));
  HCK(launch_points_tile<T>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, stream));
  if (prof) {  // bytes written per launch: 18 values (grid) or xyz (units) per point
    HCK(hipEventRecord(ev[1], stream));
    HCK(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    const std::string key = std::string(mode == kEvalFull18 ? "eval_points18_" : "eval_points_") +
                            (std::is_same<T, float>::value ? "f32" : "f64");
    KernelStat& kst = stats[key];
    kst.launches += 1;
    kst.ms += ms;
    kst.units += (double)nunits * nuv * nuv * (mode == kEvalFull18 ? 18 : 3) * sizeof(T);
  }
  return JSI_OK;
}

// Distance cell centres of both sides' regions, evaluation fused with the
// AABB (one launch when the two surfaces share a kernel variant).
template <class T>
int jsi_ctx::eval_centres(int h1, const vec4_t<T>* reg1, int nunits1, int nuv1, const EcOut<T>& o1, int h2,
                          const vec4_t<T>* reg2, int nunits2, int nuv2, const EcOut<T>& o2) {
  EcJob<T> a, b;
  RCK(surf_dev<T>(h1, &a.sd));
  RCK(surf_dev<T>(h2, &b.sd));
  a.reg = reg1;
  a.nunits = nunits1;
  a.nuv = nuv1;
  a.o = o1;
  b.reg = reg2;
  b.nunits = nunits2;
  b.nuv = nuv2;
  b.o = o2;
  HCK(launch_eval_centres<T>(a, &b, stream));
  return JSI_OK;
}

// ---------------------------------------------------------------------------
// Distance screen: exact top-k under (d2, slot, i, j)
// ---------------------------------------------------------------------------
template <class T>
int jsi_ctx::dist_topk(int nslots, const int* su1, const int* su2, int w1, int nunits1, const vec4_t<T>* C1, int w2,
                        int nunits2, const vec4_t<T>* C2, int k, int cull, Cand** rec_out, long long* n_out,
                        long long* kk_out, Rerank rr, const unsigned long long* exp_absmax,
                        unsigned long long cand_cap, TileBoxes<vec4_t<T>> tbx, Cand* last_out, int screen32) {
  using V = vec4_t<T>;
  const bool rerank = rr.C1 != nullptr;
  const int fma = rerank ? 1 : 0;  // the FP32SCREEN screen may fuse; its margin covers the rounding
  const unsigned long long maxkey_t =
      std::is_same<T, float>::value ? 0x7f800000ull : 0x7ff0000000000000ull;  // +inf bits
  const int tpu1 = ((w1 + kTile - 1) / kTile) * ((w1 + kTile - 1) / kTile);
  const int tpu2 = ((w2 + kTile - 1) / kTile) * ((w2 + kTile - 1) / kTile);
  const V *lo1 = tbx.lo1, *hi1 = tbx.hi1, *lo2 = tbx.lo2, *hi2 = tbx.hi2;
  if (!lo1) {
    V *l1, *h1, *l2, *h2;
    RCK(buf("tb.lo1", (size_t)nunits1 * tpu1, &l1));
    RCK(buf("tb.hi1", (size_t)nunits1 * tpu1, &h1));
    RCK(buf("tb.lo2", (size_t)nunits2 * tpu2, &l2));
    RCK(buf("tb.hi2", (size_t)nunits2 * tpu2, &h2));
    HCK(launch_tilebox<T>(C1, nullptr, nunits1, w1, 0, l1, h1, stream));
    HCK(launch_tilebox<T>(C2, nullptr, nunits2, w2, 0, l2, h2, stream));
    lo1 = l1;
    hi1 = h1;
    lo2 = l2;
    hi2 = h2;
  }
  const unsigned long long ntp = (unsigned long long)nslots * tpu1 * tpu2;
  if (ntp > 0xffffffffull) {
    set_err("too many tile pairs (%llu)", ntp);
    return JSI_ERR_CAPACITY;
  }
  uint64_t* list;
  unsigned* list_count;
  unsigned long long *best, *out_count, *small_count;
  T* thresh;
  Cand *cand, *small;
  SelState* st;
  unsigned* hist;
  RCK(buf("d.list", ntp, &list));
  RCK(buf("d.small", 64, &list_count));
  best = reinterpret_cast<unsigned long long*>(list_count + 2);
  out_count = reinterpret_cast<unsigned long long*>(list_count + 4);
  thresh = reinterpret_cast<T*>(list_count + 6);
  small_count = reinterpret_cast<unsigned long long*>(list_count + 10);
  RCK(buf("d.cand", cand_cap_max_, &cand));
  RCK(buf("d.topk", kSmallCap, &small));
  RCK(buf("d.selstate", 1, &st));
  RCK(buf("d.hist", 4096, &hist));
  thresh_ptr_ = thresh;
  auto widen = [&]() -> int {
    if (rerank) HCK(launch_inflate(reinterpret_cast<float*>(thresh), rr.absmax, stream));
    return JSI_OK;
  };
  // seed: the exact k-th d2 over the cell pairs of up to kSeed promising tile
  // pairs (best side-2 tile per side-1 tile, ranked by upper bound) bounds the
  // global k-th d2 from above
  const long long npos = (long long)nslots * tpu1;
  bool counters_clear = false;
  const int pairs_tp = std::min(w1 * w1, kTileCells) * std::min(w2 * w2, kTileCells);
  // round 1 (one slot of many tiles): up to 256 tile pairs, so the seeds
  // cover most regions that hold top-k pairs; refinement rounds (slots in
  // parent-rank order): about 32 k seed pairs from the first slots
  const unsigned long long want =
      ((unsigned long long)tune_.seed_pairs * (unsigned long long)k + pairs_tp - 1) / (unsigned long long)pairs_tp;
  const int nseed = nslots == 1 ? std::min(tune_.r1_seeds, tpu1) : (int)std::max(4ull, std::min(64ull, want));
  if (npos <= 8192) {
    unsigned long long *bestv = nullptr, *range, *work;
    RCK(buf("d.range", 2, &range));
    RCK(buf_raw("d.seedwork", seed_work_bytes(), reinterpret_cast<void**>(&work), 0));
    const int ordered = (nslots > 1 && tpu1 == 1 && tpu2 == 1) ? 1 : 0;
    if (!ordered) {  // ordered: the seed tile pairs are the first slots' only tile pairs
      RCK(buf("d.best", (size_t)npos, &bestv));
      HCK(launch_dist_seed<T>(nslots, su1, su2, w1, w2, lo1, hi1, lo2, hi2, k, bestv, stream));
    }
    // the seed's last kernel also clears the screen's counters (words 0-5 of
    // d.small: list count, unused, candidate count)
    HCK(launch_seed_thresh<T>(bestv, (int)npos, tpu1, nseed, ordered, list, list_count, range, work, su1, su2, w1,
                              w2, C1, C2, k, fma, thresh, stream, list_count, 6));
    RCK(widen());
    counters_clear = true;
  } else {
    HCK(launch_set_thresh<T>(thresh, cand, 0, 1, stream));
  }
  (void)best;
  // screen, re-rank and select with the candidate count kept on the device;
  // one copy brings back both the candidate count and the selected count
  const unsigned long long cap =
      cand_cap ? std::min<unsigned long long>(std::max<unsigned long long>(cand_cap, 2ull * (unsigned long long)k),
                                              cand_cap_max_)
               : cand_cap_max_;
  const unsigned small_cap = (unsigned)std::max(2048, std::min((int)kSmallCap, 2 * k));
  const unsigned long long maxkey = rerank ? 0x7ff0000000000000ull : maxkey_t;
  unsigned long long n = 0, m = 0;
  for (int attempt = 0;; ++attempt) {
    if (!counters_clear) HCK(launch_fill_u32(list_count, 0u, 1, stream));
    HCK(launch_dist_list<T>(nslots, su1, su2, w1, w2, lo1, hi1, lo2, hi2, thresh, cull, list, list_count, stream,
                            fma));
    if (!counters_clear) HCK(launch_fill_u32(reinterpret_cast<unsigned*>(out_count), 0u, 2, stream));
    counters_clear = false;  // a repeated screen clears them itself
    unsigned long long* pairs_done = nullptr;
    if (prof) {
      RCK(buf("d.pairs", kPairsSlots, &pairs_done));
      HCK(launch_fill_u32(reinterpret_cast<unsigned*>(pairs_done), 0u, 2 * kPairsSlots, stream));
      HCK(hipEventRecord(ev[0], stream));
    }
    const bool units_f = w1 <= kTile && w2 <= kTile;
    const int fgrid = units_f ? grid : std::max(1, (int)((long long)grid * tune_.r1_scale / 100));
    HCK(launch_dist_filter<T>(list, list_count, (unsigned)ntp, su1, su2, w1, w2, C1, C2, thresh, cand, out_count, cap,
                              fgrid, stream, pairs_done, fma, cull, exp_absmax, screen32, lo1, hi1, lo2, hi2,
                              tune_.filter_walk));
    if (prof) HCK(hipEventRecord(ev[1], stream));
    if (rerank) HCK(launch_rerank(cand, out_count, cap, su1, su2, w1 * w1, w2 * w2, rr.C1, rr.C2, stream));
    HCK(select_topk(cand, cap, k, maxkey, st, hist, small, small_count, small_cap, stream, 2, false, out_count));
    unsigned long long cnt[4];  // out_count, thresh, unused, small_count (adjacent in d.small)
    RCK(d2h(cnt, out_count, sizeof(cnt)));
    n = cnt[0];
    m = cnt[3];
    if (prof) {
      float ms = 0.f;
      HCK(hipEventElapsedTime(&ms, ev[0], ev[1]));
      // k_dist_filter_units runs when every unit is a single tile (refinement rounds)
      const bool units = w1 <= kTile && w2 <= kTile;
      std::string key = std::string(units ? "dist_filter_units_" : "dist_filter_tiles_") +
                        (units && screen32 ? "mfma32" : std::is_same<T, float>::value ? "f32" : "f64");
      KernelStat& kst = stats[key];
      kst.launches += 1;
      kst.ms += ms;
      unsigned long long pd[kPairsSlots];
      RCK(d2h(pd, pairs_done, sizeof(pd)));
      double pairs = 0.0;
      for (int q = 0; q < kPairsSlots; ++q) pairs += (double)pd[q];
      kst.units += pairs;
    }
    if (debug_) {
      unsigned lc = 0;
      T tv = T(0);
      RCK(d2h(&lc, list_count, sizeof(lc)));
      RCK(d2h(&tv, thresh, sizeof(tv)));
      fprintf(stderr, "[jsi] topk %s slots=%d w=%d/%d tilepairs=%llu listed=%u thresh=%.9g candidates=%llu%s\n",
              std::is_same<T, float>::value ? (rerank ? "f32s" : "f32") : "f64", nslots, w1, w2, ntp, lc,
              (double)tv, n, n > cap ? " (overflow)" : "");
    }
    if (n <= cap) break;
    if (attempt > 8) {
      set_err("distance screen could not bound its candidate set (%llu > %llu)", n, cap);
      return JSI_ERR_CAPACITY;
    }
    // the first `cap` candidates are real pairs: the k-th smallest of their
    // keys bounds the global k-th key.  When the selection compacted all the
    // records it needs (m <= small_cap), their k-th is read after a rank
    // sort; otherwise the selection's bound on it is used.
    if (m <= small_cap) {
      Cand* sorted;
      RCK(buf("d.sorted", (size_t)k, &sorted));
      HCK(rank_sort(small, small_count, small_cap, (unsigned)k, sorted, stream));
      if (rerank)
        HCK(launch_set_thresh_ru(reinterpret_cast<float*>(thresh), sorted, k - 1, stream));
      else
        HCK(launch_set_thresh<T>(thresh, sorted, k - 1, 0, stream));
    } else if (rerank) {
      HCK(launch_thresh_from_state_ru(reinterpret_cast<float*>(thresh), st, stream));
    } else {
      HCK(launch_thresh_from_state<T>(thresh, st, stream));
    }
    RCK(widen());
  }
  // two narrowing steps from the keys' own range almost always leave at most
  // small_cap records; otherwise continue until they do or the bin holding
  // rank k is a single key (then all its ties are sorted below)
  for (int extra = 0; m > small_cap && extra < 32; ++extra) {
    SelState hs;
    RCK(d2h(&hs, st, sizeof(hs)));
    if (hs.done) break;
    HCK(select_topk(cand, n, k, maxkey, st, hist, small, small_count, small_cap, stream, 2, true));
    RCK(d2h(&m, small_count, sizeof(m)));
  }
  Cand* rec = small;
  if (m > small_cap) {
    // more than kSmallCap candidates tie with the k-th key: sort them all
    const unsigned long long n2 = std::max(next_pow2(std::max(n, 1ull)), (unsigned long long)kSortBlock);
    if (n2 > cand_cap_max_) {
      set_err("distance screen: %llu candidates tie at the k-th key", n);
      return JSI_ERR_CAPACITY;
    }
    HCK(launch_fill_sentinel(cand, n, n2, stream));
    HCK(sort_cands(cand, n2, stream));
    rec = cand;
    if (last_out && n > 0)
      HCK(launch_copy_cand(last_out, cand + (std::min<unsigned long long>(n, (unsigned long long)k) - 1), stream));
  } else {
    Cand* sorted;
    RCK(buf("d.sorted", (size_t)k, &sorted));
    HCK(rank_sort(small, small_count, small_cap, (unsigned)k, sorted, stream, last_out));
    rec = sorted;
  }
  *rec_out = rec;
  *n_out = (long long)n;
  *kk_out = (long long)std::min<unsigned long long>(n, (unsigned long long)k);
  return JSI_OK;
}

static uint64_t fnv_bytes(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}
static constexpr uint64_t kFnvSeed = 1469598103934665603ull;

template <class T>
static double d2_from_bits(uint64_t b) {
  if (std::is_same<T, float>::value) {
    uint32_t u = (uint32_t)b;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
  }
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}

// Screening rounds of a distance query: dist_max_rounds, or with param_tol >
// 0 the first round whose cells span <= param_tol in u and v on both
// surfaces (round 1 cells span |stop - start| / nuv, each later round divides
// by the refinement nuv).
static int dist_rounds(const jsi_params& P, const jsi_region& r1, const jsi_region& r2) {
  if (!(P.param_tol > 0.0)) return P.dist_max_rounds;
  double w = 0.0;
  for (const jsi_region* r : {&r1, &r2}) {
    w = std::max(w, std::fabs((double)r->ustop - (double)r->ustart) / (double)r->nuv);
    w = std::max(w, std::fabs((double)r->vstop - (double)r->vstart) / (double)r->nuv);
  }
  int rounds = 1;
  while (w > P.param_tol && rounds < JSI_MAX_ROUNDS) {
    w = w / (double)P.dist_refine_nuv;
    ++rounds;
  }
  return rounds;
}

// Parameter box (ulo, vlo, uhi, vhi) of a query region.
static double4 region_box(const jsi_region& r) {
  double4 b;
  b.x = std::min((double)r.ustart, (double)r.ustop);
  b.y = std::min((double)r.vstart, (double)r.vstop);
  b.z = std::max((double)r.ustart, (double)r.ustop);
  b.w = std::max((double)r.vstart, (double)r.vstop);
  return b;
}

// FP64 closest-point polish of the first P.polish records of the final
// sorted top-k list (jsi_refine.hip); the smallest |S1 - S2| wins, the lowest
// rank on ties.
template <class T>
int jsi_ctx::polish(int h1, const jsi_region& r1, int h2, const jsi_region& r2, const jsi_params& P, const Cand* rec,
                    long long kk, const int* su1, const int* su2, const vec4_t<T>* reg1, const vec4_t<T>* reg2,
                    int nuv1, int nuv2, jsi_result* res) {
  const int m = (int)std::min<long long>(kk, P.polish);
  SurfDev<double> s1, s2;
  RCK(surf_dev<double>(h1, &s1));
  RCK(surf_dev<double>(h2, &s2));
  PolishRec* d_out;
  RCK(buf("pol.out", (size_t)m, &d_out));
  HCK(launch_polish<T>(s1, s2, rec, m, su1, su2, reg1, reg2, nuv1, nuv2, region_box(r1), region_box(r2), d_out,
                       stream));
  std::vector<PolishRec> h(m);
  RCK(d2h(h.data(), d_out, sizeof(PolishRec) * m));
  int best = 0;
  for (int e = 1; e < m; ++e)
    if (h[e].f < h[best].f) best = e;
  jsi_closest& c = res->closest;
  const PolishRec& b = h[best];
  c.distance = std::sqrt(b.f);
  c.u1 = b.x[0];
  c.v1 = b.x[1];
  c.u2 = b.x[2];
  c.v2 = b.x[3];
  for (int t = 0; t < 3; ++t) {
    c.p1[t] = b.p1[t];
    c.p2[t] = b.p2[t];
  }
  c.candidate = best;
  c.iterations = b.iters;
  c.polished = m;
  c.flags = b.flags;
  return JSI_OK;
}

template <class T>
int jsi_ctx::min_distance(int h1, const jsi_region& r1, int h2, const jsi_region& r2, const jsi_params& P,
                           jsi_result* res, jsi_pair* out, int cap_pairs) {
  using V = vec4_t<T>;
  const double t0 = now_ms();
  {
    SurfDev<float> a, b;
    RCK(surf_dev<float>(h1, &a));
    RCK(surf_dev<float>(h2, &b));
    tune_ = tuning_for(std::max(std::max(a.degu, a.degv), std::max(b.degu, b.degv)), r1.nuv);
  }
  // round-1 regions
  V g1, g2;
  g1.x = T(r1.ustart);
  g1.y = T(r1.vstart);
  g1.z = T(r1.ustop);
  g1.w = T(r1.vstop);
  g2.x = T(r2.ustart);
  g2.y = T(r2.vstart);
  g2.z = T(r2.ustop);
  g2.w = T(r2.vstop);
  int cur = 0;  // ping-pong index of the current round's unit buffers
  auto nm = [](const char* base, int side, int pp) {
    return std::string(base) + (side == 1 ? "1" : "2") + (pp ? "b" : "a");
  };
  V *reg1, *reg2;
  RCK(buf(nm("reg", 1, cur).c_str(), 1, &reg1));
  RCK(buf(nm("reg", 2, cur).c_str(), 1, &reg2));
  int nuv1 = r1.nuv, nuv2 = r2.nuv;
  int nunits1 = 1, nunits2 = 1;  // upper bounds (exact counts of later rounds stay on the device)
  const int* nunits_cur = nullptr;
  V *C1, *C2;
  // per-round results kept on the device and read once at the end: the k-th
  // record and the refinement unit counts (phase timings need a
  // synchronisation each and are taken only when profiling)
  Cand* lastrec;
  int* nunits_all;
  RCK(buf("d.lastrec", JSI_MAX_ROUNDS, &lastrec));
  RCK(buf("d.nunits", 2 * JSI_MAX_ROUNDS, &nunits_all));
  double te = now_ms();
  // double: the evaluation max-reduces the round's largest |centre
  // coordinate| (rounding margins of the screens) into amax[0]; for FP32
  // screening of round 1 it also writes the centres rounded to float
  float4 *f32c1 = nullptr, *f32c2 = nullptr;
  unsigned long long* amax = nullptr;
  int* su1 = nullptr;
  int* su2 = nullptr;
  TileBoxes<V> tb;          // tile boxes of the current round's centres (written by the evaluation)
  TileBoxes<float4> tb32;   // and of their float copies (FP32 screening of round 1)
  auto tiles_of = [](int w) {
    const int tw = (w + kTile - 1) / kTile;
    return tw * tw;
  };
  const bool dbl = std::is_same<T, double>::value;
  const bool s32_first = dbl && P.precision == JSI_PREC_FP32SCREEN;
  {
    RCK(buf(nm("C", 1, cur).c_str(), (size_t)(nuv1 - 1) * (nuv1 - 1), &C1));
    RCK(buf(nm("C", 2, cur).c_str(), (size_t)(nuv2 - 1) * (nuv2 - 1), &C2));
    if (dbl) RCK(buf("d.amax", kAmaxSlots, &amax));
    RCK(buf(nm("su", 1, cur).c_str(), 1, &su1));
    RCK(buf(nm("su", 2, cur).c_str(), 1, &su2));
    // regions, slot 0 -> unit 0 and the cleared max |coordinate| slots in one launch
    HCK(launch_query_init<T>(reg1, g1, reg2, g2, su1, su2, amax, stream));
    EcOut<T> o1, o2;
    o1.C = C1;
    o2.C = C2;
    RCK(buf(nm("tbl", 1, cur).c_str(), (size_t)tiles_of(nuv1 - 1), &o1.lo));
    RCK(buf(nm("tbh", 1, cur).c_str(), (size_t)tiles_of(nuv1 - 1), &o1.hi));
    RCK(buf(nm("tbl", 2, cur).c_str(), (size_t)tiles_of(nuv2 - 1), &o2.lo));
    RCK(buf(nm("tbh", 2, cur).c_str(), (size_t)tiles_of(nuv2 - 1), &o2.hi));
    tb.lo1 = o1.lo;
    tb.hi1 = o1.hi;
    tb.lo2 = o2.lo;
    tb.hi2 = o2.hi;
    if (s32_first) {
      RCK(buf("s32.C1", (size_t)(nuv1 - 1) * (nuv1 - 1), &f32c1));
      RCK(buf("s32.C2", (size_t)(nuv2 - 1) * (nuv2 - 1), &f32c2));
      RCK(buf("s32.tbl1", (size_t)tiles_of(nuv1 - 1), &o1.lo32));
      RCK(buf("s32.tbh1", (size_t)tiles_of(nuv1 - 1), &o1.hi32));
      RCK(buf("s32.tbl2", (size_t)tiles_of(nuv2 - 1), &o2.lo32));
      RCK(buf("s32.tbh2", (size_t)tiles_of(nuv2 - 1), &o2.hi32));
      o1.C32 = f32c1;
      o2.C32 = f32c2;
      tb32.lo1 = o1.lo32;
      tb32.hi1 = o1.hi32;
      tb32.lo2 = o2.lo32;
      tb32.hi2 = o2.hi32;
    }
    o1.absmax = o2.absmax = amax;
    RCK(eval_centres<T>(h1, reg1, 1, nuv1, o1, h2, reg2, 1, nuv2, o2));
  }
  if (prof) {
    HCK(hipStreamSynchronize(stream));
    res->ms_eval += now_ms() - te;
  }
  int nslots = 1;
  const int nref = P.dist_refine_nuv;
  const int max_rounds = dist_rounds(P, r1, r2);
  std::vector<Cand> host_rec;
  Cand* rec = nullptr;
  long long kk = 0;
  for (int round = 1;; ++round) {
    const int w1 = nuv1 - 1, w2 = nuv2 - 1;
    const double ts = now_ms();
    long long n = 0;
    if constexpr (std::is_same<T, double>::value) {
      // FP32SCREEN screens round 1 (6.8e10 pairs at nuv=512) in float.  Refined
      // cells shrink below float resolution of the world coordinates, so the
      // later rounds screen on the FP32 matrix cores relative to an origin per
      // slot (screen32) with FP64 keys; a round 1 whose float candidate set
      // cannot be bounded runs in FP64.
      int s32 = (P.precision == JSI_PREC_FP32SCREEN && round == 1) ? 1 : 0;
      if (s32) {
        Rerank rr;
        rr.C1 = C1;
        rr.C2 = C2;
        rr.absmax = amax;
        const int rc = dist_topk<float>(nslots, su1, su2, w1, nunits1, f32c1, w2, nunits2, f32c2, P.k, P.cull, &rec,
                                        &n, &kk, rr, nullptr, (unsigned long long)P.cand_cap, tb32,
                                        lastrec + (round - 1));
        if (rc == JSI_ERR_CAPACITY)
          s32 = 0;
        else if (rc != JSI_OK)
          return rc;
      }
      if (!s32) {
        RCK(dist_topk<T>(nslots, su1, su2, w1, nunits1, C1, w2, nunits2, C2, P.k, P.cull, &rec, &n, &kk, Rerank(),
                         amax, (unsigned long long)P.cand_cap, tb, lastrec + (round - 1),
                         P.precision == JSI_PREC_FP32SCREEN ? 1 : 0));
      }
    } else {
      RCK(dist_topk<T>(nslots, su1, su2, w1, nunits1, C1, w2, nunits2, C2, P.k, P.cull, &rec, &n, &kk, Rerank(),
                       nullptr, (unsigned long long)P.cand_cap, tb, lastrec + (round - 1)));
    }
    (void)nunits_cur;
    jsi_round_info& ri = res->round[round - 1];
    ri.npairs = kk;
    ri.slots = nslots;
    ri.nuv = nuv1;
    if (prof) res->ms_screen += now_ms() - ts;
    if (P.trace && kk > 0) {
      host_rec.resize(kk);
      RCK(d2h(host_rec.data(), rec, sizeof(Cand) * kk));
      uint64_t hsh = kFnvSeed;
      for (const Cand& c : host_rec) {
        const double dd = d2_from_bits<T>(c.hi);
        const int32_t s = (int32_t)(c.lo >> 40), i = (int32_t)((c.lo >> 20) & 0xfffff), j = (int32_t)(c.lo & 0xfffff);
        hsh = fnv_bytes(hsh, &dd, 8);
        hsh = fnv_bytes(hsh, &s, 4);
        hsh = fnv_bytes(hsh, &i, 4);
        hsh = fnv_bytes(hsh, &j, 4);
      }
      ri.hash = hsh;
    }
    res->rounds = round;
    res->npairs = kk;
    if (kk == 0 || round == max_rounds || round >= JSI_MAX_ROUNDS) break;
    // refinement units for the next round
    te = now_ms();
    const int nxt = cur ^ 1;
    int *fp1, *fp2, *nsu1, *nsu2;
    int* const nunits_dev = nunits_all + 2 * (round - 1);
    V *nreg1, *nreg2;
    RCK(buf_raw("fp1", sizeof(int) * (size_t)nslots * w1 * w1, reinterpret_cast<void**>(&fp1), INT_MAX));
    RCK(buf_raw("fp2", sizeof(int) * (size_t)nslots * w2 * w2, reinterpret_cast<void**>(&fp2), INT_MAX));
    RCK(buf(nm("su", 1, nxt).c_str(), (size_t)kk, &nsu1));
    RCK(buf(nm("su", 2, nxt).c_str(), (size_t)kk, &nsu2));
    RCK(buf(nm("reg", 1, nxt).c_str(), (size_t)kk, &nreg1));
    RCK(buf(nm("reg", 2, nxt).c_str(), (size_t)kk, &nreg2));
    unsigned char* dwork;
    RCK(buf_raw("d.dedupwork", dedup_work_bytes(), reinterpret_cast<void**>(&dwork), 0));
    HCK(launch_dist_dedup<T>(rec, (int)kk, w1 * w1, w2 * w2, su1, su2, reg1, reg2, nuv1, nuv2, fp1, fp2, nsu1, nsu2,
                             nreg1, nreg2, nunits_dev, dwork, stream, amax, dbl ? kAmaxSlots : 0));
    // at most kk units per side; the exact counts stay on the device and the
    // evaluation skips the tiles beyond them
    const int nu_max = (int)kk;
    V *nC1, *nC2;
    const int wr = nref - 1;
    RCK(buf(nm("C", 1, nxt).c_str(), (size_t)nu_max * wr * wr, &nC1));
    RCK(buf(nm("C", 2, nxt).c_str(), (size_t)nu_max * wr * wr, &nC2));
    EcOut<T> o1, o2;
    o1.C = nC1;
    o2.C = nC2;
    RCK(buf(nm("tbl", 1, nxt).c_str(), (size_t)nu_max, &o1.lo));  // one tile per refinement unit
    RCK(buf(nm("tbh", 1, nxt).c_str(), (size_t)nu_max, &o1.hi));
    RCK(buf(nm("tbl", 2, nxt).c_str(), (size_t)nu_max, &o2.lo));
    RCK(buf(nm("tbh", 2, nxt).c_str(), (size_t)nu_max, &o2.hi));
    o1.absmax = o2.absmax = amax;
    o1.nunits_dev = nunits_dev;
    o2.nunits_dev = nunits_dev + 1;
    RCK(eval_centres<T>(h1, nreg1, nu_max, nref, o1, h2, nreg2, nu_max, nref, o2));
    tb.lo1 = o1.lo;
    tb.hi1 = o1.hi;
    tb.lo2 = o2.lo;
    tb.hi2 = o2.hi;
    if (prof) {
      HCK(hipStreamSynchronize(stream));
      res->ms_eval += now_ms() - te;
    }
    cur = nxt;
    su1 = nsu1;
    su2 = nsu2;
    reg1 = nreg1;
    reg2 = nreg2;
    C1 = nC1;
    C2 = nC2;
    nunits1 = nu_max;
    nunits2 = nu_max;
    nslots = (int)kk;
    nuv1 = nuv2 = nref;
  }
  {
    const int nr = res->rounds;
    std::vector<Cand> lr(nr);
    std::vector<int> nu(2 * nr);
    RCK(d2h(lr.data(), lastrec, sizeof(Cand) * nr));
    RCK(d2h(nu.data(), nunits_all, sizeof(int) * 2 * nr));
    for (int r = 0; r < nr; ++r) {
      jsi_round_info& ri = res->round[r];
      ri.kth_d2 = ri.npairs > 0 ? d2_from_bits<T>(lr[r].hi) : 0.0;
      if (r + 1 < nr) {
        ri.units1 = nu[2 * r];
        ri.units2 = nu[2 * r + 1];
      }
    }
    if (kk > 0) {
      const T kd = T(d2_from_bits<T>(lr[nr - 1].hi));
      res->distance = (double)std::sqrt(kd);
    } else {
      res->distance = 0.0;
    }
  }
  if (P.polish > 0 && kk > 0) RCK(polish<T>(h1, r1, h2, r2, P, rec, kk, su1, su2, reg1, reg2, nuv1, nuv2, res));
  if (out && cap_pairs > 0 && kk > 0) {
    const long long m = std::min<long long>(kk, cap_pairs);
    host_rec.resize(m);
    RCK(d2h(host_rec.data(), rec, sizeof(Cand) * m));
    for (long long e = 0; e < m; ++e) {
      const Cand& c = host_rec[e];
      out[e].d2 = d2_from_bits<T>(c.hi);
      out[e].slot = (int32_t)(c.lo >> 40);
      out[e].i = (int32_t)((c.lo >> 20) & 0xfffff);
      out[e].j = (int32_t)(c.lo & 0xfffff);
      out[e].pad_ = 0;
    }
  }
  res->ms_total = now_ms() - t0;
  return JSI_OK;
}

// ---------------------------------------------------------------------------
// Intersection
// ---------------------------------------------------------------------------
template <class T>
int jsi_ctx::isect_units(int h, const vec4_t<T>* reg, int nunits, int nuv, int use_k, T init_u, T init_v,
                          vec4_t<T>* C, vec4_t<T>* R, const char* tag) {
  T* pts;
  unsigned long long* kmax;
  RCK(buf("ipts", (size_t)3 * nunits * nuv * nuv, &pts));
  RCK(buf((std::string("kmax.") + tag).c_str(), (size_t)3 * nunits, &kmax));
  RCK(eval_units<T>(h, reg, nunits, nuv, use_k ? kEvalPointK : kEvalPoint, 0, pts, kmax, tag));
  HCK(launch_aabb<T>(pts, nunits, nuv, 1, use_k, kmax, reg, init_u, init_v, C, R, stream));
  return JSI_OK;
}

// compute_K of the two initial regions and their patch scale, from the
// maxima the round-1 evaluation left in kmax.i1 / kmax.i2 (same arithmetic as
// k_aabb, so the float values are the ones the padding used).
template <class T>
int jsi_ctx::initial_k(int nuv1, int nuv2, const jsi_region& r1, const jsi_region& r2, const T* iu, const T* iv,
                       jsi_result* res, Copy extra) {
  // both sides' maxima (and the caller's extra read-back) with one wait
  unsigned long long* kmax1;
  unsigned long long* kmax2;
  RCK(buf("kmax.i1", 3, &kmax1));
  RCK(buf("kmax.i2", 3, &kmax2));
  unsigned long long mm[2][3];
  RCK(d2h_many({Copy{mm[0], kmax1, sizeof(mm[0])}, Copy{mm[1], kmax2, sizeof(mm[1])}, extra}));
  for (int side = 1; side <= 2; ++side) {
    const unsigned long long* m = mm[side - 1];
    T v[3];
    for (int c = 0; c < 3; ++c) {
      if (std::is_same<T, float>::value) {
        const uint32_t b = (uint32_t)m[c];
        std::memcpy(&v[c], &b, sizeof(float));
      } else {
        std::memcpy(&v[c], &m[c], sizeof(double));
      }
    }
    const int nuv = side == 1 ? nuv1 : nuv2;
    const jsi_region& r = side == 1 ? r1 : r2;
    const int sd = nuv > 1 ? nuv - 1 : 1;
    const T s2 = T(sd * sd);
    const T raw = (v[0] / s2 + T(2) * v[1] / s2 + v[2] / s2) / T(8);
    const T us = std::fabs(T(r.ustop) - T(r.ustart)) / iu[side];
    const T vs = std::fabs(T(r.vstop) - T(r.vstart)) / iv[side];
    res->isect_k_raw[side - 1] = (double)raw;
    res->isect_k_scale[side - 1] = (double)(us < vs ? vs : us);
  }
  return JSI_OK;
}

// FP32SCREEN screen of one intersection round: float records covering each
// cell's FP64 acceptance interval (and, given V*, VERSION5's float one), the
// float tile cull and overlap test, rows sorted by j, then the FP64 test of
// every candidate and the order-keeping compaction: the result is the FP64
// screen's list.  The candidates stay in *cand (screen_out).
int jsi_ctx::isect_screen32(int round, int nslots, const int* su1, const int* su2, int nunits1, int w1, int nunits2,
                            int w2, const double4* C1, const double4* R1, const double4* C2, const double4* R2,
                            const float4* V1c, const float4* V1r, const float4* V2c, const float4* V2r, int cull,
                            const char* pairs_name, IPair** pairs_out, unsigned long long* npairs_out,
                            IPair** cand_out, unsigned long long* ncand_out) {
  const int tpu1 = ((w1 + kTile - 1) / kTile) * ((w1 + kTile - 1) / kTile);
  const int tpu2 = ((w2 + kTile - 1) / kTile) * ((w2 + kTile - 1) / kTile);
  const size_t n1 = (size_t)nunits1 * w1 * w1, n2 = (size_t)nunits2 * w2 * w2;
  float4 *X1, *Y1, *X2, *Y2, *lo1, *hi1, *lo2, *hi2;
  RCK(buf("i32.X1", n1, &X1));
  RCK(buf("i32.Y1", n1, &Y1));
  RCK(buf("i32.X2", n2, &X2));
  RCK(buf("i32.Y2", n2, &Y2));
  HCK(launch_isect_boxes32(C1, R1, V1c, V1r, n1, X1, Y1, stream));
  HCK(launch_isect_boxes32(C2, R2, V2c, V2r, n2, X2, Y2, stream));
  RCK(buf("tb32.lo1", (size_t)nunits1 * tpu1, &lo1));
  RCK(buf("tb32.hi1", (size_t)nunits1 * tpu1, &hi1));
  RCK(buf("tb32.lo2", (size_t)nunits2 * tpu2, &lo2));
  RCK(buf("tb32.hi2", (size_t)nunits2 * tpu2, &hi2));
  HCK(launch_tilebox<float>(X1, Y1, nunits1, w1, 1, lo1, hi1, stream));
  HCK(launch_tilebox<float>(X2, Y2, nunits2, w2, 1, lo2, hi2, stream));
  const unsigned long long ntp = (unsigned long long)nslots * tpu1 * tpu2;
  if (ntp > (1ull << 40)) {
    set_err("intersection round %d: %d slots exceed the supported range", round, nslots);
    return JSI_ERR_CAPACITY;
  }
  uint64_t* list;
  unsigned long long* small;
  RCK(buf("i.list", ntp, &list));
  RCK(buf("i.small", 8, &small));
  HCK(launch_fill_u32(reinterpret_cast<unsigned*>(small), 0u, 2, stream));
  HCK(launch_isect_list<float>(nslots, su1, su2, w1, w2, lo1, hi1, lo2, hi2, cull, list, small, ntp, stream));
  const unsigned long long nrows = (unsigned long long)nslots * w1 * w1;
  unsigned *rowcount, *rowcursor;
  unsigned long long* rowoff;
  void* scantmp;
  RCK(buf("i.rowcount", nrows, &rowcount));
  RCK(buf("i.rowcursor", nrows, &rowcursor));
  RCK(buf("i.rowoff", nrows + 1, &rowoff));
  const size_t stb = scan_tmp_bytes(nrows);
  RCK(buf_raw("i.scantmp", stb, &scantmp, 0));
  HCK(launch_fill_u32(rowcount, 0u, nrows, stream));
  HCK(launch_isect_pass<float>(0, list, small, su1, su2, w1, w2, X1, Y1, X2, Y2, rowcount, rowoff, rowcursor, nullptr,
                               grid, stream));
  HCK(scan_u32_to_u64(rowcount, rowoff, nrows, scantmp, stb, stream));
  unsigned long long ncand = 0;
  RCK(d2h(&ncand, rowoff + nrows, sizeof(ncand)));
  IPair* cand;
  RCK(buf("i32.cand", std::max(ncand, 1ull), &cand));
  unsigned long long npairs = 0;
  if (ncand > 0) {
    HCK(launch_fill_u32(rowcursor, 0u, nrows, stream));
    HCK(launch_isect_pass<float>(1, list, small, su1, su2, w1, w2, X1, Y1, X2, Y2, rowcount, rowoff, rowcursor, cand,
                                 grid, stream));
    if (ncand > 16ull * nrows) {
      const size_t tb = segmented_sort_tmp_bytes(ncand, nrows);
      void* st;
      RCK(buf_raw("i.segsort", tb, &st, 0));
      HCK(segmented_sort_pairs(cand, ncand, rowoff, nrows, st, tb, stream));
    } else {
      HCK(launch_row_sort(rowoff, nrows, cand, stream));
    }
    unsigned* flag;
    unsigned long long* idx;
    RCK(buf("i32.flag", ncand, &flag));
    RCK(buf("i32.idx", ncand + 1, &idx));
    HCK(launch_isect_retest(cand, ncand, su1, su2, w1 * w1, w2 * w2, C1, R1, C2, R2, flag, stream));
    const size_t stb2 = scan_tmp_bytes(ncand);
    RCK(buf_raw("i.scantmp2", stb2, &scantmp, 0));
    HCK(scan_u32_to_u64(flag, idx, ncand, scantmp, stb2, stream));
    RCK(d2h(&npairs, idx + ncand, sizeof(npairs)));
    RCK(buf(pairs_name, std::max(npairs, 1ull), pairs_out));
    HCK(launch_isect_scatter(cand, flag, idx, ncand, *pairs_out, stream));
  } else {
    RCK(buf(pairs_name, 1, pairs_out));
  }
  if (prof) {
    KernelStat& kst = stats["isect_fp32_candidates"];
    kst.launches += 1;
    kst.units += (double)ncand;
  }
  *npairs_out = npairs;
  *cand_out = cand;
  *ncand_out = ncand;
  return JSI_OK;
}

template <class T>
int jsi_ctx::intersect(int h1, const jsi_region& r1, int h2, const jsi_region& r2, const jsi_params& P,
                        jsi_result* res, jsi_pair* out, int cap_pairs, CurveReq* creq) {
  using V = vec4_t<T>;
  const double t0 = now_ms();
  auto smax_h = [](T a, T b) { return (a < b) ? b : a; };
  const T tiny = T(1.0e-12f);
  T iu[3] = {T(1), T(1), T(1)}, iv[3] = {T(1), T(1), T(1)};
  iu[1] = smax_h(std::fabs(T(r1.ustop) - T(r1.ustart)), tiny);
  iv[1] = smax_h(std::fabs(T(r1.vstop) - T(r1.vstart)), tiny);
  iu[2] = smax_h(std::fabs(T(r2.ustop) - T(r2.ustart)), tiny);
  iv[2] = smax_h(std::fabs(T(r2.vstop) - T(r2.vstart)), tiny);
  const int use_k = P.use_compute_k != 0;
  auto nm = [](const char* base, int side, int pp) {
    return std::string("i") + base + (side == 1 ? "1" : "2") + (pp ? "b" : "a");
  };
  int cur = 0;
  V g1, g2;
  g1.x = T(r1.ustart);
  g1.y = T(r1.vstart);
  g1.z = T(r1.ustop);
  g1.w = T(r1.vstop);
  g2.x = T(r2.ustart);
  g2.y = T(r2.vstart);
  g2.z = T(r2.ustop);
  g2.w = T(r2.vstop);
  V *reg1, *reg2, *C1, *R1, *C2, *R2;
  int *su1, *su2;
  RCK(buf(nm("reg", 1, cur).c_str(), 1, &reg1));
  RCK(buf(nm("reg", 2, cur).c_str(), 1, &reg2));
  RCK(h2d(reg1, &g1, sizeof(V)));
  RCK(h2d(reg2, &g2, sizeof(V)));
  int nuv1 = r1.nuv, nuv2 = r2.nuv;
  double te = now_ms();
  RCK(buf(nm("C", 1, cur).c_str(), (size_t)(nuv1 - 1) * (nuv1 - 1), &C1));
  RCK(buf(nm("R", 1, cur).c_str(), (size_t)(nuv1 - 1) * (nuv1 - 1), &R1));
  RCK(buf(nm("C", 2, cur).c_str(), (size_t)(nuv2 - 1) * (nuv2 - 1), &C2));
  RCK(buf(nm("R", 2, cur).c_str(), (size_t)(nuv2 - 1) * (nuv2 - 1), &R2));
  RCK(isect_units<T>(h1, reg1, 1, nuv1, use_k, iu[1], iv[1], C1, R1, "i1"));
  RCK(isect_units<T>(h2, reg2, 1, nuv2, use_k, iu[2], iv[2], C2, R2, "i2"));
  RCK(buf(nm("su", 1, cur).c_str(), 1, &su1));
  RCK(buf(nm("su", 2, cur).c_str(), 1, &su2));
  HCK(launch_fill_i32(su1, 0, 1, stream));
  HCK(launch_fill_i32(su2, 0, 1, stream));
  // the phase timings cost a synchronisation each: taken when profiling or
  // when the legacy API prints them (CAD_TIMING_BREAKDOWN)
  const bool timing = prof || std::getenv("CAD_TIMING_BREAKDOWN") != nullptr;
  if (timing) {
    HCK(hipStreamSynchronize(stream));
    res->ms_eval += now_ms() - te;
  }
  // FP32SCREEN: the float screen's round-1 records also cover VERSION5's own
  // float boxes of the same grid (EXACT32 evaluation of the two regions)
  const bool s32 = std::is_same<T, double>::value && P.precision == JSI_PREC_FP32SCREEN;
  float4 *V1c = nullptr, *V1r = nullptr, *V2c = nullptr, *V2r = nullptr;
  if (s32) {
    const float tinyf = 1.0e-12f;
    auto ext = [&](float a, float b) { return std::max(std::fabs(b - a), tinyf); };
    const float4 f1 = make_float4(r1.ustart, r1.vstart, r1.ustop, r1.vstop);
    const float4 f2 = make_float4(r2.ustart, r2.vstart, r2.ustop, r2.vstop);
    float4 *regf1, *regf2;
    RCK(buf("i32.reg1", 1, &regf1));
    RCK(buf("i32.reg2", 1, &regf2));
    RCK(h2d(regf1, &f1, sizeof(float4)));
    RCK(h2d(regf2, &f2, sizeof(float4)));
    RCK(buf("i32.V1c", (size_t)(nuv1 - 1) * (nuv1 - 1), &V1c));
    RCK(buf("i32.V1r", (size_t)(nuv1 - 1) * (nuv1 - 1), &V1r));
    RCK(buf("i32.V2c", (size_t)(nuv2 - 1) * (nuv2 - 1), &V2c));
    RCK(buf("i32.V2r", (size_t)(nuv2 - 1) * (nuv2 - 1), &V2r));
    RCK(isect_units<float>(h1, regf1, 1, nuv1, use_k, ext(r1.ustart, r1.ustop), ext(r1.vstart, r1.vstop), V1c, V1r,
                           "x1"));
    RCK(isect_units<float>(h2, regf2, 1, nuv2, use_k, ext(r2.ustart, r2.ustop), ext(r2.vstart, r2.vstop), V2c, V2r,
                           "x2"));
  }
  int nunits1 = 1, nunits2 = 1;
  long long nslots = 1;
  int g_nuv = P.isect_refine_nuv;
  IPair* pairs = nullptr;
  unsigned long long npairs = 0;
  IPair* cand_out = nullptr;  // screen_out: the last round's float candidates
  unsigned long long ncand_out = 0;
  std::vector<IPair> hp;
  // units of nuv <= kIcMaxNuv: evaluation, cells and tile boxes in one pass
  // (JSI_ISECT_CELLS=0 keeps the grid-point kernel, the AABB and tile-box passes)
  const char* icenv = std::getenv("JSI_ISECT_CELLS");
  const bool fuse_cells = !(icenv && std::strcmp(icenv, "0") == 0);
  // rounds of at least this many slots list the slots that pass the unit-box
  // test before the screen (JSI_SLOT_LIST_MIN overrides the default)
  const char* smenv = std::getenv("JSI_SLOT_MASKS");  // 0: the emit pass repeats the tests
  const bool slot_masks = !(smenv && std::atoi(smenv) == 0);
  const char* slenv = std::getenv("JSI_SLOT_LIST_MIN");
  const long long slot_list_min = slenv ? std::atoll(slenv) : kSlotListMin;
  V *tlo1 = nullptr, *thi1 = nullptr, *tlo2 = nullptr, *thi2 = nullptr;  // tile boxes of the current units
  bool units_ready = false;  // the screen also wrote the next round's units
  unsigned long long ready_nu1 = 0, ready_nu2 = 0;
  bool f32_ready = false;  // FP32SCREEN: the evaluation wrote the units' float records and boxes
  for (int round = 1;; ++round) {
    const double ts = now_ms();
    const int w1 = nuv1 - 1, w2 = nuv2 - 1;
    const int tpu1 = ((w1 + kTile - 1) / kTile) * ((w1 + kTile - 1) / kTile);
    const int tpu2 = ((w2 + kTile - 1) / kTile) * ((w2 + kTile - 1) / kTile);
    // units of at most 64 cells are screened per slot (FP32SCREEN too, with
    // its float records; the screen_out hook keeps the candidate list form)
    const bool small = fuse_cells && w1 <= kIcMaxNuv && w2 <= kIcMaxNuv;
    SlotF32 f32;
    if (s32 && !(small && !P.screen_out)) {
      if constexpr (std::is_same<T, double>::value) {
        RCK(isect_screen32(round, (int)nslots, su1, su2, nunits1, w1, nunits2, w2, C1, R1, C2, R2,
                           round == 1 ? V1c : nullptr, round == 1 ? V1r : nullptr, round == 1 ? V2c : nullptr,
                           round == 1 ? V2r : nullptr, P.cull, nm("pairs", 1, cur).c_str(), &pairs, &npairs,
                           &cand_out, &ncand_out));
        if (round == 1 && use_k) RCK(initial_k<T>(nuv1, nuv2, r1, r2, iu, iv, res));
      }
    } else {
      if constexpr (std::is_same<T, double>::value) {
        if (s32) {  // float records of both sides (round 1: covering VERSION5's float boxes too) and their unit boxes
          const size_t n1 = (size_t)nunits1 * w1 * w1, n2 = (size_t)nunits2 * w2 * w2;
          float4 *X1, *Y1, *X2, *Y2, *fl1, *fh1, *fl2, *fh2;
          RCK(buf("i32.X1", n1, &X1));
          RCK(buf("i32.Y1", n1, &Y1));
          RCK(buf("i32.X2", n2, &X2));
          RCK(buf("i32.Y2", n2, &Y2));
          RCK(buf("tb32.lo1", (size_t)nunits1, &fl1));
          RCK(buf("tb32.hi1", (size_t)nunits1, &fh1));
          RCK(buf("tb32.lo2", (size_t)nunits2, &fl2));
          RCK(buf("tb32.hi2", (size_t)nunits2, &fh2));
          if (!f32_ready) {  // not written by the small-unit evaluation
            HCK(launch_isect_boxes32(C1, R1, round == 1 ? V1c : nullptr, round == 1 ? V1r : nullptr, n1, X1, Y1,
                                     stream));
            HCK(launch_isect_boxes32(C2, R2, round == 1 ? V2c : nullptr, round == 1 ? V2r : nullptr, n2, X2, Y2,
                                     stream));
            HCK(launch_tilebox<float>(X1, Y1, nunits1, w1, 1, fl1, fh1, stream));
            HCK(launch_tilebox<float>(X2, Y2, nunits2, w2, 1, fl2, fh2, stream));
          }
          f32.X1 = X1;
          f32.Y1 = Y1;
          f32.X2 = X2;
          f32.Y2 = Y2;
          f32.lo1 = fl1;
          f32.hi1 = fh1;
          f32.lo2 = fl2;
          f32.hi2 = fh2;
        }
      }
      V *lo1 = tlo1, *hi1 = thi1, *lo2 = tlo2, *hi2 = thi2;
      if (!lo1 && !s32) {  // tile boxes not written by the evaluation
        RCK(buf("tb.lo1", (size_t)nunits1 * tpu1, &lo1));
        RCK(buf("tb.hi1", (size_t)nunits1 * tpu1, &hi1));
        RCK(buf("tb.lo2", (size_t)nunits2 * tpu2, &lo2));
        RCK(buf("tb.hi2", (size_t)nunits2 * tpu2, &hi2));
        HCK(launch_tilebox<T>(C1, R1, nunits1, w1, 1, lo1, hi1, stream));
        HCK(launch_tilebox<T>(C2, R2, nunits2, w2, 1, lo2, hi2, stream));
      }
      const unsigned long long ntp = (unsigned long long)nslots * tpu1 * tpu2;
      if (nslots > INT_MAX || ntp > (1ull << 40)) {
        set_err("intersection round %d: %lld slots exceed the supported range", round, nslots);
        return JSI_ERR_CAPACITY;
      }
      if (small) {
        // units of at most 64 cells: per-slot counts of pairs and new units,
        // their scans, then the ordered pairs and (when a round follows) the
        // next round's units in one pass
        const SlotF32* fp = s32 ? &f32 : nullptr;
        // many slots: cull them one per lane first and list those that pass
        size_t ns = (size_t)nslots;
        int* slist = nullptr;
        if (P.cull && nslots >= slot_list_min) {
          const size_t nw = (ns + 63) / 64;
          unsigned* wcount;
          unsigned long long* woff;
          void* wtmp;
          RCK(buf("i.wcount", nw, &wcount));
          RCK(buf("i.woff", nw + 1, &woff));
          RCK(buf("i.slist", ns, &slist));
          const size_t wtb = scan_tmp_bytes(nw);
          RCK(buf_raw("i.scantmp", wtb, &wtmp, 0));
          HCK(launch_isect_slot_cull<T>(nslots, w1, w2, su1, su2, lo1, hi1, lo2, hi2, wcount, nullptr, nullptr, fp,
                                        stream));
          HCK(scan_u32_to_u64(wcount, woff, nw, wtmp, wtb, stream));
          HCK(launch_isect_slot_cull<T>(nslots, w1, w2, su1, su2, lo1, hi1, lo2, hi2, nullptr, woff, slist, fp,
                                        stream));
          unsigned long long nl = 0;
          RCK(d2h(&nl, woff + nw, sizeof(nl)));
          ns = (size_t)nl;
        }
        unsigned* cnt;
        unsigned long long *off, *tot;
        void* scantmp;
        RCK(buf("i.scount", 3 * std::max(ns, (size_t)1), &cnt));
        RCK(buf("i.soff", 3 * (ns + 1), &off));
        RCK(buf("i.stot", 3, &tot));
        const size_t stb = scan_tmp_bytes(ns, 3);
        RCK(buf_raw("i.scantmp", stb, &scantmp, 0));
        unsigned *cp = cnt, *cr = cnt + ns, *cc = cnt + 2 * ns;
        unsigned long long *op = off, *orr = off + ns + 1, *oc = off + 2 * (ns + 1);
        // the rows' hit masks, kept from the count pass for the emit pass
        // when they fit kSlotMaskBytes (a dense round's are recomputed)
        unsigned long long* masks = nullptr;
        const size_t mrows = ns * (size_t)w1 * (size_t)w1;
        if (slot_masks && mrows * sizeof(unsigned long long) <= kSlotMaskBytes)
          RCK(buf("i.masks", std::max(mrows, (size_t)1), &masks));
        HCK(launch_isect_slot_count<T>(nslots, w1, w2, su1, su2, C1, R1, C2, R2, lo1, hi1, lo2, hi2, P.cull,
                                       (long long)ns, slist, cp, cr, cc, fp, stream, masks));
        HCK(scan3_u32_to_u64(cp, ns, op, ns + 1, ns, scantmp, stb, stream, tot));  // pairs, rows, columns
        unsigned long long t3[3];
        RCK(d2h(t3, tot, sizeof(t3)));
        npairs = t3[0];
        if (round == 1 && use_k) RCK(initial_k<T>(nuv1, nuv2, r1, r2, iu, iv, res));
        RCK(buf(nm("pairs", 1, cur).c_str(), std::max(npairs, 1ull), &pairs));
        const bool follows = round < P.isect_max_rounds && round < JSI_MAX_ROUNDS && npairs > 0 &&
                             npairs <= (unsigned long long)INT_MAX;
        int *nsu1 = nullptr, *nsu2 = nullptr;
        V *nreg1 = nullptr, *nreg2 = nullptr;
        if (follows) {
          const int nxt = cur ^ 1;
          RCK(buf(nm("su", 1, nxt).c_str(), npairs, &nsu1));
          RCK(buf(nm("su", 2, nxt).c_str(), npairs, &nsu2));
          RCK(buf(nm("reg", 1, nxt).c_str(), t3[1], &nreg1));
          RCK(buf(nm("reg", 2, nxt).c_str(), t3[2], &nreg2));
          units_ready = true;
          ready_nu1 = t3[1];
          ready_nu2 = t3[2];
        }
        if (npairs > 0)
          HCK(launch_isect_slot_emit<T>(nslots, w1, w2, su1, su2, C1, R1, C2, R2, lo1, hi1, lo2, hi2, P.cull,
                                        (long long)ns, slist, op, orr, oc, pairs, reg1, reg2, nuv1, nuv2, nsu1, nsu2,
                                        nreg1, nreg2, fp, stream, masks));
      } else {
        uint64_t* list;
        unsigned long long* small;
        RCK(buf("i.list", ntp, &list));
        RCK(buf("i.small", 8, &small));
        HCK(launch_fill_u32(reinterpret_cast<unsigned*>(small), 0u, 2, stream));
        HCK(launch_isect_list<T>((int)nslots, su1, su2, w1, w2, lo1, hi1, lo2, hi2, P.cull, list, small, ntp, stream));
        const unsigned long long nrows = (unsigned long long)nslots * w1 * w1;
        unsigned *rowcount, *rowcursor;
        unsigned long long* rowoff;
        void* scantmp;
        RCK(buf("i.rowcount", nrows, &rowcount));
        RCK(buf("i.rowcursor", nrows, &rowcursor));
        RCK(buf("i.rowoff", nrows + 1, &rowoff));
        const size_t stb = scan_tmp_bytes(nrows);
        RCK(buf_raw("i.scantmp", stb, &scantmp, 0));
        HCK(launch_fill_u32(rowcount, 0u, nrows, stream));
        // the hit masks of the first listed tile pairs (four words per
        // side-1 cell, 8 KB per tile pair, up to kSlotMaskBytes), kept for
        // the write pass
        unsigned long long* tmasks = nullptr;
        const unsigned long long mask_tp =
            slot_masks ? std::min<unsigned long long>(ntp, kSlotMaskBytes / (kTileCells * 4 * sizeof(unsigned long long)))
                       : 0ull;
        if (mask_tp) RCK(buf("i.tmasks", (size_t)mask_tp * kTileCells * 4, &tmasks));
        HCK(launch_isect_pass<T>(0, list, small, su1, su2, w1, w2, C1, R1, C2, R2, rowcount, rowoff, rowcursor, nullptr,
                                 grid, stream, tmasks, mask_tp));
        HCK(scan_u32_to_u64(rowcount, rowoff, nrows, scantmp, stb, stream));
        if (round == 1 && use_k)  // the pair count comes back with the first round's maxima
          RCK(initial_k<T>(nuv1, nuv2, r1, r2, iu, iv, res, Copy{&npairs, rowoff + nrows, sizeof(npairs)}));
        else
          RCK(d2h(&npairs, rowoff + nrows, sizeof(npairs)));
        RCK(buf(nm("pairs", 1, cur).c_str(), std::max(npairs, 1ull), &pairs));
        if (npairs > 0) {
          HCK(launch_fill_u32(rowcursor, 0u, nrows, stream));
          HCK(launch_isect_pass<T>(1, list, small, su1, su2, w1, w2, C1, R1, C2, R2, rowcount, rowoff, rowcursor, pairs,
                                   grid, stream, tmasks, mask_tp));
          if (npairs > 16ull * nrows) {  // long rows: segmented radix sort
            const size_t tb = segmented_sort_tmp_bytes(npairs, nrows);
            void* st;
            RCK(buf_raw("i.segsort", tb, &st, 0));
            HCK(segmented_sort_pairs(pairs, npairs, rowoff, nrows, st, tb, stream));
          } else {
            HCK(launch_row_sort(rowoff, nrows, pairs, stream));
          }
        }
      }
    }
    if (timing) {
      HCK(hipStreamSynchronize(stream));
      res->ms_screen += now_ms() - ts;
    }
    jsi_round_info& ri = res->round[round - 1];
    ri.npairs = (int64_t)npairs;
    ri.slots = nslots;
    ri.nuv = nuv1;
    if (P.trace) {
      hp.resize(npairs);
      RCK(d2h(hp.data(), pairs, sizeof(IPair) * npairs));
      uint64_t hsh = kFnvSeed;
      for (const IPair& p : hp) {
        hsh = fnv_bytes(hsh, &p.slot, 4);
        hsh = fnv_bytes(hsh, &p.i, 4);
        hsh = fnv_bytes(hsh, &p.j, 4);
      }
      ri.hash = hsh;
    }
    res->rounds = round;
    res->npairs = (int64_t)npairs;
    if (round == P.isect_max_rounds || round >= JSI_MAX_ROUNDS) break;
    if (npairs == 0) break;
    if ((long long)npairs > P.isect_halve_above && g_nuv > P.isect_min_nuv) g_nuv /= 2;
    if (npairs > (unsigned long long)INT_MAX) {
      set_err("intersection round %d: %llu pairs exceed the supported range", round, npairs);
      return JSI_ERR_CAPACITY;
    }
    te = now_ms();
    const int nxt = cur ^ 1;
    unsigned long long nu1 = 0, nu2 = 0;
    int *nsu1, *nsu2;
    V *nreg1, *nreg2;
    if (units_ready) {  // written by the screen
      nu1 = ready_nu1;
      nu2 = ready_nu2;
      RCK(buf(nm("su", 1, nxt).c_str(), npairs, &nsu1));
      RCK(buf(nm("su", 2, nxt).c_str(), npairs, &nsu2));
      RCK(buf(nm("reg", 1, nxt).c_str(), nu1, &nreg1));
      RCK(buf(nm("reg", 2, nxt).c_str(), nu2, &nreg2));
      units_ready = false;
    } else {
      unsigned *flag1, *flag2;
      unsigned long long *idx1, *idx2;
      int* fp2;
      RCK(buf("i.flag1", npairs, &flag1));
      RCK(buf("i.flag2", npairs, &flag2));
      RCK(buf("i.idx1", npairs + 1, &idx1));
      RCK(buf("i.idx2", npairs + 1, &idx2));
      RCK(buf_raw("i.fp2", sizeof(int) * (size_t)nslots * w2 * w2, reinterpret_cast<void**>(&fp2), INT_MAX));
      HCK(launch_isect_dedup_flags(pairs, npairs, w2 * w2, fp2, flag1, flag2, stream));
      const size_t stb2 = scan_tmp_bytes(npairs);
      void* scantmp;
      RCK(buf_raw("i.scantmp2", stb2, &scantmp, 0));
      HCK(scan_u32_to_u64(flag1, idx1, npairs, scantmp, stb2, stream));
      HCK(scan_u32_to_u64(flag2, idx2, npairs, scantmp, stb2, stream));
      RCK(d2h(&nu1, idx1 + npairs, sizeof(nu1)));
      RCK(d2h(&nu2, idx2 + npairs, sizeof(nu2)));
      RCK(buf(nm("su", 1, nxt).c_str(), npairs, &nsu1));
      RCK(buf(nm("su", 2, nxt).c_str(), npairs, &nsu2));
      RCK(buf(nm("reg", 1, nxt).c_str(), nu1, &nreg1));
      RCK(buf(nm("reg", 2, nxt).c_str(), nu2, &nreg2));
      HCK(launch_isect_units<T>(pairs, npairs, w2 * w2, su1, su2, reg1, reg2, nuv1, nuv2, flag1, flag2, idx1, idx2,
                                fp2, nsu1, nsu2, nreg1, nreg2, stream));
    }
    ri.units1 = (int64_t)nu1;
    ri.units2 = (int64_t)nu2;
    const int wr = g_nuv - 1;
    V *nC1, *nR1, *nC2, *nR2;
    RCK(buf(nm("C", 1, nxt).c_str(), (size_t)nu1 * wr * wr, &nC1));
    RCK(buf(nm("R", 1, nxt).c_str(), (size_t)nu1 * wr * wr, &nR1));
    RCK(buf(nm("C", 2, nxt).c_str(), (size_t)nu2 * wr * wr, &nC2));
    RCK(buf(nm("R", 2, nxt).c_str(), (size_t)nu2 * wr * wr, &nR2));
    tlo1 = thi1 = tlo2 = thi2 = nullptr;
    f32_ready = false;
    if (fuse_cells && g_nuv >= 2 && g_nuv <= kIcMaxNuv) {
      IcJob<T> j1, j2;
      RCK(surf_dev<T>(h1, &j1.sd));
      RCK(surf_dev<T>(h2, &j2.sd));
      j1.reg = nreg1;
      j2.reg = nreg2;
      j1.nunits = (int)nu1;
      j2.nunits = (int)nu2;
      j1.nuv = j2.nuv = g_nuv;
      j1.init_u = iu[1];
      j1.init_v = iv[1];
      j2.init_u = iu[2];
      j2.init_v = iv[2];
      j1.C = nC1;
      j1.R = nR1;
      j2.C = nC2;
      j2.R = nR2;
      if (!s32) {  // the FP32 screen builds its own boxes from its float records
        RCK(buf("tb.lo1", (size_t)nu1, &tlo1));
        RCK(buf("tb.hi1", (size_t)nu1, &thi1));
        RCK(buf("tb.lo2", (size_t)nu2, &tlo2));
        RCK(buf("tb.hi2", (size_t)nu2, &thi2));
        j1.lo = tlo1;
        j1.hi = thi1;
        j2.lo = tlo2;
        j2.hi = thi2;
      }
      if constexpr (std::is_same<T, double>::value) {
        if (s32 && !P.screen_out) {  // the next screen's float records and unit boxes, from the same pass
          const size_t n1 = (size_t)nu1 * wr * wr, n2 = (size_t)nu2 * wr * wr;
          RCK(buf("i32.X1", n1, &j1.X));
          RCK(buf("i32.Y1", n1, &j1.Y));
          RCK(buf("i32.X2", n2, &j2.X));
          RCK(buf("i32.Y2", n2, &j2.Y));
          RCK(buf("tb32.lo1", (size_t)nu1, &j1.flo));
          RCK(buf("tb32.hi1", (size_t)nu1, &j1.fhi));
          RCK(buf("tb32.lo2", (size_t)nu2, &j2.flo));
          RCK(buf("tb32.hi2", (size_t)nu2, &j2.fhi));
          f32_ready = true;
        }
      }
      if (prof) HCK(hipEventRecord(ev[0], stream));
      HCK(launch_isect_cells<T>(j1, &j2, use_k, stream));
      if (prof) {
        HCK(hipEventRecord(ev[1], stream));
        HCK(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        HCK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        KernelStat& kst = stats[std::is_same<T, float>::value ? "isect_cells_f32" : "isect_cells_f64"];
        kst.launches += 1;
        kst.ms += ms;
        kst.units += (double)(nu1 + nu2);
      }
    } else {
      RCK(isect_units<T>(h1, nreg1, (int)nu1, g_nuv, use_k, iu[1], iv[1], nC1, nR1, "i1"));
      RCK(isect_units<T>(h2, nreg2, (int)nu2, g_nuv, use_k, iu[2], iv[2], nC2, nR2, "i2"));
    }
    if (timing) {
      HCK(hipStreamSynchronize(stream));
      res->ms_eval += now_ms() - te;
    }
    cur = nxt;
    su1 = nsu1;
    su2 = nsu2;
    reg1 = nreg1;
    reg2 = nreg2;
    C1 = nC1;
    R1 = nR1;
    C2 = nC2;
    R2 = nR2;
    nunits1 = (int)nu1;
    nunits2 = (int)nu2;
    nslots = (long long)npairs;
    nuv1 = nuv2 = g_nuv;
  }
  res->isect_refine_nuv = g_nuv;
  if (creq) {
    // project every final pair onto the intersection curve, then order the points
    creq->points.clear();
    creq->ncurves = 0;
    if (npairs > 0) {
      SurfDev<double> s1, s2;
      RCK(surf_dev<double>(h1, &s1));
      RCK(surf_dev<double>(h2, &s2));
      CurveRec* d_cr;
      RCK(buf("crv.out", npairs, &d_cr));
      if (prof) HCK(hipEventRecord(ev[0], stream));
      HCK(launch_curve_proj<T>(s1, s2, pairs, (long long)npairs, su1, su2, reg1, reg2, nuv1, nuv2, region_box(r1),
                               region_box(r2), d_cr, stream));
      if (prof) {
        HCK(hipEventRecord(ev[1], stream));
        HCK(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        HCK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        KernelStat& kst = stats["curve_proj"];
        kst.launches += 1;
        kst.ms += ms;
        kst.units += (double)npairs;
      }
      const double tc0 = now_ms();
      void* hc;
      RCK(pinned(sizeof(CurveRec) * npairs, &hc));
      RCK(d2h(hc, d_cr, sizeof(CurveRec) * npairs));
      if (prof) host_stat("curve_records_d2h", now_ms() - tc0);
      RCK(trace_curves(static_cast<const CurveRec*>(hc), (long long)npairs, s1, s2, region_box(r1), region_box(r2), creq));
    }
  }
  if (s32 && P.screen_out) {  // report the last round's float candidates instead of the tested pairs
    pairs = cand_out;
    npairs = ncand_out;
    res->npairs = (int64_t)ncand_out;
  }
  if (out && cap_pairs > 0 && npairs > 0) {
    const unsigned long long m = std::min<unsigned long long>(npairs, (unsigned long long)cap_pairs);
    hp.resize(m);
    RCK(d2h(hp.data(), pairs, sizeof(IPair) * m));
    for (unsigned long long e = 0; e < m; ++e) out[e] = jsi_pair{0.0, hp[e].slot, hp[e].i, hp[e].j, 0};
  }
  HCK(hipStreamSynchronize(stream));  // the query's last kernels have finished when it returns
  res->ms_total = now_ms() - t0;
  return JSI_OK;
}

// Curves from the projected pairs: seeds (host), forward marches of every
// seed and backward marches of the chain heads (device, a counting pass and a
// writing pass each), chaining (host).
int jsi_ctx::trace_curves(const CurveRec* recs, long long n, const SurfDev<double>& s1, const SurfDev<double>& s2,
                          double4 box1, double4 box2, CurveReq* creq) {
  std::vector<MarchSeed> seeds;
  double th = now_ms();
  const double hstep = curve_seeds(recs, n, creq->tol, creq->spacing, seeds);
  if (prof) host_stat("curve_seeds", now_ms() - th);
  const int ns = (int)seeds.size();
  if (ns == 0) return JSI_OK;
  MarchSeed* d_seeds;
  int* d_starts;
  MarchOut* d_mo;
  long long* d_off;
  RCK(buf("crv.seeds", ns, &d_seeds));
  RCK(buf("crv.starts", ns, &d_starts));
  RCK(buf("crv.mo", ns, &d_mo));
  RCK(buf("crv.off", ns, &d_off));
  RCK(h2d(d_seeds, seeds.data(), sizeof(MarchSeed) * ns));
  auto march = [&](const std::vector<int>& starts, double dir, std::vector<MarchOut>& mo, std::vector<long long>& off,
                   std::vector<MarchPt>& pts) -> int {
    const int m = (int)starts.size();
    mo.assign(m, MarchOut());
    off.assign(m, 0);
    pts.clear();
    if (m == 0) return JSI_OK;
    RCK(h2d(d_starts, starts.data(), sizeof(int) * m));
    HCK(launch_march(s1, s2, d_seeds, ns, d_starts, m, dir, hstep, box1, box2, nullptr, nullptr, d_mo, stream));
    RCK(d2h(mo.data(), d_mo, sizeof(MarchOut) * m));
    long long total = 0;
    for (int k = 0; k < m; ++k) {
      off[k] = total;
      total += mo[k].n;
    }
    if (total == 0) return JSI_OK;
    MarchPt* d_pts;
    RCK(buf("crv.pts", (size_t)total, &d_pts));
    RCK(h2d(d_off, off.data(), sizeof(long long) * m));
    HCK(launch_march(s1, s2, d_seeds, ns, d_starts, m, dir, hstep, box1, box2, d_off, d_pts, d_mo, stream));
    pts.resize(total);
    RCK(d2h(pts.data(), d_pts, sizeof(MarchPt) * total));
    return JSI_OK;
  };
  std::vector<int> all(ns);
  for (int k = 0; k < ns; ++k) all[k] = k;
  std::vector<MarchOut> fwd, bwd;
  std::vector<long long> fwd_off, bwd_off;
  std::vector<MarchPt> fwd_pts, bwd_pts;
  th = now_ms();
  RCK(march(all, 1.0, fwd, fwd_off, fwd_pts));
  if (prof) host_stat("curve_march_fwd", now_ms() - th);
  std::vector<int> indeg(ns, 0), heads;
  for (int k = 0; k < ns; ++k)
    if (fwd[k].succ >= 0) ++indeg[fwd[k].succ];
  for (int k = 0; k < ns; ++k)
    if (indeg[k] == 0) heads.push_back(k);
  th = now_ms();
  RCK(march(heads, -1.0, bwd, bwd_off, bwd_pts));
  if (prof) host_stat("curve_march_bwd", now_ms() - th);
  th = now_ms();
  assemble_curves(seeds, fwd, fwd_off, fwd_pts, heads, bwd, bwd_off, bwd_pts, creq->points, &creq->ncurves);
  if (prof) host_stat("curve_assemble", now_ms() - th);
  return JSI_OK;
}

template <class T>
int jsi_ctx::eval_grid(int h, T ustep, T vstep, T ustart, T vstart, int nuv, T* host_out) {
  using V = vec4_t<T>;
  V g;
  g.x = ustart;
  g.y = vstart;
  g.z = ustep;
  g.w = vstep;
  V* reg;
  T* out;
  RCK(buf("eg.reg", 1, &reg));
  RCK(h2d(reg, &g, sizeof(V)));
  RCK(buf("eg.out", (size_t)nuv * nuv * 18, &out));
  RCK(eval_units<T>(h, reg, 1, nuv, kEvalFull18, 1, out, nullptr, "eg"));
  RCK(d2h(host_out, out, sizeof(T) * (size_t)nuv * nuv * 18));
  return JSI_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

int validate_surface(const int* iv, int ndata) {
  if (!iv) return JSI_ERR_INVALID;
  if (iv[1] < 1 || iv[4] < 1 || iv[2] <= iv[1] || iv[5] <= iv[4] || iv[3] != iv[2] + iv[1] + 1 ||
      iv[6] != iv[5] + iv[4] + 1) {
    set_err("inconsistent ivec {%d,%d,%d,%d,%d,%d,%d}", iv[0], iv[1], iv[2], iv[3], iv[4], iv[5], iv[6]);
    return JSI_ERR_INVALID;
  }
  const long long need = (long long)iv[3] + iv[6] + 3ll * iv[2] * iv[5] + ((iv[0] & 2) ? (long long)iv[2] * iv[5] : 0);
  if (ndata < need) {
    set_err("ndata %d < %lld floats required by ivec", ndata, need);
    return JSI_ERR_INVALID;
  }
  return JSI_OK;
}

// Knot vectors must be non-decreasing (FindSpan's binary search assumes it).
int validate_knots(const int* iv, const float* host_data) {
  for (int d = 0; d < 2; ++d) {
    const float* k = host_data + (d == 0 ? 0 : iv[3]);
    const int n = d == 0 ? iv[3] : iv[6];
    for (int i = 1; i < n; ++i)
      if (!(k[i - 1] <= k[i])) {
        set_err("knot vector %c is not non-decreasing at %d", d == 0 ? 'u' : 'v', i);
        return JSI_ERR_INVALID;
      }
  }
  return JSI_OK;
}

int validate_query(jsi_ctx* ctx, int h1, const jsi_region* r1, int h2, const jsi_region* r2, const jsi_params* P) {
  if (!ctx || !r1 || !r2 || !P) {
    set_err("null argument");
    return JSI_ERR_INVALID;
  }
  for (const jsi_region* r : {r1, r2})
    if (r->nuv < 2 || r->nuv > JSI_MAX_NUV) {
      set_err("nuv %d outside [2, %d]", r->nuv, JSI_MAX_NUV);
      return JSI_ERR_UNSUPPORTED;
    }
  if (P->precision < JSI_PREC_EXACT32 || P->precision > JSI_PREC_FP32SCREEN) {
    set_err("unknown precision %d", P->precision);
    return JSI_ERR_INVALID;
  }
  for (int h : {h1, h2})
    {
      std::lock_guard<std::mutex> lock(ctx->surfaces->mu);
      if (h < 0 || h >= (int)ctx->surfaces->recs.size() || !ctx->surfaces->recs[h].live) {
        set_err("invalid surface handle %d", h);
        return JSI_ERR_INVALID;
      }
    }
  return JSI_OK;
}

}  // namespace

extern "C" {

void jsi_params_default(jsi_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->k = 8;
  p->dist_refine_nuv = 16;
  p->dist_max_rounds = 4;
  p->isect_max_rounds = 6;
  p->isect_refine_nuv = 16;
  p->isect_halve_above = 512;
  p->isect_min_nuv = 4;
  p->use_compute_k = 1;
  p->precision = JSI_PREC_EXACT32;
  p->cull = 1;
  p->trace = 0;
}

const char* jsi_status_string(int s) {
  switch (s) {
    case JSI_OK: return "ok";
    case JSI_ERR_INVALID: return "invalid argument";
    case JSI_ERR_NOMEM: return "out of device memory";
    case JSI_ERR_DEVICE: return "device error";
    case JSI_ERR_CAPACITY: return "capacity exceeded";
    case JSI_ERR_UNSUPPORTED: return "unsupported";
    case JSI_ERR_BUSY: return "busy";
    default: return "unknown status";
  }
}

const char* jsi_last_error(void) { return g_err; }

int jsi_ctx_create(int device, size_t hbm_cap_bytes, jsi_ctx** out) {
  if (!out) return JSI_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  HCK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_err("device %d not present (%d devices)", device, ndev);
    return JSI_ERR_INVALID;
  }
  HCK(hipSetDevice(device));
  jsi_ctx* c = new jsi_ctx();
  c->device = device;
  c->debug_ = std::getenv("JSI_DEBUG") != nullptr;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_err("hipStreamCreate: %s", hipGetErrorString(e));
    return JSI_ERR_DEVICE;
  }
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  c->pool->cap = hbm_cap_bytes ? hbm_cap_bytes : (free_b / 4) * 3;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->ncu = prop.multiProcessorCount;
  if (c->ncu <= 0) c->ncu = 256;
  c->grid = c->ncu * 8;
  if (hipHostMalloc(&c->pin_small_, jsi_ctx::kPinSmall, hipHostMallocDefault) != hipSuccess) c->pin_small_ = nullptr;
  *out = c;
  return JSI_OK;
}

// A worker context: own stream and buffers, the parent's surfaces and cap.
int jsi_ctx_create_worker(jsi_ctx* parent, jsi_ctx** out) {
  if (!parent || !out) return JSI_ERR_INVALID;
  HCK(hipSetDevice(parent->device));
  jsi_ctx* c = new jsi_ctx();
  c->device = parent->device;
  c->debug_ = parent->debug_;
  c->worker = true;
  c->surfaces = parent->surfaces;
  c->pool = parent->pool;  // one account for the parent and all its workers
  c->root = parent->root ? parent->root : parent;
  c->ncu = parent->ncu;
  c->grid = parent->grid;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_err("hipStreamCreate: %s", hipGetErrorString(e));
    return JSI_ERR_DEVICE;
  }
  if (hipHostMalloc(&c->pin_small_, jsi_ctx::kPinSmall, hipHostMallocDefault) != hipSuccess) c->pin_small_ = nullptr;
  *out = c;
  return JSI_OK;
}

int jsi_ctx_destroy(jsi_ctx* c) {
  if (!c) return JSI_ERR_INVALID;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  jsi_ctx_drop_batcher(c);  // its plans release their memory to the pool first
  for (auto& kv : c->bufs)
    if (kv.second.p) {
      (void)hipFree(kv.second.p);
      c->pool->release(kv.second.bytes);
    }
  if (!c->worker) {
    std::lock_guard<std::mutex> lock(c->surfaces->mu);
    for (auto& s : c->surfaces->recs) {
      if (s.owned && s.d_data) (void)hipFree(s.d_data);
      if (s.d_data64) (void)hipFree(s.d_data64);
    }
    c->surfaces->recs.clear();
  }
  for (hipEvent_t e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->pinned_) (void)hipHostFree(c->pinned_);
  if (c->pin_small_) (void)hipHostFree(c->pin_small_);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return JSI_OK;
}

int jsi_ctx_profile(jsi_ctx* c, int enable) {
  if (!c) return JSI_ERR_INVALID;
  HCK(hipSetDevice(c->device));
  if (enable && !c->ev[0]) {
    // device-scope release: no system-wide cache writeback inside the timed
    // interval, so the pair brackets the kernel itself
    HCK(hipEventCreateWithFlags(&c->ev[0], hipEventReleaseToDevice));
    HCK(hipEventCreateWithFlags(&c->ev[1], hipEventReleaseToDevice));
  }
  c->prof = enable ? 1 : 0;
  c->stats.clear();
  return JSI_OK;
}

int jsi_ctx_kernel_stats(jsi_ctx* c, const char* name, int64_t* launches, double* ms, double* units) {
  if (!c || !name) return JSI_ERR_INVALID;
  std::lock_guard<std::mutex> lock(c->stats_mu);
  auto it = c->stats.find(name);
  const KernelStat st = it == c->stats.end() ? KernelStat() : it->second;
  if (launches) *launches = st.launches;
  if (ms) *ms = st.ms;
  if (units) *units = st.units;
  return JSI_OK;
}

int jsi_ctx_sync(jsi_ctx* c) {
  if (!c) return JSI_ERR_INVALID;
  HCK(hipStreamSynchronize(c->stream));
  return JSI_OK;
}

int jsi_surface_upload(jsi_ctx* c, const int ivec[7], const float* data, int ndata, int* handle) {
  if (!c || !ivec || !data || !handle) return JSI_ERR_INVALID;
  RCK(validate_surface(ivec, ndata));
  RCK(validate_knots(ivec, data));
  HCK(hipSetDevice(c->device));
  SurfaceRec s;
  s.live = true;
  std::memcpy(s.ivec, ivec, sizeof(s.ivec));
  s.ndata = ndata;
  s.owned = true;
  const size_t bytes = sizeof(float) * (size_t)ndata;
  if (!c->pool->reserve(bytes)) {
    set_err("HBM cap reached uploading a surface");
    return JSI_ERR_NOMEM;
  }
  {
    hipError_t e = hipMalloc(&s.d_data, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      c->pool->release(bytes);
      set_err("hipMalloc(%zu) for a surface failed: %s", bytes, hipGetErrorString(e));
      return JSI_ERR_NOMEM;
    }
  }
  HCK(hipMemcpy(s.d_data, data, bytes, hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> lock(c->surfaces->mu);
  c->surfaces->recs.push_back(s);
  *handle = (int)c->surfaces->recs.size() - 1;
  return JSI_OK;
}

int jsi_surface_register_device(jsi_ctx* c, const int* d_ivec, const float* d_data, int ndata, int* handle) {
  if (!c || !d_ivec || !d_data || !handle) return JSI_ERR_INVALID;
  HCK(hipSetDevice(c->device));
  SurfaceRec s;
  HCK(hipMemcpy(s.ivec, d_ivec, sizeof(s.ivec), hipMemcpyDeviceToHost));
  RCK(validate_surface(s.ivec, ndata));
  {
    std::vector<float> knots((size_t)s.ivec[3] + s.ivec[6]);
    HCK(hipMemcpy(knots.data(), d_data, sizeof(float) * knots.size(), hipMemcpyDeviceToHost));
    RCK(validate_knots(s.ivec, knots.data()));
  }
  s.live = true;
  s.ndata = ndata;
  s.owned = false;
  s.d_data = const_cast<float*>(d_data);
  std::lock_guard<std::mutex> lock(c->surfaces->mu);
  c->surfaces->recs.push_back(s);
  *handle = (int)c->surfaces->recs.size() - 1;
  return JSI_OK;
}

int jsi_surface_release(jsi_ctx* c, int h) {
  if (!c) return JSI_ERR_INVALID;
  HCK(hipStreamSynchronize(c->stream));
  std::lock_guard<std::mutex> lock(c->surfaces->mu);
  auto& tab = c->surfaces->recs;
  if (h < 0 || h >= (int)tab.size() || !tab[h].live) return JSI_ERR_INVALID;
  SurfaceRec& s = tab[h];
  if (s.owned && s.d_data) {
    HCK(hipFree(s.d_data));
    c->pool->release(sizeof(float) * (size_t)s.ndata);
  }
  if (s.d_data64) {
    HCK(hipFree(s.d_data64));
    c->pool->release(sizeof(double) * (size_t)s.ndata);
  }
  s = SurfaceRec();
  return JSI_OK;
}

int jsi_eval_grid_f32(jsi_ctx* c, int h, float ustep, float vstep, float ustart, float vstart, int nuv, float* out) {
  if (!c || !out || nuv < 1 || nuv > 16384) return JSI_ERR_INVALID;
  HCK(hipSetDevice(c->device));
  return c->eval_grid<float>(h, ustep, vstep, ustart, vstart, nuv, out);
}

int jsi_eval_grid_f64(jsi_ctx* c, int h, double ustep, double vstep, double ustart, double vstart, int nuv,
                      double* out) {
  if (!c || !out || nuv < 1 || nuv > 16384) return JSI_ERR_INVALID;
  HCK(hipSetDevice(c->device));
  return c->eval_grid<double>(h, ustep, vstep, ustart, vstart, nuv, out);
}

static int validate_distance(jsi_ctx* c, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                             const jsi_params* P, jsi_result* res) {
  RCK(validate_query(c, h1, r1, h2, r2, P));
  if (!res || P->k < 1 || P->k > JSI_MAX_K || P->dist_max_rounds < 1 || P->dist_refine_nuv < 2 ||
      P->dist_refine_nuv > JSI_MAX_NUV || P->polish < 0 || P->polish > JSI_MAX_K || !(P->param_tol >= 0.0) ||
      !std::isfinite(P->param_tol)) {
    set_err("invalid distance parameters");
    return JSI_ERR_INVALID;
  }
  if (P->param_tol > 0.0 && P->precision == JSI_PREC_EXACT32) {
    set_err("param_tol needs FP64 parameters (precision FP64 or FP32SCREEN)");
    return JSI_ERR_INVALID;
  }
  return JSI_OK;
}

extern "C++" int min_distance_in_scratch(jsi_ctx* c, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                            const jsi_params* P, jsi_result* res, jsi_pair* pairs_out, int pairs_cap, void* scratch,
                            size_t bytes, unsigned long long cand_cap) {
  RCK(validate_distance(c, h1, r1, h2, r2, P, res));
  std::memset(res, 0, sizeof(*res));
  HCK(hipSetDevice(c->device));
  c->scratch_.base = static_cast<char*>(scratch);
  c->scratch_.cap = bytes;
  c->scratch_.off = 0;
  c->scratch_.bufs.clear();
  const unsigned long long saved_cap = c->cand_cap_max_;
  c->cand_cap_max_ = cand_cap;
  int rc = P->precision == JSI_PREC_EXACT32 ? c->min_distance<float>(h1, *r1, h2, *r2, *P, res, pairs_out, pairs_cap)
                                            : c->min_distance<double>(h1, *r1, h2, *r2, *P, res, pairs_out, pairs_cap);
  // the stream has drained on success (results were read back); after a
  // failure, wait for what was launched before handing the scratch back
  if (rc != JSI_OK) (void)hipStreamSynchronize(c->stream);
  if (c->debug_)
    fprintf(stderr, "[jsi] scratch: %zu of %zu bytes used\n", c->scratch_.off, c->scratch_.cap);
  c->scratch_.base = nullptr;
  c->scratch_.bufs.clear();
  c->cand_cap_max_ = saved_cap;
  return rc;
}

int jsi_min_distance(jsi_ctx* c, int h1, const jsi_region* r1, int h2, const jsi_region* r2, const jsi_params* P,
                     jsi_result* res, jsi_pair* pairs_out, int pairs_cap) {
  RCK(validate_distance(c, h1, r1, h2, r2, P, res));
  std::memset(res, 0, sizeof(*res));
  HCK(hipSetDevice(c->device));
  if (P->precision == JSI_PREC_EXACT32) return c->min_distance<float>(h1, *r1, h2, *r2, *P, res, pairs_out, pairs_cap);
  return c->min_distance<double>(h1, *r1, h2, *r2, *P, res, pairs_out, pairs_cap);
}

int jsi_intersect(jsi_ctx* c, int h1, const jsi_region* r1, int h2, const jsi_region* r2, const jsi_params* P,
                  jsi_result* res, jsi_pair* pairs_out, int pairs_cap) {
  RCK(validate_query(c, h1, r1, h2, r2, P));
  if (!res || P->isect_max_rounds < 1 || P->isect_refine_nuv < 2 || P->isect_refine_nuv > JSI_MAX_NUV ||
      P->isect_min_nuv < 1) {
    set_err("invalid intersection parameters");
    return JSI_ERR_INVALID;
  }
  std::memset(res, 0, sizeof(*res));
  HCK(hipSetDevice(c->device));
  if (P->precision == JSI_PREC_EXACT32) return c->intersect<float>(h1, *r1, h2, *r2, *P, res, pairs_out, pairs_cap);
  return c->intersect<double>(h1, *r1, h2, *r2, *P, res, pairs_out, pairs_cap);
}

int jsi_intersect_curves(jsi_ctx* c, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                         const jsi_params* P, double spacing, double tol, jsi_result* res, jsi_curve_point* pts,
                         int cap, int* npts, int* ncurves) {
  RCK(validate_query(c, h1, r1, h2, r2, P));
  if (!res || !npts || !ncurves || (cap > 0 && !pts) || P->isect_max_rounds < 1 || P->isect_refine_nuv < 2 ||
      P->isect_refine_nuv > JSI_MAX_NUV || P->isect_min_nuv < 1 || !std::isfinite(spacing) ||
      !std::isfinite(tol) || tol < 0.0) {
    set_err("invalid intersection-curve parameters");
    return JSI_ERR_INVALID;
  }
  std::memset(res, 0, sizeof(*res));
  *npts = 0;
  *ncurves = 0;
  HCK(hipSetDevice(c->device));
  CurveReq req;
  req.spacing = spacing;
  req.tol = tol > 0.0 ? tol : 1.0e-9;
  int rc;
  if (P->precision == JSI_PREC_EXACT32)
    rc = c->intersect<float>(h1, *r1, h2, *r2, *P, res, nullptr, 0, &req);
  else
    rc = c->intersect<double>(h1, *r1, h2, *r2, *P, res, nullptr, 0, &req);
  if (rc != JSI_OK) return rc;
  const size_t n = req.points.size();
  *npts = (int)std::min<size_t>(n, INT_MAX);
  *ncurves = req.ncurves;
  if (cap > 0) std::memcpy(pts, req.points.data(), sizeof(jsi_curve_point) * std::min<size_t>(n, (size_t)cap));
  return JSI_OK;
}

}  // extern "C"

// used by the operator ABI (jsi_opabi.hip)
template int jsi_ctx::dist_topk<float>(int, const int*, const int*, int, int, const float4*, int, int, const float4*,
                                       int, int, Cand**, long long*, long long*, Rerank, const unsigned long long*,
                                       unsigned long long, TileBoxes<float4>, Cand*, int);

// Used by the batched distance path (jsi_batch.hip).
template int jsi_ctx::surf_dev<float>(int, SurfDev<float>*);
template int jsi_ctx::surf_dev<double>(int, SurfDev<double>*);
template int jsi_ctx::min_distance<float>(int, const jsi_region&, int, const jsi_region&, const jsi_params&,
                                          jsi_result*, jsi_pair*, int);
template int jsi_ctx::min_distance<double>(int, const jsi_region&, int, const jsi_region&, const jsi_params&,
                                           jsi_result*, jsi_pair*, int);

extern "C" int jsi_tune_set(jsi_ctx* ctx, int degree, int nuv, const char* knob, int value) {
  if (!ctx || !knob || degree < 1 || degree > 24 || nuv < 2 || nuv > JSI_MAX_NUV) return JSI_ERR_INVALID;
  jsi_ctx* r = ctx->root ? ctx->root : ctx;
  std::lock_guard<std::mutex> lock(r->stats_mu);
  jsi::Tuning& t = r->tune_table[std::make_pair(degree, nuv)];
  const std::string k(knob);
  if (k == "seed_pairs" && value >= 1 && value <= 256) {
    t.seed_pairs = value;
  } else if (k == "filter_walk" && value >= 1 && value <= 16) {
    t.filter_walk = value;
  } else if (k == "r1_scale" && value >= 10 && value <= 1000) {
    t.r1_scale = value;
  } else if (k == "r1_seeds" && value >= 1 && value <= 256) {
    t.r1_seeds = value;
  } else {
    return JSI_ERR_INVALID;
  }
  return JSI_OK;
}

extern "C" int jsi_tune_clear(jsi_ctx* ctx) {
  if (!ctx) return JSI_ERR_INVALID;
  jsi_ctx* r = ctx->root ? ctx->root : ctx;
  std::lock_guard<std::mutex> lock(r->stats_mu);
  r->tune_table.clear();
  return JSI_OK;
}

