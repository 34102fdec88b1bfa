// jsi_runtime.hpp - the context object behind the C ABI handle jsi_ctx.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <initializer_list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/jsi_mi355x.h"
#include "jsi_internal.hpp"

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct SurfaceRec {
  bool live = false;
  int ivec[7] = {0, 0, 0, 0, 0, 0, 0};
  int ndata = 0;
  float* d_data = nullptr;      // NurbsFace data in HBM (owned unless registered from device pointers)
  bool owned = false;
  double* d_data64 = nullptr;   // promoted copy for the FP64 recipe, made on first FP64 use
};

// One context per device: a stream, a grow-only set of named device buffers
// (reused across queries, so steady-state queries allocate nothing) and the
// resident surfaces.  `cap` bounds the bytes the context may hold.
// Resident surfaces, shared by a context and the worker contexts of its
// batch runner / server (each worker has its own stream and buffers).
struct SurfaceTable {
  std::mutex mu;
  std::vector<SurfaceRec> recs;
};

// FP32SCREEN: float centres are screened, FP64 centres re-rank the candidates.
struct RerankSpec {
  const double4* C1 = nullptr;
  const double4* C2 = nullptr;
  const unsigned long long* absmax = nullptr;
};

// Per-tile boxes of both sides' centres, when the evaluation already wrote them.
template <class V>
struct TileBoxes {
  const V* lo1 = nullptr;
  const V* hi1 = nullptr;
  const V* lo2 = nullptr;
  const V* hi2 = nullptr;
};

// Intersection-curve request of jsi_intersect_curves.
struct CurveReq {
  double spacing = 0.0;
  double tol = 1.0e-9;
  std::vector<jsi_curve_point> points;
  int ncurves = 0;
};

// Device bytes held by a context and every worker context made from it (one
// counter for all of them), bounded by `cap`.
struct HbmPool {
  std::atomic<size_t> used{0};
  size_t cap = 0;
  bool reserve(size_t bytes) {
    size_t u = used.load();
    do {
      if (u + bytes > cap) return false;
    } while (!used.compare_exchange_weak(u, u + bytes));
    return true;
  }
  void release(size_t bytes) { used.fetch_sub(bytes); }
};

namespace jsi {
class DistBatcher;
struct DistBatcherDelete {  // defined where DistBatcher is complete (jsi_batch.hip)
  void operator()(DistBatcher* b) const;
};
}

struct KernelStat {
  long long launches = 0;
  double ms = 0.0;     // summed launch durations (HIP events on the context stream)
  double units = 0.0;  // summed algorithmic units (cell pairs for the screens)
};

struct jsi_ctx {
  int device = 0;
  int prof = 0;  // time the screen kernels with HIP events
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::map<std::string, KernelStat> stats;
  hipStream_t stream = nullptr;
  std::shared_ptr<HbmPool> pool = std::make_shared<HbmPool>();
  int ncu = 0;
  int grid = 0;  // persistent-kernel grid size
  std::shared_ptr<SurfaceTable> surfaces = std::make_shared<SurfaceTable>();
  bool worker = false;  // shares its parent's surfaces; destroy frees only its own buffers
  std::unique_ptr<jsi::DistBatcher, jsi::DistBatcherDelete> batcher;  // batched distance queries, made on first use
  std::map<std::string, DevBuf> bufs;
  // Caller-provided scratch (the allocation-free distance entry point): while
  // `base` is set, buf_raw carves every named buffer of the query from
  // [base, base + cap) in request order (a buffer requested again reuses its
  // piece when it is large enough), and no device memory is allocated.
  struct Scratch {
    char* base = nullptr;
    size_t cap = 0, off = 0;
    std::map<std::string, DevBuf> bufs;
  } scratch_;
  unsigned long long cand_cap_max_ = 1ull << 22;  // candidate records per screening pass (power of two)
  void* thresh_ptr_ = nullptr;  // device threshold of the screen in flight
  bool debug_ = false;          // JSI_DEBUG: per-screen counts on stderr

  // Counters shared by a context and its workers (kept on the root context,
  // read with jsi_ctx_kernel_stats): batched queries, batches, fallbacks.
  jsi_ctx* root = nullptr;
  std::mutex stats_mu;
  // Autotuned launch and seed choices per (degree, nuv) (kept on the root
  // context, guarded by stats_mu), and the choices of the query in flight.
  std::map<std::pair<int, int>, jsi::Tuning> tune_table;
  jsi::Tuning tune_;
  jsi::Tuning tuning_for(int degree, int nuv) {
    jsi_ctx* r = root ? root : this;
    std::lock_guard<std::mutex> lock(r->stats_mu);
    const jsi::Tuning* best = nullptr;
    long long bd = 0;
    for (const auto& e : r->tune_table) {  // same degree, nearest nuv
      if (e.first.first != degree) continue;
      const long long d = std::llabs((long long)e.first.second - nuv);
      if (!best || d < bd) {
        best = &e.second;
        bd = d;
      }
    }
    return best ? *best : jsi::Tuning();
  }
  void shared_stat(const char* name, double units) {
    jsi_ctx* r = root ? root : this;
    std::lock_guard<std::mutex> lock(r->stats_mu);
    KernelStat& st = r->stats[name];
    st.launches += 1;
    st.units += units;
  }
  // profiling: host-side time of a named step (launches = calls, ms = summed time)
  void host_stat(const char* name, double ms) {
    KernelStat& st = stats[name];
    st.launches += 1;
    st.ms += ms;
  }
  int buf_raw(const char* name, size_t bytes, void** out, int fill_int);
  template <class X>
  int buf(const char* name, size_t count, X** out) {
    void* p = nullptr;
    int r = buf_raw(name, count * sizeof(X), &p, 0);
    *out = static_cast<X*>(p);
    return r;
  }
  // page-locked host staging (grown on demand, freed with the context)
  void* pinned_ = nullptr;
  size_t pinned_bytes_ = 0;
  // page-locked staging of the copies of up to 1 MB a query makes (round
  // totals, regions, result lists), allocated with the context: a copy to or
  // from pageable memory goes through the runtime's own staging and waits
  // longer
  static constexpr size_t kPinSmall = size_t(1) << 20;
  void* pin_small_ = nullptr;
  int pinned(size_t bytes, void** out);
  int d2h(void* dst, const void* src, size_t bytes);
  // several small device -> host copies with one wait (together at most 1 MB)
  struct Copy {
    void* dst;
    const void* src;
    size_t bytes;
  };
  int d2h_many(std::initializer_list<Copy> copies);
  int h2d(void* dst, const void* src, size_t bytes);

  template <class T>
  int surf_dev(int h, jsi::SurfDev<T>* out);
  template <class T>
  int eval_units(int h, const jsi::vec4_t<T>* reg, int nunits, int nuv, int mode, int steps_given, T* out,
                 unsigned long long* kmax, const char* tag);
  using Rerank = RerankSpec;
  template <class T>
  int eval_centres(int h1, const jsi::vec4_t<T>* reg1, int nunits1, int nuv1, const jsi::EcOut<T>& o1, int h2,
                   const jsi::vec4_t<T>* reg2, int nunits2, int nuv2, const jsi::EcOut<T>& o2);
  int isect_screen32(int round, int nslots, const int* su1, const int* su2, int nunits1, int w1, int nunits2, int w2,
                     const double4* C1, const double4* R1, const double4* C2, const double4* R2,
                     const float4* V1c, const float4* V1r, const float4* V2c, const float4* V2r, int cull,
                     const char* pairs_name, jsi::IPair** pairs_out, unsigned long long* npairs_out,
                     jsi::IPair** cand_out, unsigned long long* ncand_out);
  template <class T>
  int dist_topk(int nslots, const int* su1, const int* su2, int w1, int nunits1, const jsi::vec4_t<T>* C1, int w2,
                int nunits2, const jsi::vec4_t<T>* C2, int k, int cull, jsi::Cand** rec_out, long long* n_out,
                long long* kk_out, Rerank rr = Rerank(), const unsigned long long* exp_absmax = nullptr,
                unsigned long long cand_cap = 0, TileBoxes<jsi::vec4_t<T>> tbx = TileBoxes<jsi::vec4_t<T>>(),
                jsi::Cand* last_out = nullptr, int screen32 = 0);
  template <class T>
  int min_distance(int h1, const jsi_region& r1, int h2, const jsi_region& r2, const jsi_params& P,
                   jsi_result* res, jsi_pair* out, int cap_pairs);
  template <class T>
  int isect_units(int h, const jsi::vec4_t<T>* reg, int nunits, int nuv, int use_k, T init_u, T init_v,
                  jsi::vec4_t<T>* C, jsi::vec4_t<T>* R, const char* tag);
  template <class T>
  int polish(int h1, const jsi_region& r1, int h2, const jsi_region& r2, const jsi_params& P, const jsi::Cand* rec,
             long long kk, const int* su1, const int* su2, const jsi::vec4_t<T>* reg1, const jsi::vec4_t<T>* reg2,
             int nuv1, int nuv2, jsi_result* res);
  template <class T>
  int initial_k(int nuv1, int nuv2, const jsi_region& r1, const jsi_region& r2, const T* iu, const T* iv,
                jsi_result* res, Copy extra = Copy{nullptr, nullptr, 0});
  template <class T>
  int intersect(int h1, const jsi_region& r1, int h2, const jsi_region& r2, const jsi_params& P, jsi_result* res,
                jsi_pair* out, int cap_pairs, CurveReq* creq = nullptr);
  int trace_curves(const jsi::CurveRec* recs, long long n, const jsi::SurfDev<double>& s1,
                   const jsi::SurfDev<double>& s2, double4 box1, double4 box2, CurveReq* creq);
  template <class T>
  int eval_grid(int h, T ustep, T vstep, T ustart, T vstart, int nuv, T* host_out);
};

// Worker context sharing the parent's surfaces (batch runner / server).
extern "C" int jsi_ctx_create_worker(jsi_ctx* parent, jsi_ctx** out);

// Runs n queries on one context: distance queries of equal shape as batches
// (jsi_batch.hip), the others one at a time.
__attribute__((visibility("hidden"))) int jsi_ctx_run_distance_batch(jsi_ctx* c, int n, const jsi_query* const* qs,
                                                                     jsi_result* const* res, int* status);
// Shape key of a distance query that can run batched ("" = it runs alone).
__attribute__((visibility("hidden"))) std::string jsi_ctx_batch_key(jsi_ctx* c, const jsi_query& q);
// Frees the context's batch plans.
__attribute__((visibility("hidden"))) void jsi_ctx_drop_batcher(jsi_ctx* c);
// Device bytes one batched distance query holds.
__attribute__((visibility("hidden"))) size_t jsi_batch_query_bytes(const jsi_query& q);

// jsi_min_distance with every device buffer carved from `scratch` (device
// memory of `bytes` bytes that the caller owns); allocates no device memory.
// Candidate records per screening pass are capped at `cand_cap` (a power of
// two; more candidates take the overflow path, results unchanged).
__attribute__((visibility("hidden"))) int min_distance_in_scratch(jsi_ctx* c, int h1, const jsi_region* r1, int h2,
                                                                  const jsi_region* r2, const jsi_params* P,
                                                                  jsi_result* res, jsi_pair* pairs_out,
                                                                  int pairs_cap, void* scratch, size_t bytes,
                                                                  unsigned long long cand_cap);
