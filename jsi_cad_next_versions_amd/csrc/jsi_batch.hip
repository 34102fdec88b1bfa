// jsi_batch.hip - batched distance queries.
//
// Queries of one shape (precision, grids, k, rounds, polish, kernel variants)
// run together: every kernel of the query sequence is launched once for the
// whole batch with blockIdx.y = query (jsi_kernels.hip, record_dist_batch),
// so the small per-query kernels (seeds, selection, rank sort, dedup, polish)
// fill the chip instead of occupying a hardware queue each, and the screens
// of all queries share one launch.  A plan holds the device buffers of a
// shape for a fixed number of queries and the launch sequence captured as a
// HIP graph; each execution rewrites the host image of the launch arguments
// (surfaces and regions change from batch to batch), and the graph copies it
// to the device, runs every launch and copies the results back, with no host
// round trip in between.  A query that overflowed a bounded buffer is run
// again on the single-query path, which handles it by repeating steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jsi_mi355x.h"
#include "jsi_internal.hpp"
#include "jsi_runtime.hpp"

namespace jsi {

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int tiles_of(int w) {
  const int tw = (w + kTile - 1) / kTile;
  return tw * tw;
}

// Rounds of a distance query (dist_max_rounds, or param_tol's count).
int query_rounds(const jsi_params& P, const jsi_region& r1, const jsi_region& r2) {
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

uint64_t fnv_bytes(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

template <class T>
double d2_of(uint64_t b) {
  if (std::is_same<T, float>::value) {
    const uint32_t u = (uint32_t)b;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
  }
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}

double4 box_of(const jsi_region& r) {
  double4 b;
  b.x = std::min((double)r.ustart, (double)r.ustop);
  b.y = std::min((double)r.vstart, (double)r.vstop);
  b.z = std::max((double)r.ustart, (double)r.ustop);
  b.w = std::max((double)r.vstart, (double)r.vstop);
  return b;
}

}  // namespace

// Device buffers of nq queries of one shape, the argument arena, the graph.
struct DistPlan {
  DistBatchShape sh;
  int nq = 0;
  int dbl = 0;
  size_t per_q = 0;                // bytes of one query's working buffers
  char* dev = nullptr;             // working buffers, then the outputs
  size_t dev_bytes = 0;
  unsigned* fail = nullptr;        // [nq]
  Cand* lastrec = nullptr;         // [nq][rounds]
  int* nunits = nullptr;           // [nq][2 rounds]
  Cand* sorted = nullptr;          // [rounds][nq][k]
  PolishRec* pol = nullptr;        // [nq][m]
  size_t out_bytes = 0;            // fail .. pol, copied back in one piece
  unsigned char* d_arena = nullptr;
  unsigned char* h_arena = nullptr;  // page-locked image of the launch arguments
  size_t arena_bytes = 0;
  void* h_out = nullptr;           // page-locked copy of the outputs
  unsigned* h_fail0 = nullptr;     // page-locked initial failure flags: 1 marks an unused slot
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool no_graph = false;           // capture failed: launch directly
  HbmPool* pool = nullptr;
  size_t reserved = 0;
  uint64_t last = 0;               // LRU clock of the plan cache

  ~DistPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dev) (void)hipFree(dev);
    if (d_arena) (void)hipFree(d_arena);
    if (h_arena) (void)hipHostFree(h_arena);
    if (h_out) (void)hipHostFree(h_out);
    if (h_fail0) (void)hipHostFree(h_fail0);
    if (pool && reserved) pool->release(reserved);
  }
};

// Offsets of one query's working buffers (see DistBatchQ).
struct QLayout {
  size_t reg[2][2], su[2][2], C[2][2], tbl[2][2], tbh[2][2], c32[2], tb32l[2], tb32h[2];
  size_t amax, list, small, cand, topk, st, hist, range, seedwork, best, fp1, fp2, dwork;
  size_t bytes = 0;
  size_t take(size_t b) {
    const size_t o = align256(bytes);
    bytes = o + std::max(b, (size_t)256);
    return o;
  }
};

static QLayout layout_of(const DistBatchShape& sh, size_t vsz) {
  QLayout L;
  const int k = sh.k;
  const size_t wr = (size_t)(sh.nref - 1) * (sh.nref - 1);
  const int nuv[2] = {sh.nuv1, sh.nuv2};
  const bool s32 = sh.prec == JSI_PREC_FP32SCREEN;
  for (int s = 0; s < 2; ++s) {
    const size_t cells1 = (size_t)(nuv[s] - 1) * (nuv[s] - 1);
    const size_t t1 = (size_t)tiles_of(nuv[s] - 1);
    for (int p = 0; p < 2; ++p) {
      L.reg[s][p] = L.take(vsz * (size_t)k);
      L.su[s][p] = L.take(sizeof(int) * (size_t)k);
    }
    // one buffer of centres and tile boxes per side: a round's centres are
    // last read by its screen, before the next round's evaluation writes
    // them (regions and slot maps stay ping-pong: the dedup reads one, writes
    // the other)
    L.C[s][0] = L.C[s][1] = L.take(vsz * std::max(cells1, (size_t)k * wr));
    L.tbl[s][0] = L.tbl[s][1] = L.take(vsz * std::max(t1, (size_t)k));
    L.tbh[s][0] = L.tbh[s][1] = L.take(vsz * std::max(t1, (size_t)k));
    L.c32[s] = s32 ? L.take(sizeof(float4) * cells1) : 0;
    L.tb32l[s] = s32 ? L.take(sizeof(float4) * t1) : 0;
    L.tb32h[s] = s32 ? L.take(sizeof(float4) * t1) : 0;
  }
  const size_t tp1 = (size_t)tiles_of(sh.nuv1 - 1), tp2 = (size_t)tiles_of(sh.nuv2 - 1);
  L.amax = L.take(sizeof(unsigned long long) * kAmaxSlots);
  L.list = L.take(sizeof(uint64_t) * std::max({tp1 * tp2, (size_t)k, (size_t)256}));
  L.small = L.take(64 * sizeof(unsigned));
  L.cand = L.take(sizeof(Cand) * sh.cand_cap);
  L.topk = L.take(sizeof(Cand) * sh.small_cap);
  L.st = L.take(sizeof(SelState));
  L.hist = L.take(sizeof(unsigned) * 4096);
  L.range = L.take(2 * sizeof(unsigned long long));
  L.seedwork = L.take(seed_work_bytes());
  L.best = L.take(sizeof(unsigned long long) * std::max(tp1, (size_t)k));
  L.fp1 = L.take(sizeof(int) * std::max((size_t)(sh.nuv1 - 1) * (sh.nuv1 - 1), (size_t)k * wr));
  L.fp2 = L.take(sizeof(int) * std::max((size_t)(sh.nuv2 - 1) * (sh.nuv2 - 1), (size_t)k * wr));
  L.dwork = L.take(dedup_work_bytes());
  L.bytes = align256(L.bytes);
  return L;
}

// Batch plans shared by a context and all its workers: a worker checks a plan
// out for one batch and returns it; idle plans beyond kIdleBudget bytes are
// freed, least recently used first.
class PlanCache {
 public:
  std::unique_ptr<DistPlan> checkout(const std::string& key) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = idle_.find(key);
    if (it == idle_.end()) return nullptr;
    std::unique_ptr<DistPlan> p = std::move(it->second);
    idle_bytes_ -= p->dev_bytes;
    idle_.erase(it);
    return p;
  }
  void checkin(const std::string& key, std::unique_ptr<DistPlan> p) {
    std::vector<std::unique_ptr<DistPlan>> drop;
    {
      std::lock_guard<std::mutex> lock(mu_);
      seen_.insert(key);
      p->last = ++clock_;
      idle_bytes_ += p->dev_bytes;
      idle_.emplace(key, std::move(p));
      while (idle_bytes_ > kIdleBudget && idle_.size() > 1) {
        auto victim = idle_.begin();
        for (auto e = idle_.begin(); e != idle_.end(); ++e)
          if (e->second->last < victim->second->last) victim = e;
        idle_bytes_ -= victim->second->dev_bytes;
        drop.push_back(std::move(victim->second));
        idle_.erase(victim);
      }
    }
    // freed outside the lock (hipFree synchronises the device)
  }
  bool seen(const std::string& key) {
    std::lock_guard<std::mutex> lock(mu_);
    return seen_.count(key) != 0;
  }

 private:
  static constexpr size_t kIdleBudget = size_t(48) << 30;
  std::mutex mu_;
  std::multimap<std::string, std::unique_ptr<DistPlan>> idle_;
  std::set<std::string> seen_;
  size_t idle_bytes_ = 0;
  uint64_t clock_ = 0;
};

class DistBatcher {
 public:
  DistBatcher(jsi_ctx* c, std::shared_ptr<PlanCache> cache) : c_(c), cache_(std::move(cache)) {
    const char* g = std::getenv("JSI_BATCH_GRAPH");
    graphs_ = !(g && std::strcmp(g, "0") == 0);
  }
  const std::shared_ptr<PlanCache>& cache() const { return cache_; }

  // Runs n distance queries of one class (batch_key equal); results and
  // statuses per query.
  int run(const jsi_query* const* qs, int n, jsi_result* const* res, int* status);

  static bool eligible(const jsi_query& q) {
    return q.kind == JSI_QUERY_DISTANCE && q.params.cand_cap == 0 && q.params.k >= 1 && q.params.k <= 8192 &&
           q.params.dist_refine_nuv >= 2 && q.params.dist_refine_nuv <= 17 && q.r1.nuv >= 2 && q.r2.nuv >= 2;
  }
  std::string key(const jsi_query& q);

 private:
  template <class T>
  int run_t(const jsi_query* const* qs, int n, jsi_result* const* res, int* status);
  template <class T>
  int make_plan(const DistBatchShape& sh, int nq, std::unique_ptr<DistPlan>* out);
  template <class T>
  void fill_q(const DistPlan& pl, const QLayout& L, int q, DistBatchQ<T>* d);
  Tuning tuning_of(const jsi_query& q);

  jsi_ctx* c_;
  std::shared_ptr<PlanCache> cache_;
  bool graphs_ = true;
};

std::string DistBatcher::key(const jsi_query& q) {
  const jsi_params& P = q.params;
  const bool dbl = P.precision != JSI_PREC_EXACT32;
  int v1 = -1, v2 = -1, pv = -1;
  if (dbl) {
    SurfDev<double> a, b;
    if (c_->surf_dev<double>(q.h1, &a) != JSI_OK || c_->surf_dev<double>(q.h2, &b) != JSI_OK) return std::string();
    v1 = eval_variant<double>(a);
    v2 = eval_variant<double>(b);
    if (P.polish > 0) pv = polish_variant(a, b);
  } else {
    SurfDev<float> a, b;
    if (c_->surf_dev<float>(q.h1, &a) != JSI_OK || c_->surf_dev<float>(q.h2, &b) != JSI_OK) return std::string();
    v1 = eval_variant<float>(a);
    v2 = eval_variant<float>(b);
    if (P.polish > 0) {
      SurfDev<double> pa, pb;
      if (c_->surf_dev<double>(q.h1, &pa) != JSI_OK || c_->surf_dev<double>(q.h2, &pb) != JSI_OK)
        return std::string();
      pv = polish_variant(pa, pb);
    }
  }
  const Tuning tn = tuning_of(q);
  char buf[256];
  snprintf(buf, sizeof(buf), "p%d n%d,%d r%d R%d k%d c%d m%d v%d,%d,%d t%d,%d,%d,%d", P.precision, q.r1.nuv,
           q.r2.nuv, P.dist_refine_nuv, query_rounds(P, q.r1, q.r2), P.k, P.cull ? 1 : 0, std::min(P.k, P.polish), v1,
           v2, pv, tn.seed_pairs, tn.filter_walk, tn.r1_scale, tn.r1_seeds);
  return buf;
}

// The autotuned choices for a query: by the larger degree of its surfaces
// and its round-1 nuv.
Tuning DistBatcher::tuning_of(const jsi_query& q) {
  SurfDev<float> a, b;
  if (c_->surf_dev<float>(q.h1, &a) != JSI_OK || c_->surf_dev<float>(q.h2, &b) != JSI_OK) return Tuning();
  return c_->tuning_for(std::max(std::max(a.degu, a.degv), std::max(b.degu, b.degv)), q.r1.nuv);
}

template <class T>
int DistBatcher::make_plan(const DistBatchShape& sh, int nq, std::unique_ptr<DistPlan>* out) {
  auto pl = std::make_unique<DistPlan>();
  pl->sh = sh;
  pl->nq = nq;
  pl->dbl = std::is_same<T, double>::value;
  const QLayout L = layout_of(sh, sizeof(vec4_t<T>));
  pl->per_q = L.bytes;
  const int R = sh.rounds, m = std::max(0, std::min(sh.k, sh.polish));
  size_t o = (size_t)nq * L.bytes;
  const size_t out0 = o;
  const size_t o_fail = o;
  o = align256(o + sizeof(unsigned) * nq);
  const size_t o_last = o;
  o = align256(o + sizeof(Cand) * (size_t)nq * R);
  const size_t o_nu = o;
  o = align256(o + sizeof(int) * (size_t)nq * 2 * R);
  const size_t o_pol = o;
  o = align256(o + sizeof(PolishRec) * (size_t)nq * std::max(m, 1));
  pl->out_bytes = o - out0;
  const size_t o_sorted = o;
  o = align256(o + sizeof(Cand) * (size_t)R * nq * sh.k);
  pl->dev_bytes = o;
  if (!c_->pool->reserve(pl->dev_bytes)) {
    return JSI_ERR_NOMEM;
  }
  pl->pool = c_->pool.get();
  pl->reserved = pl->dev_bytes;
  if (hipMalloc(&pl->dev, pl->dev_bytes) != hipSuccess) {
    (void)hipGetLastError();
    pl->dev = nullptr;
    return JSI_ERR_NOMEM;
  }
  pl->fail = reinterpret_cast<unsigned*>(pl->dev + o_fail);
  pl->lastrec = reinterpret_cast<Cand*>(pl->dev + o_last);
  pl->nunits = reinterpret_cast<int*>(pl->dev + o_nu);
  pl->pol = reinterpret_cast<PolishRec*>(pl->dev + o_pol);
  pl->sorted = reinterpret_cast<Cand*>(pl->dev + o_sorted);
  if (hipHostMalloc(&pl->h_out, pl->out_bytes, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&pl->h_fail0), sizeof(unsigned) * nq, hipHostMallocDefault) != hipSuccess)
    return JSI_ERR_NOMEM;
  // first-appearance tables start at INT_MAX; every use restores them
  for (int q = 0; q < nq; ++q) {
    char* base = pl->dev + (size_t)q * L.bytes;
    const size_t n1 = std::max((size_t)(sh.nuv1 - 1) * (sh.nuv1 - 1), (size_t)sh.k * (sh.nref - 1) * (sh.nref - 1));
    const size_t n2 = std::max((size_t)(sh.nuv2 - 1) * (sh.nuv2 - 1), (size_t)sh.k * (sh.nref - 1) * (sh.nref - 1));
    if (launch_fill_i32(reinterpret_cast<int*>(base + L.fp1), INT_MAX, n1, c_->stream) != hipSuccess ||
        launch_fill_i32(reinterpret_cast<int*>(base + L.fp2), INT_MAX, n2, c_->stream) != hipSuccess)
      return JSI_ERR_DEVICE;
  }
  if (hipStreamSynchronize(c_->stream) != hipSuccess) return JSI_ERR_DEVICE;
  *out = std::move(pl);
  return JSI_OK;
}

template <class T>
void DistBatcher::fill_q(const DistPlan& pl, const QLayout& L, int q, DistBatchQ<T>* d) {
  using V = vec4_t<T>;
  char* b = pl.dev + (size_t)q * L.bytes;
  for (int s = 0; s < 2; ++s) {
    for (int p = 0; p < 2; ++p) {
      d->reg[s][p] = reinterpret_cast<V*>(b + L.reg[s][p]);
      d->su[s][p] = reinterpret_cast<int*>(b + L.su[s][p]);
      d->C[s][p] = reinterpret_cast<V*>(b + L.C[s][p]);
      d->tbl[s][p] = reinterpret_cast<V*>(b + L.tbl[s][p]);
      d->tbh[s][p] = reinterpret_cast<V*>(b + L.tbh[s][p]);
    }
    const bool s32 = pl.sh.prec == JSI_PREC_FP32SCREEN;
    d->c32[s] = s32 ? reinterpret_cast<float4*>(b + L.c32[s]) : nullptr;
    d->tb32l[s] = s32 ? reinterpret_cast<float4*>(b + L.tb32l[s]) : nullptr;
    d->tb32h[s] = s32 ? reinterpret_cast<float4*>(b + L.tb32h[s]) : nullptr;
  }
  d->amax = reinterpret_cast<unsigned long long*>(b + L.amax);
  d->list = reinterpret_cast<uint64_t*>(b + L.list);
  d->small = reinterpret_cast<unsigned*>(b + L.small);
  d->cand = reinterpret_cast<Cand*>(b + L.cand);
  d->topk = reinterpret_cast<Cand*>(b + L.topk);
  d->st = reinterpret_cast<SelState*>(b + L.st);
  d->hist = reinterpret_cast<unsigned*>(b + L.hist);
  d->range = reinterpret_cast<unsigned long long*>(b + L.range);
  d->seedwork = reinterpret_cast<unsigned long long*>(b + L.seedwork);
  d->best = reinterpret_cast<unsigned long long*>(b + L.best);
  d->fp1 = reinterpret_cast<int*>(b + L.fp1);
  d->fp2 = reinterpret_cast<int*>(b + L.fp2);
  d->dwork = reinterpret_cast<unsigned char*>(b + L.dwork);
  const int R = pl.sh.rounds, m = std::max(1, std::min(pl.sh.k, pl.sh.polish));
  d->lastrec = pl.lastrec + (size_t)q * R;
  d->nunits = pl.nunits + (size_t)q * 2 * R;
  for (int r = 0; r < JSI_MAX_ROUNDS; ++r)
    d->sorted[r] = r < R ? pl.sorted + ((size_t)r * pl.nq + q) * pl.sh.k : nullptr;
  d->pol = pl.pol + (size_t)q * m;
  d->fail = pl.fail + q;
}

// Device bytes a plan of one shape may hold; larger shapes run in smaller
// batches (the 4x-scaled nuv 1024 query holds ~250 MB, most of it the
// round-1 tile-pair list).
constexpr size_t kPlanBudget = size_t(8) << 30;

int DistBatcher::run(const jsi_query* const* qs, int n, jsi_result* const* res, int* status) {
  if (n <= 0) return JSI_OK;
  const size_t per = jsi_batch_query_bytes(*qs[0]);
  const int cap = per * 64 <= kPlanBudget ? 64 : (per * 16 <= kPlanBudget ? 16 : 4);
  if (n == 1 && !cache_->seen(key(*qs[0]) + " q4")) return JSI_ERR_UNSUPPORTED;  // alone: not worth a plan
  for (int a = 0; a < n; a += cap) {
    const int m = std::min(cap, n - a);
    const int rc = qs[0]->params.precision == JSI_PREC_EXACT32 ? run_t<float>(qs + a, m, res + a, status + a)
                                                               : run_t<double>(qs + a, m, res + a, status + a);
    if (rc != JSI_OK) return rc;
  }
  return JSI_OK;
}

template <class T>
int DistBatcher::run_t(const jsi_query* const* qs, int n, jsi_result* const* res, int* status) {
  const double t0 = now_ms();
  const jsi_query& q0 = *qs[0];
  const jsi_params& P = q0.params;
  DistBatchShape sh;
  sh.prec = P.precision;
  sh.nuv1 = q0.r1.nuv;
  sh.nuv2 = q0.r2.nuv;
  sh.nref = P.dist_refine_nuv;
  sh.rounds = query_rounds(P, q0.r1, q0.r2);
  sh.k = P.k;
  sh.cull = P.cull ? 1 : 0;
  sh.polish = P.polish;
  sh.small_cap = (unsigned)std::max(2048, std::min(16384, 2 * sh.k));
  sh.cand_cap = std::max<unsigned long long>(1ull << 19, 128ull * (unsigned long long)sh.k);
  sh.grid = c_->grid;
  sh.tune = tuning_of(q0);
  // plans hold 4, 16 or 64 queries; slots past n are marked failed from the
  // start, so every kernel skips them
  const int nq = n <= 4 ? 4 : (n <= 16 ? 16 : 64);
  if (n > 64) return JSI_ERR_INVALID;
  sh.nq = nq;
  const std::string pk = key(q0) + " q" + std::to_string(nq);
  std::unique_ptr<DistPlan> plp = cache_->checkout(pk);
  if (!plp) {
    const int rc = make_plan<T>(sh, nq, &plp);
    if (rc != JSI_OK) return rc;
  }
  // back to the cache when this batch is done, whatever happens
  struct Return {
    PlanCache* c;
    const std::string& k;
    std::unique_ptr<DistPlan>& p;
    ~Return() {
      if (p) c->checkin(k, std::move(p));
    }
  } ret{cache_.get(), pk, plp};
  DistPlan& pl = *plp;
  const QLayout L = layout_of(sh, sizeof(vec4_t<T>));
  std::vector<DistBatchQ<T>> dq(nq);
  for (int q = 0; q < nq; ++q) {
    const jsi_query& Q = *qs[q < n ? q : 0];
    DistBatchQ<T>& d = dq[q];
    std::memset(&d, 0, sizeof(d));
    int rc = c_->surf_dev<T>(Q.h1, &d.sd1);
    if (rc == JSI_OK) rc = c_->surf_dev<T>(Q.h2, &d.sd2);
    if (rc == JSI_OK && sh.polish > 0) rc = c_->surf_dev<double>(Q.h1, &d.p1);
    if (rc == JSI_OK && sh.polish > 0) rc = c_->surf_dev<double>(Q.h2, &d.p2);
    if (rc != JSI_OK) return rc;
    d.g1.x = T(Q.r1.ustart);
    d.g1.y = T(Q.r1.vstart);
    d.g1.z = T(Q.r1.ustop);
    d.g1.w = T(Q.r1.vstop);
    d.g2.x = T(Q.r2.ustart);
    d.g2.y = T(Q.r2.vstart);
    d.g2.z = T(Q.r2.ustop);
    d.g2.w = T(Q.r2.vstop);
    d.box1 = box_of(Q.r1);
    d.box2 = box_of(Q.r2);
    fill_q<T>(pl, L, q, &d);
  }
  BatchProgram prog;
  int rc = record_dist_batch<T>(prog, sh, dq);
  if (rc == JSI_OK) rc = record_polish_batch<T>(prog, sh, dq);
  if (rc != JSI_OK) return rc;
  if (!pl.d_arena) {
    pl.arena_bytes = prog.arena.size();
    if (hipMalloc(&pl.d_arena, std::max(pl.arena_bytes, (size_t)256)) != hipSuccess ||
        hipHostMalloc(&pl.h_arena, std::max(pl.arena_bytes, (size_t)256), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return JSI_ERR_NOMEM;
    }
  }
  if (prog.arena.size() != pl.arena_bytes) return JSI_ERR_INVALID;  // same shape, same image size
  std::memcpy(pl.h_arena, prog.arena.data(), pl.arena_bytes);
  for (int q = 0; q < nq; ++q) pl.h_fail0[q] = q < n ? 0u : 1u;
  hipStream_t s = c_->stream;
  hipError_t e = hipSuccess;
  const char* stage = "launch";
  if (graphs_ && !pl.exec && !pl.no_graph) {
    // the graph is built node by node (no stream capture, which other
    // threads' work on their own streams could disturb): argument image and
    // initial flags in, every launch in order, outputs out
    stage = "graph build";
    hipGraph_t g = nullptr;
    e = hipGraphCreate(&g, 0);
    hipGraphNode_t prev = nullptr, node = nullptr;
    auto dep = [&]() { return prev ? 1 : 0; };
    if (e == hipSuccess)
      e = hipGraphAddMemcpyNode1D(&node, g, nullptr, 0, pl.d_arena, pl.h_arena, pl.arena_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      prev = node;
      e = hipGraphAddMemcpyNode1D(&node, g, &prev, dep(), pl.fail, pl.h_fail0, sizeof(unsigned) * nq,
                                  hipMemcpyHostToDevice);
    }
    for (size_t li = 0; e == hipSuccess && li < prog.launches.size(); ++li) {
      prev = node;
      const BatchLaunch& l = prog.launches[li];
      const void* a = pl.d_arena + l.off;
      void* args[] = {&a};
      hipKernelNodeParams kp{};
      kp.func = const_cast<void*>(l.func);
      kp.gridDim = l.grid;
      kp.blockDim = l.block;
      kp.sharedMemBytes = 0;
      kp.kernelParams = args;
      kp.extra = nullptr;
      e = hipGraphAddKernelNode(&node, g, &prev, 1, &kp);
    }
    if (e == hipSuccess) {
      prev = node;
      e = hipGraphAddMemcpyNode1D(&node, g, &prev, 1, pl.h_out, pl.fail, pl.out_bytes, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess) {
      stage = "graph instantiate";
      e = hipGraphInstantiate(&pl.exec, g, nullptr, nullptr, 0);
    }
    if (e == hipSuccess) {
      pl.graph = g;
    } else {  // run this plan without a graph from now on
      fprintf(stderr, "[jsi] batch %s failed (%s); launching directly\n", stage, hipGetErrorString(e));
      (void)hipGetLastError();
      if (g) (void)hipGraphDestroy(g);
      pl.exec = nullptr;
      pl.no_graph = true;
      e = hipSuccess;
    }
  }
  if (graphs_ && pl.exec) {
    stage = "graph launch";
    e = hipGraphLaunch(pl.exec, s);
  } else {
    stage = "launch";
    e = hipMemcpyAsync(pl.d_arena, pl.h_arena, pl.arena_bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(pl.fail, pl.h_fail0, sizeof(unsigned) * nq, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = prog.run(pl.d_arena, s);
    if (e == hipSuccess) e = hipMemcpyAsync(pl.h_out, pl.fail, pl.out_bytes, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) {
    stage = "synchronize";
    e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    fprintf(stderr, "[jsi] batch %s failed: %s\n", stage, hipGetErrorString(e));
    (void)hipGetLastError();
    return JSI_ERR_DEVICE;
  }
  // results
  const int R = sh.rounds, k = sh.k, m = std::min(sh.k, std::max(0, sh.polish));
  const char* ho = static_cast<const char*>(pl.h_out);
  const size_t o_last = align256(sizeof(unsigned) * nq);
  const size_t o_nu = o_last + align256(sizeof(Cand) * (size_t)nq * R);
  const size_t o_pol = o_nu + align256(sizeof(int) * (size_t)nq * 2 * R);
  const unsigned* hfail = reinterpret_cast<const unsigned*>(ho);
  const Cand* hlast = reinterpret_cast<const Cand*>(ho + o_last);
  const int* hnu = reinterpret_cast<const int*>(ho + o_nu);
  const PolishRec* hpol = reinterpret_cast<const PolishRec*>(ho + o_pol);
  const double ms = now_ms() - t0;
  c_->shared_stat("batch_launches", (double)prog.launches.size());
  {
    int nfail = 0;
    for (int q = 0; q < n; ++q) {
      nfail += hfail[q] ? 1 : 0;
      if (hfail[q] && c_->debug_)
        fprintf(stderr, "[jsi] batch query %d rerun alone: %s%s%s\n", q, hfail[q] & 1 ? "candidate overflow " : "",
                hfail[q] & 2 ? "fewer than k pairs " : "", hfail[q] & 4 ? "ties beyond the compaction" : "");
    }
    c_->shared_stat("batches", (double)n);
    c_->shared_stat("batch_fallbacks", (double)nfail);
  }
  std::vector<Cand> lists;
  for (int q = 0; q < n; ++q) {
    jsi_result* r = res[q];
    const jsi_params& Pq = qs[q]->params;
    std::memset(r, 0, sizeof(*r));
    if (hfail[q]) {  // the single-query path repeats what overflowed
      status[q] = c_->min_distance<T>(qs[q]->h1, qs[q]->r1, qs[q]->h2, qs[q]->r2, Pq, r, nullptr, 0);
      continue;
    }
    if (Pq.trace) {
      lists.resize((size_t)R * k);
      for (int rr = 0; rr < R; ++rr)
        if (hipMemcpy(lists.data() + (size_t)rr * k, dq[q].sorted[rr], sizeof(Cand) * k, hipMemcpyDeviceToHost) !=
            hipSuccess)
          return JSI_ERR_DEVICE;
    }
    r->rounds = R;
    r->npairs = k;
    for (int rr = 0; rr < R; ++rr) {
      jsi_round_info& ri = r->round[rr];
      ri.npairs = k;
      ri.slots = rr == 0 ? 1 : k;
      ri.nuv = rr == 0 ? sh.nuv1 : sh.nref;
      ri.kth_d2 = d2_of<T>(hlast[(size_t)q * R + rr].hi);
      if (rr + 1 < R) {
        ri.units1 = hnu[(size_t)q * 2 * R + 2 * rr];
        ri.units2 = hnu[(size_t)q * 2 * R + 2 * rr + 1];
      }
      if (Pq.trace) {
        uint64_t h = 1469598103934665603ull;
        for (int e2 = 0; e2 < k; ++e2) {
          const Cand& cd = lists[(size_t)rr * k + e2];
          const double dd = d2_of<T>(cd.hi);
          const int32_t sl = (int32_t)(cd.lo >> 40), i = (int32_t)((cd.lo >> 20) & 0xfffff),
                        j = (int32_t)(cd.lo & 0xfffff);
          h = fnv_bytes(h, &dd, 8);
          h = fnv_bytes(h, &sl, 4);
          h = fnv_bytes(h, &i, 4);
          h = fnv_bytes(h, &j, 4);
        }
        ri.hash = h;
      }
    }
    const T kd = T(d2_of<T>(hlast[(size_t)q * R + R - 1].hi));
    r->distance = (double)std::sqrt(kd);
    if (m > 0) {
      const PolishRec* pr = hpol + (size_t)q * m;
      int best = 0;
      for (int e2 = 1; e2 < m; ++e2)
        if (pr[e2].f < pr[best].f) best = e2;
      jsi_closest& cl = r->closest;
      const PolishRec& b = pr[best];
      cl.distance = std::sqrt(b.f);
      cl.u1 = b.x[0];
      cl.v1 = b.x[1];
      cl.u2 = b.x[2];
      cl.v2 = b.x[3];
      for (int t = 0; t < 3; ++t) {
        cl.p1[t] = b.p1[t];
        cl.p2[t] = b.p2[t];
      }
      cl.candidate = best;
      cl.iterations = b.iters;
      cl.polished = m;
      cl.flags = b.flags;
    }
    r->ms_total = ms;
    status[q] = JSI_OK;
  }
  return JSI_OK;
}

}  // namespace jsi

void jsi::DistBatcherDelete::operator()(jsi::DistBatcher* b) const { delete b; }

// The context's batcher; a worker's shares its root context's plan cache.
static jsi::DistBatcher* batcher_of(jsi_ctx* c) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (!c->batcher) {
    jsi_ctx* r = c->root ? c->root : c;
    if (!r->batcher) r->batcher.reset(new jsi::DistBatcher(r, std::make_shared<jsi::PlanCache>()));
    if (r != c) c->batcher.reset(new jsi::DistBatcher(c, r->batcher->cache()));
  }
  return c->batcher.get();
}

void jsi_ctx_drop_batcher(jsi_ctx* c) { c->batcher.reset(); }

// Batched distance queries of one context: groups the eligible queries by
// shape and runs each group (up to 64 queries at a time) as one batch; other
// queries run one at a time.  Used by jsi_run_batch and the server.
int jsi_ctx_run_distance_batch(jsi_ctx* c, int n, const jsi_query* const* qs, jsi_result* const* res, int* status) {
  jsi::DistBatcher* const bat = batcher_of(c);
  if (hipSetDevice(c->device) != hipSuccess) return JSI_ERR_DEVICE;
  std::map<std::string, std::vector<int>> groups;
  for (int i = 0; i < n; ++i) {
    if (jsi::DistBatcher::eligible(*qs[i])) {
      const std::string k = bat->key(*qs[i]);
      if (!k.empty()) {
        groups[k].push_back(i);
        continue;
      }
    }
    std::memset(res[i], 0, sizeof(jsi_result));
    status[i] = qs[i]->kind == JSI_QUERY_DISTANCE
                    ? jsi_min_distance(c, qs[i]->h1, &qs[i]->r1, qs[i]->h2, &qs[i]->r2, &qs[i]->params, res[i],
                                       nullptr, 0)
                    : jsi_intersect(c, qs[i]->h1, &qs[i]->r1, qs[i]->h2, &qs[i]->r2, &qs[i]->params, res[i], nullptr,
                                    0);
  }
  int first = JSI_OK;
  for (auto& g : groups) {
    const std::vector<int>& idx = g.second;
    for (size_t a = 0; a < idx.size(); a += 64) {
      const int cnt = (int)std::min<size_t>(64, idx.size() - a);
      std::vector<const jsi_query*> bq(cnt);
      std::vector<jsi_result*> br(cnt);
      std::vector<int> bs(cnt, JSI_OK);
      for (int j = 0; j < cnt; ++j) {
        bq[j] = qs[idx[a + j]];
        br[j] = res[idx[a + j]];
      }
      const int rc = bat->run(bq.data(), cnt, br.data(), bs.data());
      for (int j = 0; j < cnt; ++j) {
        int st = rc == JSI_OK ? bs[j] : rc;
        if (rc != JSI_OK) {  // the batch could not run (e.g. memory): each query alone
          std::memset(br[j], 0, sizeof(jsi_result));
          st = jsi_min_distance(c, bq[j]->h1, &bq[j]->r1, bq[j]->h2, &bq[j]->r2, &bq[j]->params, br[j], nullptr, 0);
        }
        status[idx[a + j]] = st;
      }
    }
  }
  for (int i = 0; i < n; ++i)
    if (status[i] != JSI_OK && first == JSI_OK) first = status[i];
  return first;
}

std::string jsi_ctx_batch_key(jsi_ctx* c, const jsi_query& q) {
  if (!jsi::DistBatcher::eligible(q)) return std::string();
  if (std::getenv("JSI_BATCH") && std::strcmp(std::getenv("JSI_BATCH"), "0") == 0) return std::string();
  return batcher_of(c)->key(q);
}

size_t jsi_batch_query_bytes(const jsi_query& q) {
  jsi::DistBatchShape sh;
  sh.prec = q.params.precision;
  sh.nuv1 = q.r1.nuv;
  sh.nuv2 = q.r2.nuv;
  sh.nref = q.params.dist_refine_nuv;
  sh.k = q.params.k;
  sh.small_cap = (unsigned)std::max(2048, std::min(16384, 2 * sh.k));
  sh.cand_cap = std::max<unsigned long long>(1ull << 19, 128ull * (unsigned long long)sh.k);
  const size_t v = q.params.precision == JSI_PREC_EXACT32 ? sizeof(float4) : sizeof(double4);
  const size_t R = (size_t)jsi::query_rounds(q.params, q.r1, q.r2);
  return jsi::layout_of(sh, v).bytes + R * sh.k * sizeof(jsi::Cand) + (size_t)4096;
}
