// jsi_serve.hip - batch runner and serving loop (include/jsi_mi355x.h).
//
// Worker contexts share their parent's resident surfaces and own a stream and
// a set of device buffers each, so the kernels of the queries in flight overlap
// on the device.  A query's result never depends on what else runs: every
// kernel is deterministic and each worker has private buffers.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/jsi_mi355x.h"
#include "jsi_runtime.hpp"

namespace {

struct Job {
  int64_t ticket = 0;
  jsi_query q;
  jsi_result res;
  int status = JSI_OK;
  bool done = false;
  size_t est_bytes = 0;
  std::string key;  // shape of a batchable distance query ("" = runs alone)
};

// Device bytes a query may hold while it runs: a batched distance query its
// share of the batch's buffers (cells of both ping-pong rounds, tile lists,
// candidates, first-appearance tables); otherwise grid points, cells, tile
// lists, the 64 MiB candidate buffer and, for intersections, pair lists.
size_t estimate_bytes(const jsi_query& q, bool batched) {
  if (batched) return jsi_batch_query_bytes(q);
  const size_t n1 = (size_t)q.r1.nuv * q.r1.nuv, n2 = (size_t)q.r2.nuv * q.r2.nuv;
  const size_t scal = q.params.precision == JSI_PREC_EXACT32 ? 4 : 8;
  size_t b = (n1 + n2) * scal * 3 * 2 + (n1 + n2) * scal * 8 + ((size_t)64 << 20);
  if (q.kind == JSI_QUERY_INTERSECT) b += (size_t)512 << 20;  // pair lists, row hit masks
  return b;
}

constexpr int kMaxBatch = 64;  // distance queries a worker takes in one batch

}  // namespace



struct jsi_server {
  jsi_ctx* parent = nullptr;
  std::vector<jsi_ctx*> workers;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Job>> jobs;
  int64_t next_ticket = 1;
  size_t hbm_cap = 0, inflight = 0;
  int max_pending = 0;
  bool stop = false;
  int64_t n_admitted = 0, n_refused = 0, n_completed = 0, n_hits = 0, n_evictions = 0;
  // LRU cache of resident surfaces keyed by the caller
  struct CacheEnt {
    int handle;
    size_t bytes;
    uint64_t last;
    int refs;
  };
  std::unordered_map<uint64_t, CacheEnt> cache;
  std::unordered_map<int, uint64_t> handle_key;
  uint64_t clock = 0;
  size_t cache_bytes = 0, cache_cap = 0;

  void touch_refs(const jsi_query& q, int delta) {
    for (int h : {q.h1, q.h2}) {
      auto it = handle_key.find(h);
      if (it != handle_key.end()) cache[it->second].refs += delta;
    }
  }

  // Continuous batching: a free worker takes the oldest request and, when it
  // is a batchable distance query, every queued request of the same shape
  // (up to kMaxBatch), so the requests that arrived while the workers were
  // busy run as one batch on the next free worker.
  void worker_loop(int w) {
    jsi_ctx* c = workers[w];
    (void)hipSetDevice(c->device);
    while (true) {
      std::vector<std::shared_ptr<Job>> take;
      {
        std::unique_lock<std::mutex> lock(mu);
        cv_work.wait(lock, [&] { return stop || !queue.empty(); });
        if (stop && queue.empty()) return;
        take.push_back(queue.front());
        queue.pop_front();
        const std::string& key = take[0]->key;
        if (!key.empty())
          for (auto it = queue.begin(); it != queue.end() && (int)take.size() < kMaxBatch;) {
            if ((*it)->key == key) {
              take.push_back(*it);
              it = queue.erase(it);
            } else {
              ++it;
            }
          }
      }
      if (take.size() == 1 && take[0]->key.empty()) {
        Job* job = take[0].get();
        job->status = job->q.kind == JSI_QUERY_DISTANCE
                          ? jsi_min_distance(c, job->q.h1, &job->q.r1, job->q.h2, &job->q.r2, &job->q.params,
                                             &job->res, nullptr, 0)
                          : jsi_intersect(c, job->q.h1, &job->q.r1, job->q.h2, &job->q.r2, &job->q.params, &job->res,
                                          nullptr, 0);
      } else {
        std::vector<const jsi_query*> qs;
        std::vector<jsi_result*> rs;
        std::vector<int> st(take.size(), JSI_OK);
        for (auto& j : take) {
          qs.push_back(&j->q);
          rs.push_back(&j->res);
        }
        (void)jsi_ctx_run_distance_batch(c, (int)take.size(), qs.data(), rs.data(), st.data());
        for (size_t i = 0; i < take.size(); ++i) take[i]->status = st[i];
      }
      {
        std::lock_guard<std::mutex> lock(mu);
        for (auto& job : take) {
          job->done = true;
          inflight -= job->est_bytes;
          touch_refs(job->q, -1);
          ++n_completed;
        }
      }
      cv_done.notify_all();
    }
  }
};

extern "C" {

int jsi_server_create(jsi_ctx* ctx, int slots, size_t hbm_cap_bytes, int max_pending, size_t cache_cap_bytes,
                      jsi_server** out) {
  if (!ctx || !out || slots < 1 || slots > 64) return JSI_ERR_INVALID;
  auto* s = new jsi_server();
  s->parent = ctx;
  // admission: the caller's cap, or what the context's device memory account
  // has left when the server starts
  {
    const size_t used = ctx->pool->used.load(), cap = ctx->pool->cap;
    s->hbm_cap = hbm_cap_bytes ? hbm_cap_bytes : (cap > used ? cap - used : 0);
  }
  s->max_pending = max_pending > 0 ? max_pending : 1 << 20;
  s->cache_cap = cache_cap_bytes;
  for (int w = 0; w < slots; ++w) {
    jsi_ctx* c = nullptr;
    const int rc = jsi_ctx_create_worker(ctx, &c);
    if (rc != JSI_OK) {
      jsi_server_destroy(s);
      return rc;
    }
    s->workers.push_back(c);
  }
  for (int w = 0; w < slots; ++w) s->threads.emplace_back([s, w] { s->worker_loop(w); });
  *out = s;
  return JSI_OK;
}

int jsi_server_destroy(jsi_server* s) {
  if (!s) return JSI_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->threads) t.join();
  for (jsi_ctx* c : s->workers) jsi_ctx_destroy(c);
  for (auto& kv : s->cache) jsi_surface_release(s->parent, kv.second.handle);
  delete s;
  return JSI_OK;
}

int jsi_server_surface(jsi_server* s, uint64_t key, const int ivec[7], const float* data, int ndata, int* handle) {
  if (!s || !handle) return JSI_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->cache.find(key);
  if (it != s->cache.end()) {
    it->second.last = ++s->clock;
    *handle = it->second.handle;
    ++s->n_hits;
    return JSI_OK;
  }
  const size_t bytes = sizeof(float) * (size_t)ndata;
  // evict least-recently-used surfaces no query in flight uses
  while (s->cache_cap && s->cache_bytes + bytes > s->cache_cap) {
    auto victim = s->cache.end();
    for (auto e = s->cache.begin(); e != s->cache.end(); ++e)
      if (e->second.refs == 0 && (victim == s->cache.end() || e->second.last < victim->second.last)) victim = e;
    if (victim == s->cache.end()) break;
    jsi_surface_release(s->parent, victim->second.handle);
    s->handle_key.erase(victim->second.handle);
    s->cache_bytes -= victim->second.bytes;
    s->cache.erase(victim);
    ++s->n_evictions;
  }
  int h = -1;
  const int rc = jsi_surface_upload(s->parent, ivec, data, ndata, &h);
  if (rc != JSI_OK) return rc;
  s->cache[key] = {h, bytes, ++s->clock, 0};
  s->handle_key[h] = key;
  s->cache_bytes += bytes;
  *handle = h;
  return JSI_OK;
}

static std::shared_ptr<Job> make_job(jsi_server* s, const jsi_query* q) {
  auto job = std::make_shared<Job>();
  job->q = *q;
  std::memset(&job->res, 0, sizeof(job->res));
  job->key = jsi_ctx_batch_key(s->parent, *q);
  job->est_bytes = estimate_bytes(*q, !job->key.empty());
  return job;
}

// Admission and queueing (caller holds s->mu).
static int enqueue_locked(jsi_server* s, const std::shared_ptr<Job>& job, int64_t* ticket) {
  // a request larger than the whole cap is still admitted when nothing else runs
  if ((int)s->queue.size() >= s->max_pending ||
      (s->hbm_cap && s->inflight > 0 && s->inflight + job->est_bytes > s->hbm_cap)) {
    ++s->n_refused;
    return JSI_ERR_BUSY;
  }
  job->ticket = s->next_ticket++;
  s->inflight += job->est_bytes;
  s->touch_refs(job->q, +1);
  s->jobs[job->ticket] = job;
  s->queue.push_back(job);
  ++s->n_admitted;
  *ticket = job->ticket;
  return JSI_OK;
}

int jsi_server_submit(jsi_server* s, const jsi_query* q, int64_t* ticket) {
  if (!s || !q || !ticket) return JSI_ERR_INVALID;
  auto job = make_job(s, q);
  {
    std::lock_guard<std::mutex> lock(s->mu);
    const int rc = enqueue_locked(s, job, ticket);
    if (rc != JSI_OK) return rc;
  }
  s->cv_work.notify_one();
  return JSI_OK;
}

int jsi_server_poll(jsi_server* s, int64_t ticket, jsi_result* result, int wait) {
  if (!s) return JSI_ERR_INVALID;
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) return JSI_ERR_INVALID;
  std::shared_ptr<Job> job = it->second;
  if (!job->done) {
    if (!wait) return 0;
    s->cv_done.wait(lock, [&] { return job->done; });
  }
  s->jobs.erase(ticket);
  if (result) *result = job->res;
  return job->status == JSI_OK ? 1 : job->status;
}

int jsi_server_counters(jsi_server* s, int64_t* admitted, int64_t* refused, int64_t* completed, int64_t* hits,
                        int64_t* evictions) {
  if (!s) return JSI_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  if (admitted) *admitted = s->n_admitted;
  if (refused) *refused = s->n_refused;
  if (completed) *completed = s->n_completed;
  if (hits) *hits = s->n_hits;
  if (evictions) *evictions = s->n_evictions;
  return JSI_OK;
}

int jsi_run_batch(jsi_ctx* ctx, int n, const jsi_query* queries, jsi_result* results, int* status,
                  int concurrency) {
  if (!ctx || n < 0 || (n > 0 && (!queries || !results)) || concurrency < 1) return JSI_ERR_INVALID;
  jsi_server* s = nullptr;
  int rc = jsi_server_create(ctx, concurrency, 0, 0, 0, &s);
  if (rc != JSI_OK) return rc;
  std::vector<int64_t> tickets(n);
  {
    // every query is queued before a worker takes any, so equal shapes batch
    std::vector<std::shared_ptr<Job>> jobs;
    for (int i = 0; i < n; ++i) jobs.push_back(make_job(s, &queries[i]));
    std::lock_guard<std::mutex> lock(s->mu);
    const size_t cap = s->hbm_cap;
    s->hbm_cap = 0;  // a batch run admits everything it was given
    for (int i = 0; i < n && rc == JSI_OK; ++i) rc = enqueue_locked(s, jobs[i], &tickets[i]);
    s->hbm_cap = cap;
  }
  s->cv_work.notify_all();
  if (rc != JSI_OK) {
    jsi_server_destroy(s);
    return rc;
  }
  int first = JSI_OK;
  for (int i = 0; i < n; ++i) {
    const int r = jsi_server_poll(s, tickets[i], &results[i], 1);
    const int st = r == 1 ? JSI_OK : r;
    if (status) status[i] = st;
    if (st != JSI_OK && first == JSI_OK) first = st;
  }
  jsi_server_destroy(s);
  return first;
}

}  // extern "C"
