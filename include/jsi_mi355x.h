/*
 * jsi_mi355x.h - C ABI of the MI355X minimum-distance / surface-surface
 * intersection library (libjsi_mi355x.so).
 *
 * Every entry point returns an int status (JSI_OK == 0, negative on error) and
 * never exits the process.  Pointers are plain host pointers unless the name
 * says "device".  No torch or HIP types appear in these signatures.
 *
 * Surfaces use the JSI_CAD_NEXT NurbsFace layout
 * (JSI_CAD_NEXT_VERSION5/include/nurbs_data.hpp:1-10):
 *   ivec[7] = {flags, degU, nCPu, nKu, degV, nCPv, nKv}
 *   data    = U[nKu] | V[nKv] | CP xyz[3*nCPu*nCPv] (u fastest) | w[nCPu*nCPv] if flags&2
 * (layout read by BuildEvalContext, JSI_CAD_NEXT_VERSION5/src/evaluation/evaluation.cu:239-257).
 */
#ifndef JSI_MI355X_H
#define JSI_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define JSI_OK 0
#define JSI_ERR_INVALID (-1)     /* bad argument, handle or surface */
#define JSI_ERR_NOMEM (-2)       /* device arena or HBM cap exhausted */
#define JSI_ERR_DEVICE (-3)      /* HIP runtime error */
#define JSI_ERR_CAPACITY (-4)    /* an internal bounded buffer overflowed */
#define JSI_ERR_UNSUPPORTED (-5) /* periodic surfaces, degree >= 25, nuv > 1025 ... */
#define JSI_ERR_BUSY (-6)        /* serving queue full */

/* ---- precision modes ---------------------------------------------------- */
/* EXACT32: FP32 with the reference's operation order and no contraction;
 *          bit-exact with VERSION5 (which computes in float throughout).
 * FP64:    the same recipe in double ("plain FP64").
 * FP32SCREEN: FP64 evaluation, screening in FP32 with a rounding margin that
 *          makes the candidate set a superset of the FP64 one, then FP64
 *          re-ranking / re-testing; results equal FP64 mode.  Distance: round
 *          1 on float copies of the centres, rounds 2.. on the FP32 matrix
 *          cores relative to per-slot origins.  Intersection: float records
 *          covering the FP64 acceptance intervals (round 1: also VERSION5's
 *          float ones), so the candidates include VERSION5's round-1 pairs.   */
#define JSI_PREC_EXACT32 0
#define JSI_PREC_FP64 1
#define JSI_PREC_FP32SCREEN 2

#define JSI_MAX_ROUNDS 16
#define JSI_MAX_NUV 1025 /* cell index must fit 20 bits */
#define JSI_MAX_K 8192

/* Region of one surface, the value part of EvalAndConstructTask
 * (JSI_CAD_NEXT_VERSION5/include/jsi_cad.hpp:11-23). side is 1 or 2. */
typedef struct jsi_region {
  float ustart, vstart, ustop, vstop;
  int nuv;
  int side;
} jsi_region;

/* Query parameters.  jsi_params_default() fills the VERSION5 values
 * (distance_analysis.cu:33-36, intersection.cu:36,368-370, jsi_cad.hpp:35-36). */
typedef struct jsi_params {
  int k;                 /* distance top-k (k_of_topk = 8) */
  int dist_refine_nuv;   /* nuv of distance refinement units (next_nuv = 16) */
  int dist_max_rounds;   /* distance screening rounds (max_layers = 4) */
  int isect_max_rounds;  /* intersection screening rounds (6) */
  int isect_refine_nuv;  /* starting intersection refinement nuv (static global_nuv = 16) */
  int isect_halve_above; /* halve refinement nuv when pairs > this (512) ... */
  int isect_min_nuv;     /* ... and nuv > this (4) */
  int use_compute_k;     /* second-derivative AABB padding (CAD_INTERSECT_USE_K, default on) */
  int precision;         /* JSI_PREC_* */
  int cull;              /* conservative tile culling inside the screens (results unchanged) */
  int trace;             /* fill jsi_round_info.hash (costs a device->host copy per round) */
  int cand_cap;          /* distance screen candidate capacity, 0 = default (4 Mi records), at least
                            2k is used; a smaller value exercises the overflow path (threshold
                            tightening and a repeated screen), results unchanged */
  int polish;            /* distance: FP64 closest-point polish of the first `polish` records of the
                            final top-k list (0 = off, at most JSI_MAX_K); fills jsi_result.closest */
  int screen_out;        /* intersection, FP32SCREEN only: 1 = pairs_out and jsi_result.npairs hold the
                            float screen's candidates of the last round instead of its FP64-tested
                            pairs (a check of the screen's superset property; rounds unchanged) */
  double param_tol;      /* distance, FP64 / FP32SCREEN only: > 0 replaces dist_max_rounds by the
                            number of rounds after which a cell spans <= param_tol in u and v on both
                            surfaces (at most JSI_MAX_ROUNDS); 0 = off */
} jsi_params;

typedef struct jsi_round_info {
  int64_t npairs;  /* distance: size of the top-k list; intersection: overlap pairs */
  int64_t slots;   /* aligned slots screened in this round */
  int64_t units1;  /* unique side-1 refinement regions made from this round's pairs */
  int64_t units2;  /* unique side-2 refinement regions */
  double kth_d2;   /* distance: largest d2 of the round's top-k (k-th smallest) */
  uint64_t hash;   /* FNV-1a 64 over the ordered pair list (trace only) */
  int nuv;         /* sampling nuv of the regions screened in this round */
  int pad_;
} jsi_round_info;

/* FP64 closest points of the two surfaces (jsi_params.polish > 0).  Each
 * polished record starts at the parameter centres of its two cells and runs
 * a damped Newton iteration on |S1(u1,v1) - S2(u2,v2)|^2 inside the query
 * regions; the record with the smallest result wins (lowest rank on ties).
 * No counterpart in the reference, whose result is the k-th centre distance. */
typedef struct jsi_closest {
  double distance;    /* |S1(u1,v1) - S2(u2,v2)| */
  double u1, v1, u2, v2;
  double p1[3];       /* S1(u1, v1) */
  double p2[3];       /* S2(u2, v2) */
  int32_t candidate;  /* rank of the winning record in the final top-k list */
  int32_t iterations; /* iterations of the winner (accepted and rejected steps) */
  int32_t polished;   /* records polished (0 = polish off) */
  int32_t flags;      /* JSI_POLISH_* termination of the winner */
} jsi_closest;
#define JSI_POLISH_STATIONARY 1 /* gradient exactly zero */
#define JSI_POLISH_STALLED 2    /* damping exceeded its cap without a decrease */
#define JSI_POLISH_FIXED 4      /* the projected step left the parameters unchanged */
#define JSI_POLISH_CONVERGED 8  /* Newton step <= 1e-15 (1 + max |parameter|), or a step
                                   <= 1e-9 (1 + max |parameter|) that does not decrease f */

typedef struct jsi_result {
  double distance;       /* distance: sqrt(k-th smallest centre d2) of the last round */
  int64_t npairs;        /* pairs of the last round */
  int rounds;            /* rounds screened */
  int isect_refine_nuv;  /* intersection: refinement nuv after the query (halving state) */
  double ms_total;       /* host wall clock for the whole query */
  double ms_eval;        /* evaluation + AABB construction (measured in profiling mode */
  double ms_screen;      /* screening + top-k selection     or with CAD_TIMING_BREAKDOWN set) */
  jsi_round_info round[JSI_MAX_ROUNDS];
  jsi_closest closest;   /* distance with jsi_params.polish > 0 */
  /* intersection with use_compute_k: compute_K of the two initial regions
   * (side 1, side 2) and their patch scale max(|du|/U0, |dv|/V0), the values
   * VERSION5 prints as "Compute_K initial side%d raw=%g scale=%g padding=%g"
   * (intersection.cu:187-205); 0 otherwise */
  double isect_k_raw[2];
  double isect_k_scale[2];
} jsi_result;

/* One candidate pair: slot and cell ids as in DistPairRef / IntersectPairRef
 * (distance_analysis.cu:190-195, intersection.cu:248-252). d2 = 0 for
 * intersection pairs. */
typedef struct jsi_pair {
  double d2;
  int32_t slot;
  int32_t i;
  int32_t j;
  int32_t pad_;
} jsi_pair;

typedef struct jsi_ctx jsi_ctx;

void jsi_params_default(jsi_params* p);
const char* jsi_status_string(int status);
const char* jsi_last_error(void);

/* Context: one per device.  hbm_cap_bytes bounds the device arena the
 * context may hold (0 = 3/4 of free HBM).  Replaces init_cad/deinit_cad
 * (JSI_CAD_NEXT_VERSION5/src/common/common.cu:35-61). */
int jsi_ctx_create(int device, size_t hbm_cap_bytes, jsi_ctx** out);
int jsi_ctx_destroy(jsi_ctx* ctx);
int jsi_ctx_sync(jsi_ctx* ctx);
/* Per-kernel timing (HIP events on the context stream) of the screen
 * kernels: enable = 1 clears the counters and starts recording.  Names:
 * "dist_filter_f32", "dist_filter_f64".  units = cell pairs evaluated.
 * Always kept, for the context and every batch runner / server made from it:
 * "batches" (launches = batches run, units = queries in them) and
 * "batch_fallbacks" (units = batched queries rerun alone after a bounded
 * buffer of the batch overflowed). */
int jsi_ctx_profile(jsi_ctx* ctx, int enable);
int jsi_ctx_kernel_stats(jsi_ctx* ctx, const char* name, int64_t* launches, double* ms, double* units);

/* Copies a NurbsFace (host pointers) to device memory; the surface stays
 * resident until released.  *handle >= 0 on success. */
int jsi_surface_upload(jsi_ctx* ctx, const int ivec[7], const float* data, int ndata, int* handle);
/* Registers a surface whose ivec/data already live in device memory (the
 * EvalAndConstructTask::d_ivec/d_data convention of VERSION5). */
int jsi_surface_register_device(jsi_ctx* ctx, const int* d_ivec, const float* d_data, int ndata, int* handle);
int jsi_surface_release(jsi_ctx* ctx, int handle);

/* Grid evaluation, the evalWithCuda contract
 * (JSI_CAD_NEXT_VERSION5/src/evaluation/evaluation.cu:391-405 and :358-374):
 * point (idx, idy) at u = ustep*idx + ustart, v = vstep*idy + vstart, written
 * to out[(idx + idy*nuv)*18 + c], c = [S, Su, Sv, Suu, Suv, Svv] x xyz.
 * out is host memory.  _f32 runs EXACT32, _f64 the FP64 recipe. */
int jsi_eval_grid_f32(jsi_ctx* ctx, int handle, float ustep, float vstep, float ustart, float vstart, int nuv,
                      float* out);
int jsi_eval_grid_f64(jsi_ctx* ctx, int handle, double ustep, double vstep, double ustart, double vstart,
                      int nuv, double* out);

/* Minimum distance, the run_dist path: minimum_distance
 * (JSI_CAD_NEXT_VERSION5/src/distance_analysis/distance_analysis.cu:356-506).
 * pairs_out (optional, host) receives the last round's sorted top-k. */
int jsi_min_distance(jsi_ctx* ctx, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                     const jsi_params* params, jsi_result* result, jsi_pair* pairs_out, int pairs_cap);

/* jsi_min_distance for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and errors as
 * jsi_min_distance; no device memory is allocated: every buffer of the query
 * comes from a 256 MiB scratch array that the library holds on each GPU it
 * runs on until the process ends (queries up to nuv 512 with k 4096 fit; a
 * larger one fails with JSI_ERR_NOMEM and a message naming the limit, before
 * or while it runs).  Calls on one GPU share that array: calls that overlap
 * on one GPU race. */
int jsi_min_distance_noalloc(jsi_ctx* ctx, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                             const jsi_params* params, jsi_result* result, jsi_pair* pairs_out, int pairs_cap);

/* Autotuned launch and seed choices of the distance screens, per (degree,
 * nuv): a query uses the entry of the larger degree of its surfaces with the
 * nearest round-1 nuv (the defaults when its degree has none).  knob:
 * "seed_pairs" (refinement seed cell pairs per unit of k, 1..256, default
 * 32), "filter_walk" (slots a refinement-filter workgroup walks, 1..16,
 * default 8), "r1_scale" (round-1 screen workgroups, percent of the default
 * grid, 10..1000), "r1_seeds" (round-1 seed tile pairs whose exact k-th d2
 * bounds the threshold, 1..256, default 256).  None of them changes a result.  The table is shared by a
 * context and its server workers.  No counterpart in the reference (which
 * has fixed launch shapes); autotune.py writes the table. */
int jsi_tune_set(jsi_ctx* ctx, int degree, int nuv, const char* knob, int value);
int jsi_tune_clear(jsi_ctx* ctx); /* back to the defaults everywhere */

/* Surface-surface intersection screening, the run_intersect path:
 * intersection (JSI_CAD_NEXT_VERSION5/src/intersection/intersection.cu:338-475).
 * pairs_out (optional, host) receives the last round's ordered pairs. */
int jsi_intersect(jsi_ctx* ctx, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                  const jsi_params* params, jsi_result* result, jsi_pair* pairs_out, int pairs_cap);

/* One sample of an intersection curve (jsi_intersect_curves). */
typedef struct jsi_curve_point {
  double x, y, z;        /* S1(u1, v1) */
  double u1, v1, u2, v2; /* parameters on the two surfaces */
  double residual;       /* |S1(u1,v1) - S2(u2,v2)| */
  int32_t curve;         /* curve id, 0-based, in order of creation */
  int32_t flags;         /* JSI_CURVE_CLOSED when the curve is closed */
} jsi_curve_point;
#define JSI_CURVE_CLOSED 1

/* Intersection curves: the run_intersect rounds of jsi_intersect, then every
 * pair of the last round is projected onto the intersection curve in FP64
 * (Gauss-Newton on S1(u1,v1) - S2(u2,v2) = 0 from the centres of its two
 * cells, inside the query regions); projections with residual <= tol
 * (0 = 1e-9) are thinned to a spacing of `spacing` in 3-space (<= 0 = 1e-3 of
 * their bounding-box diagonal), linked within 3 x spacing and ordered along
 * each curve by its tangent.  pts (host) receives min(cap, *npts) samples,
 * curve by curve in walk order; *npts and *ncurves receive the totals.  No
 * counterpart in the reference, whose intersection() only prints pair counts
 * (JSI_CAD_NEXT_VERSION5/src/intersection/intersection.cu:367). */
int jsi_intersect_curves(jsi_ctx* ctx, int h1, const jsi_region* r1, int h2, const jsi_region* r2,
                         const jsi_params* params, double spacing, double tol, jsi_result* result,
                         jsi_curve_point* pts, int cap, int* npts, int* ncurves);

/* ---- batches and serving ------------------------------------------------ */
#define JSI_QUERY_DISTANCE 0
#define JSI_QUERY_INTERSECT 1

/* One query of a batch or a server: surfaces are handles of the context the
 * batch / server was created on. */
typedef struct jsi_query {
  int kind; /* JSI_QUERY_* */
  int h1, h2;
  jsi_region r1, r2;
  jsi_params params;
} jsi_query;

/* Runs n independent queries with up to `concurrency` of them in flight, each
 * on its own stream (their kernels overlap on the device).  results[i] is
 * identical to running queries[i] alone; status[i] (optional) receives each
 * query's status.  Returns the first non-OK status, or JSI_OK.  No
 * counterpart in the reference (VERSION5 runs one query at a time). */
int jsi_run_batch(jsi_ctx* ctx, int n, const jsi_query* queries, jsi_result* results, int* status,
                  int concurrency);

/* Serving loop: `slots` worker streams take admitted queries as they free up
 * (continuous batching: a free worker takes the oldest request and every
 * queued distance request of the same shape, up to 64, and runs them as one
 * batch); submissions are refused with JSI_ERR_BUSY when the estimated device
 * memory of the queries in flight would exceed hbm_cap_bytes (0 = what the
 * context's device memory cap leaves free when the server starts; a request
 * is always admitted when nothing is in flight) or max_pending requests wait.
 * Surfaces can be cached by a caller key with LRU eviction beyond
 * cache_cap_bytes. */
typedef struct jsi_server jsi_server;
int jsi_server_create(jsi_ctx* ctx, int slots, size_t hbm_cap_bytes, int max_pending, size_t cache_cap_bytes,
                      jsi_server** out);
int jsi_server_destroy(jsi_server* server);
/* cached upload: same key -> same resident surface (LRU touch) */
int jsi_server_surface(jsi_server* server, uint64_t key, const int ivec[7], const float* data, int ndata,
                       int* handle);
int jsi_server_submit(jsi_server* server, const jsi_query* query, int64_t* ticket);
/* 1 = done (result filled), 0 = still pending (wait = 0), < 0 = the query's error status */
int jsi_server_poll(jsi_server* server, int64_t ticket, jsi_result* result, int wait);
/* counters: admitted, refused (busy), completed, cache hits, cache evictions */
int jsi_server_counters(jsi_server* server, int64_t* admitted, int64_t* refused, int64_t* completed,
                        int64_t* cache_hits, int64_t* cache_evictions);

#ifdef __cplusplus
}
#endif

#endif /* JSI_MI355X_H */
