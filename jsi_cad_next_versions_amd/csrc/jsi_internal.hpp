// jsi_internal.hpp - types and kernel launchers shared by the HIP kernels
// (jsi_kernels.hip) and the host runtime (jsi_runtime.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>

#include <tuple>
#include <vector>

#include "../../include/jsi_mi355x.h"

namespace jsi {

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};
template <class T>
using vec4_t = typename Vec4<T>::type;

// One distance candidate.  hi = bit pattern of d2 (non-negative, so the
// unsigned order of the bits is the numeric order), lo = slot<<40 | i<<20 | j.
// Lexicographic (hi, lo) order is the total order (d2, slot, i, j) of
// dist_pair_ref_less (distance_analysis.cu:226-231).
struct Cand {
  uint64_t hi;
  uint64_t lo;
};

// One intersection pair (IntersectPairRef, intersection.cu:248-252).
struct IPair {
  int32_t slot;
  int32_t i;
  int32_t j;
};

constexpr int kTile = 16;          // cells per tile edge in the screens
constexpr int kTileCells = kTile * kTile;
constexpr int kPairsSlots = 32;      // profiling counters of evaluated cell pairs (summed on the host)
constexpr int kSortBlock = 2048;   // records per LDS bitonic block
constexpr int kAmaxSlots = 64;     // slots of a max |coordinate| reduction (read as their max)

inline __host__ __device__ uint64_t pack_idx(uint32_t slot, uint32_t i, uint32_t j) {
  return (uint64_t(slot) << 40) | (uint64_t(i) << 20) | uint64_t(j);
}

// Geometry of one evaluation batch: `units` regions of the same surface
// sampled on the same nuv x nuv grid.
template <class T>
struct SurfDev {
  const T* U;   // knots u
  const T* V;   // knots v
  const T* CP;  // xyz, u fastest
  const T* W;   // weights or nullptr
  int degu, degv, ncpu, ncpv, nku, nkv;
  int rational;
  int zero;  // periodic / degree >= 25: all-zero output (evaluation.cu:271-272)
};

enum EvalMode { kEvalFull18 = 0, kEvalPoint = 1, kEvalPointK = 2 };

// ---- launchers (defined in jsi_kernels.hip) --------------------------------
template <class T>
hipError_t launch_promote(const float* src, T* dst, int n, hipStream_t s);

// Evaluates nunits x nuv x nuv points, one workgroup per 16 x 16-point tile
// (spans and basis rows computed in the kernel).  mode kEvalFull18: out = AoS
// 18 per point (evalWithCuda layout, unit-major).  kEvalPoint / kEvalPointK:
// out = SoA xyz per unit [unit][3][nuv*nuv]; kEvalPointK also max-reduces
// |Suu|,|Suv|,|Svv| into kmax[unit*3 + {0,1,2}] (ordered bit patterns,
// cleared by the caller).  steps_given: regions hold (ustart, vstart, ustep,
// vstep) instead of (ustart, vstart, ustop, vstop).
template <class T>
hipError_t launch_points_tile(const SurfDev<T>& sd, const vec4_t<T>* reg, int nunits, int nuv, int mode,
                              int steps_given, T* out, unsigned long long* kmax, hipStream_t s);

// Outputs of the fused evaluation + distance centres.  C as launch_aabb(isect
// = 0) writes it; optional: lo/hi, the per-tile box of the centres
// (launch_tilebox isect = 0); T = double only: absmax max-reduces the largest
// |coordinate| (double bits, kAmaxSlots partial maxima), C32 (requires
// absmax) the centres rounded to float and lo32/hi32 their tile boxes.
// nunits_dev (optional): the true unit count on the device, <= nunits.
template <class T>
struct EcOut {
  vec4_t<T>* C = nullptr;
  vec4_t<T>* lo = nullptr;
  vec4_t<T>* hi = nullptr;
  float4* C32 = nullptr;
  float4* lo32 = nullptr;
  float4* hi32 = nullptr;
  unsigned long long* absmax = nullptr;
  const int* nunits_dev = nullptr;
};
// One surface's regions for the fused evaluation + distance centres.
template <class T>
struct EcJob {
  SurfDev<T> sd;
  const vec4_t<T>* reg = nullptr;
  int nunits = 0, nuv = 0;
  EcOut<T> o;
};
// b may be null; two jobs of the same kernel variant share one launch.
template <class T>
hipError_t launch_eval_centres(const EcJob<T>& a, const EcJob<T>* b, hipStream_t s);

// Cell AABBs of every unit from its SoA points.  isect = 0: distance centres
// (construct_aabbs_distance_host); isect = 1: centre + padded radius
// (construct_aabbs_host) with the compute_K padding from kmax.
template <class T>
hipError_t launch_aabb(const T* pts, int nunits, int nuv, int isect, int use_k, const unsigned long long* kmax,
                       const vec4_t<T>* regions, T init_u, T init_v, vec4_t<T>* C, vec4_t<T>* R, hipStream_t s);

// Per-tile bounding boxes.  isect = 0: box of centres; isect = 1: box of the
// padded cells, widened outward so the tile test is conservative.
template <class T>
hipError_t launch_tilebox(const vec4_t<T>* C, const vec4_t<T>* R, int nunits, int w, int isect, vec4_t<T>* lo,
                          vec4_t<T>* hi, hipStream_t s);

// Distance screen.
template <class T>
hipError_t launch_dist_seed(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                            const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, int k,
                            unsigned long long* best, hipStream_t s);
template <class T>
hipError_t launch_dist_list(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                            const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, const T* thresh,
                            int cull, uint64_t* list, unsigned* list_count, hipStream_t s, int fma = 0);
// Seed threshold: picks up to nseed tile pairs from best[] (npos =
// nslots * tiles per side-1 unit; ordered = the first slots, one tile per
// unit, best[] unused), histograms the d2 of their cell pairs and writes the
// upper edge of the bin holding rank k to *thresh (+inf if they hold fewer
// than k pairs).
template <class T>
hipError_t launch_seed_thresh(const unsigned long long* best, int npos, int tpu1, int nseed, int ordered,
                              uint64_t* list, unsigned* list_count, unsigned long long* range,
                              unsigned long long* work, const int* su1, const int* su2, int w1, int w2,
                              const vec4_t<T>* C1, const vec4_t<T>* C2, int k, int fma, T* thresh, hipStream_t s,
                              unsigned* clear = nullptr, int nclear = 0);
size_t seed_work_bytes();
template <class T>
hipError_t launch_dist_filter(const uint64_t* list, const unsigned* list_count, unsigned list_cap, const int* su1,
                              const int* su2, int w1, int w2, const vec4_t<T>* C1, const vec4_t<T>* C2,
                              const T* thresh, Cand* out, unsigned long long* out_count, unsigned long long cap,
                              int grid, hipStream_t s, unsigned long long* pairs_done = nullptr, int fma = 0,
                              int cull = 1, const unsigned long long* absmax = nullptr, int screen32 = 0,
                              const vec4_t<T>* blo1 = nullptr, const vec4_t<T>* bhi1 = nullptr,
                              const vec4_t<T>* blo2 = nullptr, const vec4_t<T>* bhi2 = nullptr, int walk = 8);
// absmax (double only): the expanded-form test (three FMAs per pair with a
// rounding margin from the largest |centre coordinate|, exact keys).
// screen32 (with absmax, units of one tile): the FP32 matrix-core screen of
// the refinement rounds (FP32SCREEN), same records; blo1 .. bhi2: the units'
// centre boxes (one tile per unit).
template <class T>
hipError_t launch_set_thresh(T* thresh, const Cand* rec, int idx, int use_inf, hipStream_t s);

// FP32SCREEN helpers (the float centres and max |coordinate| come from the
// evaluation): widen a float threshold by the rounding bound, and replace
// candidate keys by their FP64 d2.
hipError_t launch_inflate(float* thresh, const unsigned long long* absmax_bits, hipStream_t s);
hipError_t launch_rerank(Cand* cand, const unsigned long long* n_dev, unsigned long long cap, const int* su1,
                         const int* su2, int cells1, int cells2, const double4* C1, const double4* C2, hipStream_t s);

// Operator ABI: flat AoS arrays -> padded w x w float4 grids; endpoint masks.
hipError_t launch_pack_cells(const float* c, const float* r, long long n, int w, float4* C, float4* R,
                             hipStream_t s);
hipError_t launch_pair_mask(int mode, const uint64_t* list, const unsigned long long* list_count,
                            const unsigned* list_count32, int w1, int w2, const float4* C1, const float4* R1,
                            const float4* C2, const float4* R2, const float* thresh, long long n1, long long n2,
                            int* m1, int* m2, int grid, hipStream_t s);

// Histogram radix selection of the k smallest candidate keys (see jsi_kernels.hip).
struct SelState {
  unsigned long long lo, hi, width, need, below, final_hi;
  int done;
  int pad_;
};
// Runs `iters` narrowing steps (resume: continue a previous selection on the
// same records instead of starting one), then compacts the records with key
// <= the current bound into out when out is given.  Converged when the
// compacted count is <= small_cap or the state's done flag is set.
// n_dev (optional): the record count on the device; n is then its upper bound.
hipError_t select_topk(const Cand* cand, unsigned long long n, unsigned long long k, unsigned long long maxkey,
                       SelState* st, unsigned* hist, Cand* out, unsigned long long* out_count, unsigned small_cap,
                       hipStream_t s, int iters = 2, bool resume = false, const unsigned long long* n_dev = nullptr);
hipError_t launch_thresh_from_state_ru(float* thresh, const SelState* st, hipStream_t s);
hipError_t launch_set_thresh_ru(float* thresh, const Cand* rec, int idx, hipStream_t s);
hipError_t launch_copy_cand(Cand* dst, const Cand* src, hipStream_t s);
template <class T>
hipError_t launch_query_init(vec4_t<T>* reg1, vec4_t<T> g1, vec4_t<T>* reg2, vec4_t<T> g2, int* su1, int* su2,
                             unsigned long long* amax, hipStream_t s);
template <class T>
hipError_t launch_thresh_from_state(T* thresh, const SelState* st, hipStream_t s);

// Sorting of candidate records (bitonic, n2 = power of two >= kSortBlock).
hipError_t launch_fill_sentinel(Cand* a, unsigned long long from, unsigned long long n2, hipStream_t s);
hipError_t sort_cands(Cand* a, unsigned long long n2, hipStream_t s);
// Sorted first `keep` of m (device count, at most m_cap) distinct records.
// last_out (optional) receives the last kept record (rank min(keep, m) - 1).
hipError_t rank_sort(const Cand* in, const unsigned long long* m_dev, unsigned m_cap, unsigned keep, Cand* out,
                     hipStream_t s, Cand* last_out = nullptr);

// Distance refinement: first-appearance dedup of the sorted top-k endpoints,
// child regions and the next round's slot -> unit maps (three passes).
template <class T>
hipError_t launch_dist_dedup(const Cand* rec, int n, int cells1, int cells2, const int* su1, const int* su2,
                             const vec4_t<T>* reg1, const vec4_t<T>* reg2, int parent_nuv1, int parent_nuv2,
                             int* firstpos1, int* firstpos2, int* nsu1, int* nsu2, vec4_t<T>* nreg1,
                             vec4_t<T>* nreg2, int* nunits, unsigned char* work, hipStream_t s,
                             unsigned long long* zero = nullptr, int nzero = 0);
size_t dedup_work_bytes();  // size of `work` for launch_dist_dedup

// Intersection screen.  launch_isect_pass masks (optional, mask_tp x 256 x 4
// words): the count pass (write = 0) keeps the hits of the first mask_tp
// listed tile pairs there and the write pass reads them.
template <class T>
hipError_t launch_isect_list(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                             const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, int cull,
                             uint64_t* list, unsigned long long* list_count, unsigned long long list_cap,
                             hipStream_t s);
template <class T>
hipError_t launch_isect_pass(int write, const uint64_t* list, const unsigned long long* list_count, const int* su1,
                             const int* su2, int w1, int w2, const vec4_t<T>* C1, const vec4_t<T>* R1,
                             const vec4_t<T>* C2, const vec4_t<T>* R2, unsigned* rowcount,
                             const unsigned long long* rowoff, unsigned* rowcursor, IPair* out, int grid,
                             hipStream_t s, unsigned long long* masks = nullptr, unsigned long long mask_tp = 0);
hipError_t launch_row_sort(const unsigned long long* rowoff, unsigned long long nrows, IPair* out, hipStream_t s);
// FP32SCREEN intersection: float records covering the FP64 (and, given Cv/Rv,
// the VERSION5 float) acceptance intervals of each cell; FP64 re-test of the
// float candidates and the order-keeping compaction of the accepted ones.
hipError_t launch_isect_boxes32(const double4* C, const double4* R, const float4* Cv, const float4* Rv,
                                unsigned long long n, float4* X, float4* Y, hipStream_t s);
hipError_t launch_isect_retest(const IPair* cand, unsigned long long n, const int* su1, const int* su2, int cells1,
                               int cells2, const double4* C1, const double4* R1, const double4* C2, const double4* R2,
                               unsigned* flag, hipStream_t s);
hipError_t launch_isect_scatter(const IPair* cand, const unsigned* flag, const unsigned long long* idx,
                                unsigned long long n, IPair* out, hipStream_t s);
// Long rows: segmented radix sort of the cell ids j (jsi_sort.hip); tmp of
// segmented_sort_tmp_bytes(n, nrows) bytes.
size_t segmented_sort_tmp_bytes(unsigned long long n, unsigned long long nrows);
hipError_t segmented_sort_pairs(IPair* pairs, unsigned long long n, const unsigned long long* rowoff,
                                unsigned long long nrows, void* tmp, size_t tmp_bytes, hipStream_t s);
hipError_t scan_u32_to_u64(const unsigned* in, unsigned long long* out, unsigned long long n, void* tmp,
                           size_t tmp_bytes, hipStream_t s);
// three scans of n entries in the same launches: in + y in_stride -> out + y
// out_stride (in_stride >= n, out_stride >= n + 1), y = 0..2; totals
// (optional): the three totals side by side
hipError_t scan3_u32_to_u64(const unsigned* in, size_t in_stride, unsigned long long* out, size_t out_stride,
                            unsigned long long n, void* tmp, size_t tmp_bytes, hipStream_t s,
                            unsigned long long* totals = nullptr);
// scratch of a scan of n entries over `arrays` arrays at once
size_t scan_tmp_bytes(unsigned long long n, int arrays = 1);
hipError_t launch_isect_dedup_flags(const IPair* pairs, unsigned long long n, int cells2, int* firstpos2,
                                    unsigned* flag1, unsigned* flag2, hipStream_t s);
template <class T>
hipError_t launch_isect_units(const IPair* pairs, unsigned long long n, int cells2, const int* su1, const int* su2,
                              const vec4_t<T>* reg1, const vec4_t<T>* reg2, int parent_nuv1, int parent_nuv2,
                              const unsigned* flag1, const unsigned* flag2, const unsigned long long* idx1,
                              const unsigned long long* idx2, int* firstpos2, int* nsu1, int* nsu2,
                              vec4_t<T>* nreg1, vec4_t<T>* nreg2, hipStream_t s);

// One side of an intersection refinement round whose units have nuv <=
// kIcMaxNuv (jsi_isect_cells.hip): evaluation, the compute_K padding, the
// padded cell AABBs and each unit's tile box in one pass, several units per
// wavefront.  lo / hi (one box per unit) may be null.
constexpr int kIcMaxNuv = 8;
template <class T>
struct IcJob {
  SurfDev<T> sd;
  const vec4_t<T>* reg = nullptr;
  int nunits = 0, nuv = 0;
  T init_u = T(1), init_v = T(1);
  vec4_t<T>* C = nullptr;
  vec4_t<T>* R = nullptr;
  vec4_t<T>* lo = nullptr;
  vec4_t<T>* hi = nullptr;
  // FP32SCREEN (T = double): each cell's float record (record32_axis) and
  // each unit's float box of them, the values k_isect_boxes32 and
  // launch_tilebox<float> would write
  float4* X = nullptr;
  float4* Y = nullptr;
  float4* flo = nullptr;
  float4* fhi = nullptr;
};
// Overlap screen of a round whose units have at most kIcMaxNuv^2 cells per
// side, per slot (jsi_isect_cells.hip).  Cull (rounds with many slots): with
// wcount, the number of slots whose unit boxes meet in each run of 64; with
// woff (the exclusive scan of wcount) and list, those slots in order.  Count,
// over n entries (slots 0..n-1, or list[0..n-1] when list is non-null): per
// entry its pairs, its non-empty rows and its distinct side-2 cells (without
// a list and with cull, slots whose unit boxes lo/hi are disjoint have none).
// Emit, given the exclusive scans of the three counts: the pairs in (slot, i,
// j) order and, with nsu1 non-null, the next round's units (per pair its
// units, per unit its region: cells of the parents reg1 / reg2 on pnuv1 /
// pnuv2 grids), numbered by first appearance.  masks (optional, n x w1^2
// words): the count pass writes each row's hit mask there and the emit pass
// reads it instead of repeating the tests.
// FP32SCREEN inputs of the slot screen (T = double): the float records of
// both sides (k_isect_boxes32) and their unit boxes; the float test picks the
// candidates, the FP64 test decides them.
struct SlotF32 {
  const float4 *X1 = nullptr, *Y1 = nullptr, *X2 = nullptr, *Y2 = nullptr;
  const float4 *lo1 = nullptr, *hi1 = nullptr, *lo2 = nullptr, *hi2 = nullptr;
};
template <class T>
hipError_t launch_isect_slot_cull(long long nslots, int w1, int w2, const int* su1, const int* su2,
                                  const vec4_t<T>* lo1, const vec4_t<T>* hi1, const vec4_t<T>* lo2,
                                  const vec4_t<T>* hi2, unsigned* wcount, const unsigned long long* woff, int* list,
                                  const SlotF32* f, hipStream_t s);
template <class T>
hipError_t launch_isect_slot_count(long long nslots, int w1, int w2, const int* su1, const int* su2,
                                   const vec4_t<T>* C1, const vec4_t<T>* R1, const vec4_t<T>* C2,
                                   const vec4_t<T>* R2, const vec4_t<T>* lo1, const vec4_t<T>* hi1,
                                   const vec4_t<T>* lo2, const vec4_t<T>* hi2, int cull, long long n,
                                   const int* list, unsigned* npair, unsigned* nrow, unsigned* ncol, const SlotF32* f,
                                   hipStream_t s, unsigned long long* masks = nullptr);
template <class T>
hipError_t launch_isect_slot_emit(long long nslots, int w1, int w2, const int* su1, const int* su2,
                                  const vec4_t<T>* C1, const vec4_t<T>* R1, const vec4_t<T>* C2, const vec4_t<T>* R2,
                                  const vec4_t<T>* lo1, const vec4_t<T>* hi1, const vec4_t<T>* lo2,
                                  const vec4_t<T>* hi2, int cull, long long n, const int* list,
                                  const unsigned long long* offp, const unsigned long long* offr,
                                  const unsigned long long* offc, IPair* out, const vec4_t<T>* reg1,
                                  const vec4_t<T>* reg2, int pnuv1, int pnuv2, int* nsu1, int* nsu2,
                                  vec4_t<T>* nreg1, vec4_t<T>* nreg2, const SlotF32* f, hipStream_t s,
                                  const unsigned long long* masks = nullptr);
// b may be null; two jobs of the same kernel variant share one launch.
template <class T>
hipError_t launch_isect_cells(const IcJob<T>& a, const IcJob<T>* b, int use_k, hipStream_t s);

// FP64 solves on the final cells (jsi_refine.hip).  x = (u1, v1, u2, v2).
struct PolishRec {  // closest-point polish of one top-k record
  double x[4];
  double f;         // |S1 - S2|^2 at x
  double p1[3], p2[3];
  int iters, flags; // JSI_POLISH_*
};
struct CurveRec {   // projection of one overlap pair onto the intersection curve
  double x[4];
  double p[3];      // S1(u1, v1)
  double t[3];      // (S1u x S1v) x (S2u x S2v)
  double r2;        // |S1 - S2|^2 at x
  int iters, flags;
};
// box = (ulo, vlo, uhi, vhi) of the query region; TR = scalar of the unit regions.
template <class TR>
hipError_t launch_polish(const SurfDev<double>& s1, const SurfDev<double>& s2, const Cand* rec, int m, const int* su1,
                         const int* su2, const vec4_t<TR>* reg1, const vec4_t<TR>* reg2, int nuv1, int nuv2,
                         double4 box1, double4 box2, PolishRec* out, hipStream_t s);
template <class TR>
hipError_t launch_curve_proj(const SurfDev<double>& s1, const SurfDev<double>& s2, const IPair* pairs, long long n,
                             const int* su1, const int* su2, const vec4_t<TR>* reg1, const vec4_t<TR>* reg2, int nuv1,
                             int nuv2, double4 box1, double4 box2, CurveRec* out, hipStream_t s);
// Curve tracing from the projections (jsi_refine.hip).
struct MarchSeed {
  double x[4];
  double p[3];
  double T[3];  // unit tangent
  double r2;
};
struct MarchPt {
  double x[4];
  double p[3];
  double r2;
};
struct MarchOut {
  int succ;  // seed reached, -1 none
  int n;     // points recorded
  int flags; // JSI_MARCH_*
  int pad;
};
constexpr int JSI_MARCH_SEED = 1, JSI_MARCH_EDGE = 2, JSI_MARCH_FAILED = 4, JSI_MARCH_CAP = 8;
// Seeds: accepted projections thinned to spacing h; returns h.
double curve_seeds(const CurveRec* recs, long long n, double tol, double spacing, std::vector<MarchSeed>& seeds);
// Marches from seeds[starts[k]] along dir * T with step h (counting pass: pts
// = nullptr; writing pass: points at pts + offs[k]).
hipError_t launch_march(const SurfDev<double>& s1, const SurfDev<double>& s2, const MarchSeed* seeds, int nseeds,
                        const int* starts, int nstart, double dir, double h, double4 box1, double4 box2,
                        const long long* offs, MarchPt* pts, MarchOut* out, hipStream_t s);
// Chains the marches into ordered curves.
void assemble_curves(const std::vector<MarchSeed>& seeds, const std::vector<MarchOut>& fwd,
                     const std::vector<long long>& fwd_off, const std::vector<MarchPt>& fwd_pts,
                     const std::vector<int>& heads, const std::vector<MarchOut>& bwd,
                     const std::vector<long long>& bwd_off, const std::vector<MarchPt>& bwd_pts,
                     std::vector<jsi_curve_point>& out, int* ncurves);

hipError_t launch_fill_i32(int* p, int value, unsigned long long n, hipStream_t s);
hipError_t launch_fill_u32(unsigned* p, unsigned value, unsigned long long n, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched distance queries.  A batch is a set of queries of one shape (same
// precision, grids, k, rounds, polish); every kernel of the query runs once
// for the whole batch with blockIdx.y = query, reading that query's arguments
// from a device array.  A query whose screen overflowed its candidate buffer,
// or whose selection could not bound its top-k in two narrowing steps, sets
// its failure flag; every later kernel of that query returns at once and the
// host runs the query again on the single-query path.
// ---------------------------------------------------------------------------
template <class... A>
struct BArgs {  // one query's arguments of a batched kernel
  std::tuple<A...> t;
  const unsigned* fail;
};

// Launches recorded for a batch and the host image of their arguments; the
// image is copied to `dev` (arena) before the launches run, so the same
// program runs again (or replays as a graph) after the image is rewritten.
// Every recorded kernel takes one argument: a pointer into the arena.
struct BatchLaunch {
  const void* func;  // the kernel (host-side handle)
  dim3 grid, block;
  size_t off;        // its argument array in the arena
};
struct BatchProgram {
  std::vector<unsigned char> arena;
  std::vector<BatchLaunch> launches;
  template <class X>
  size_t put(const std::vector<X>& v) {
    size_t off = (arena.size() + 63) & ~(size_t)63;
    arena.resize(off + sizeof(X) * v.size());
    std::memcpy(arena.data() + off, v.data(), sizeof(X) * v.size());
    return off;
  }
  void add(const void* func, dim3 grid, dim3 block, size_t off) { launches.push_back({func, grid, block, off}); }
  hipError_t run(const unsigned char* dev, hipStream_t s) const {
    for (const BatchLaunch& l : launches) {
      const void* a = dev + l.off;
      void* args[] = {&a};
      const hipError_t e = hipLaunchKernel(l.func, l.grid, l.block, args, 0, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
};

// Launch and seed choices of the distance screens, picked per (degree, nuv)
// by the autotuner (autotune.py) and set through jsi_tune_set; the defaults
// are the hand-set values.  None of them changes a result.
struct Tuning {
  int seed_pairs = 32;  // refinement rounds: seed cell pairs per unit of k (from the first slots)
  int filter_walk = 8;  // refinement filter: slots a workgroup walks (1..16)
  int r1_scale = 100;   // round-1 screen: workgroups, percent of the default grid
  int r1_seeds = 256;   // round 1: seed tile pairs (the best side-1 tiles' best partners)
};

struct DistBatchShape {
  int prec = 0;                       // JSI_PREC_*
  int nuv1 = 0, nuv2 = 0;             // round-1 grids
  int nref = 16;                      // refinement grid
  int rounds = 4;
  int k = 8;
  int cull = 1;
  int polish = 0;                     // records polished (0 = none)
  unsigned long long cand_cap = 0;    // candidate records per query and screen
  unsigned small_cap = 0;             // records compacted by the selection
  int grid = 0;                       // resident workgroups of the chip
  int nq = 0;                         // queries
  Tuning tune;                        // launch and seed choices
};

// Per-query inputs and device buffers of a distance batch.  [s][p]: side s,
// ping-pong buffer p (round r reads p = (r - 1) % 2).
template <class T>
struct DistBatchQ {
  SurfDev<T> sd1, sd2;
  SurfDev<double> p1, p2;  // FP64 surfaces of the polish
  vec4_t<T> g1, g2;        // round-1 regions (ustart, vstart, ustop, vstop)
  double4 box1, box2;      // parameter boxes of the polish
  vec4_t<T>* reg[2][2];
  int* su[2][2];
  vec4_t<T>* C[2][2];
  vec4_t<T>* tbl[2][2];
  vec4_t<T>* tbh[2][2];
  float4* c32[2];          // FP32SCREEN round 1: centres rounded to float, their tile boxes
  float4* tb32l[2];
  float4* tb32h[2];
  unsigned long long* amax;
  Cand* lastrec;           // [rounds]
  int* nunits;             // [2 rounds]
  uint64_t* list;
  unsigned* small;         // counters (list count, candidate count, threshold, compacted count)
  Cand* cand;
  Cand* topk;
  SelState* st;
  unsigned* hist;
  unsigned long long* range;
  unsigned long long* seedwork;
  unsigned long long* best;
  Cand* sorted[JSI_MAX_ROUNDS];  // each round's ordered top-k (k records)
  int* fp1;
  int* fp2;
  unsigned char* dwork;
  PolishRec* pol;
  unsigned* fail;
};

// Records every launch of a distance batch (all rounds, not the polish).
template <class T>
int record_dist_batch(BatchProgram& prog, const DistBatchShape& sh, const std::vector<DistBatchQ<T>>& qs);
// Records the polish of every query of a batch (jsi_refine.hip).
template <class T>
int record_polish_batch(BatchProgram& prog, const DistBatchShape& sh, const std::vector<DistBatchQ<T>>& qs);
// Kernel variants a query's launches depend on (batches group equal keys).
template <class T>
int eval_variant(const SurfDev<T>& sd);
int polish_variant(const SurfDev<double>& s1, const SurfDev<double>& s2);

}  // namespace jsi
