// jsi_kernels.hip - hand-written gfx950 kernels of the minimum-distance /
// surface-surface-intersection hot path.
//
// Every arithmetic expression that produces a result the reference also
// produces keeps the reference's operation order; the library is compiled
// with -ffp-contract=off, so float kernels are bit-exact with VERSION5 and
// double kernels with the FP64 oracle.  File:line citations refer to
// JSI_CAD_NEXT_VERSION5/.
#include <algorithm>
#include <type_traits>
#include <float.h>
#include <limits.h>

#include "jsi_device.hpp"
#include "jsi_internal.hpp"
#include "jsi_nurbs.hpp"

namespace jsi {

namespace {

constexpr int kMaxDeg = 25;  // MAXDEG, evaluation.cu:13

__device__ __forceinline__ unsigned ord32_up(float x) { return __float_as_uint(x); }
__device__ __forceinline__ unsigned ord32_up(double x) { return __float_as_uint(__double2float_ru(x)); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

inline unsigned blocks_for(unsigned long long n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// Knot span and basis functions (FindSpan / DersBasisFuns, evaluation.cu:152-219;
// The NURBS Book A2.1 and A2.3): find_span and basis_ders are in jsi_nurbs.hpp.
// ---------------------------------------------------------------------------
// Values row only (ders[0] of basis_ders): the same recurrence, kept in
// registers.  N[r] holds ndu[r][j-1] while stage j runs, and each stage
// performs exactly the operations basis_ders performs on ndu, so the results
// are bit-identical.  Loops are unrolled up to MAXP and cut at p.
template <class T, int MAXP>
__device__ __forceinline__ void basis_vals(int span, int p, T u, const T* U, T (&N)[MAXP + 1]) {
  T left[MAXP + 1], right[MAXP + 1];
  constexpr int UR = MAXP <= 7 ? MAXP + 1 : 1;  // unroll low degrees only
  N[0] = T(1);
#pragma unroll UR
  for (int j = 1; j <= MAXP; ++j) {
    if (j > p) break;
    left[j] = u - U[span + 1 - j];
    right[j] = U[span + j] - u;
    T saved = T(0);
#pragma unroll UR
    for (int r = 0; r < j; ++r) {
      const T den = right[r + 1] + left[j - r];
      const T temp = N[r] / den;
      N[r] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    N[j] = saved;
  }
}

// Fused evaluation + distance AABB centres.  One wavefront per 16x16-cell
// tile of a unit (a refinement unit is one tile).  The tile's 17 corner
// columns and 17 corner rows get their spans and basis values (A2.1, A2.3
// order 0: the values row ders[0] does not depend on the derivative order);
// EvaluateOneUV's S row (evaluation.cu:281-300 / :313-348) is then formed in
// its two contractions: the u-sum `row` of control row r for corner column e
// depends on (e, r) only, so it is computed once per (e, r) into LDS and
// shared by every corner of that column, and each corner accumulates its q+1
// rows in the reference order, followed by the 1/W scaling of the quotient
// rule; each cell's four corners reduce to its centre
// (construct_aabbs_distance_host, distance_analysis.cu:45-86).  The operations
// on every value are those of the reference, so results are bit-identical.
// Control rows are processed in chunks of kRows (a tile normally needs one).
// KL: the knot vectors are staged in LDS (chosen on the host from their
// length), so the span search reads LDS instead of chaining HBM loads.
// A single wavefront keeps ~8-12 KB of LDS, so many tiles are resident per
// CU and their memory latencies overlap.  Double only: with absmax the
// largest |centre coordinate| is max-reduced into absmax[block % kAmaxSlots]
// (rounding margins of the screens); with C32 the centres are also written
// rounded to float (FP32 screening).
static_assert(kAmaxSlots == 64, "the filter reduces the slots across one wavefront");
template <int MAXD>
struct EcRows {  // control rows per chunk of u-contractions
  static constexpr int value = MAXD <= 3 ? 8 : 16;
};

template <class T, int MAXD, bool RAT, bool KL>
__device__ __forceinline__ void ec_tile(const SurfDev<T>& sd, const vec4_t<T>* regions, int nunits, int nuv,
                                        const EcOut<T>& out, const int bid) {
  vec4_t<T>* const C = out.C;
  float4* const C32 = out.C32;
  unsigned long long* const absmax = out.absmax;
  const int* const nunits_dev = out.nunits_dev;
  constexpr int E = kTile + 1;  // corner points per tile edge
  constexpr int kRows = EcRows<MAXD>::value;
  constexpr int kCorners = (E * E + 63) / 64;  // corners per lane
  constexpr int kK = KL ? 2 * kEcKnots : 1;
  constexpr int kPts = 3 * E * E;
  // LDS strides are padded so that the lanes of a wavefront, which read
  // different corner columns, fall on different banks
  constexpr int kNb = MAXD + 2;           // basis row stride
  constexpr int kRs = 4 * kRows + 1;      // u-contraction column stride
  // one buffer, three lives: knots (phases 0-1), u-contractions (phases
  // 2-3), corner points x[], y[], z[] (phases 4-5); barriers separate them
  constexpr int kRsz = E * kRs;
  constexpr int kA = kK > kPts ? (kK > kRsz ? kK : kRsz) : (kPts > kRsz ? kPts : kRsz);
  __shared__ T sNu[E][kNb], sNv[E][kNb];
  __shared__ int sSu[E], sSv[E];
  __shared__ T sA[kA];
  T* const sKu = sA;
  T* const sKv = sA + kEcKnots;
  T(*const sR)[kRs] = reinterpret_cast<T(*)[kRs]>(sA);  // [column][4 * row + coordinate]
  T* const sP = sA;
  const int lane = threadIdx.x;
  const int w = nuv - 1;
  const int tw = (w + kTile - 1) / kTile, tpu = tw * tw;
  const int unit = bid / tpu, t = bid % tpu;
  if (unit >= nunits || (nunits_dev && unit >= *nunits_dev)) return;  // whole wavefront
  const int row0 = (t / tw) * kTile, col0 = (t % tw) * kTile;
  const int ht = min(kTile, w - row0), wd = min(kTile, w - col0);  // cells
  // the tile's (ht + 1) x (wd + 1) corners, numbered row by row: x = ey cw +
  // ex (a 15 x 15-cell unit has 256, four per lane)
  const int cw = wd + 1, ncorn = (ht + 1) * cw;
  const float rcw = 1.f / (float)cw;  // ey = floor((x + 0.5) / cw), exact for x < 2^12
  const int p = sd.degu, q = sd.degv;
  const vec4_t<T> r = regions[unit];
  // phase 0: knot vectors in LDS
  if (KL && !sd.zero) {
    for (int x = lane; x < sd.nku; x += 64) sKu[x] = sd.U[x];
    for (int x = lane; x < sd.nkv; x += 64) sKv[x] = sd.V[x];
  }
  __syncthreads();
  // phase 1: spans and basis values of the corner columns and rows
  if (!sd.zero)
    for (int task = lane; task < 2 * E; task += 64) {
      const int dir = task >= E;
      const int e = dir ? task - E : task;
      if (e > (dir ? ht : wd)) continue;
      const T start = dir ? r.y : r.x;
      const T step = ((dir ? r.w : r.z) - start) / T(nuv);
      const int idx = (dir ? row0 : col0) + e;
      const T tt = step * T(idx) + start;
      const int deg = dir ? q : p;
      const T* knots = KL ? (dir ? sKv : sKu) : (dir ? sd.V : sd.U);
      const int span = find_span(dir ? sd.nkv : sd.nku, deg, tt, knots);
      T nb[MAXD + 1];
      basis_vals<T, MAXD>(span, deg, tt, knots, nb);
      constexpr int UB = MAXD <= 7 ? MAXD + 1 : 1;
#pragma unroll UB
      for (int j = 0; j <= MAXD; ++j)
        if (j <= deg) (dir ? sNv : sNu)[e][j] = nb[j];
      (dir ? sSv : sSu)[e] = span;
    }
  __syncthreads();
  // control rows the tile reads: [vlo - q, vhi] (a wavefront reduction over
  // the corner rows' spans)
  int vlo = (!sd.zero && lane <= ht) ? sSv[lane] : INT_MAX;
  int vhi = (!sd.zero && lane <= ht) ? sSv[lane] : INT_MIN;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    vlo = min(vlo, __shfl_xor(vlo, off));
    vhi = max(vhi, __shfl_xor(vhi, off));
  }
  const int v0 = vlo - q, nrows = sd.zero ? 0 : vhi - vlo + q + 1;
  T acc[kCorners][4];
#pragma unroll
  for (int c = 0; c < kCorners; ++c) acc[c][0] = acc[c][1] = acc[c][2] = acc[c][3] = T(0);
  for (int r0 = 0; r0 < nrows; r0 += kRows) {
    const int nr = min(kRows, nrows - r0);
    // phase 2: u-contractions of rows [r0, r0 + nr) for every corner column
    for (int task = lane; task < (wd + 1) * nr; task += 64) {
      const int ex = task / nr, rr = task % nr;
      const int su = sSu[ex];
      const int base = sd.ncpu * (v0 + r0 + rr) + su - p;
      T row[4] = {T(0), T(0), T(0), T(0)};
      constexpr int UR = MAXD <= 7 ? MAXD + 1 : 1;
#pragma unroll UR
      for (int j = 0; j <= MAXD; ++j) {
        if (j > p) break;
        const int id = base + j;
        const T n = sNu[ex][j];
        if (RAT) {
          const T nw = n * sd.W[id];
          row[0] += nw * sd.CP[3 * id + 0];
          row[1] += nw * sd.CP[3 * id + 1];
          row[2] += nw * sd.CP[3 * id + 2];
          row[3] += nw;
        } else {
          row[0] += n * sd.CP[3 * id + 0];
          row[1] += n * sd.CP[3 * id + 1];
          row[2] += n * sd.CP[3 * id + 2];
        }
      }
      sR[ex][4 * rr + 0] = row[0];
      sR[ex][4 * rr + 1] = row[1];
      sR[ex][4 * rr + 2] = row[2];
      sR[ex][4 * rr + 3] = row[3];
    }
    __syncthreads();
    // phase 3: each corner adds the rows of this chunk, in s2 order; when one
    // chunk holds every row the tile reads (nrows <= kRows, r0 = 0), a
    // corner's rows are sv - vlo + s2, all present, and need no check
    const bool one = nrows <= kRows;
#pragma unroll
    for (int c = 0; c < kCorners; ++c) {
      if (64 * c >= ncorn) break;  // uniform
      const int x = lane + 64 * c;
      const int ey = (int)(((float)x + 0.5f) * rcw), ex = x - ey * cw;
      if (x >= ncorn) continue;
      const int sv = sSv[ey];
      constexpr int UR = MAXD <= 7 ? MAXD + 1 : 1;
      if (one) {
        const int rb = sv - vlo;
#pragma unroll UR
        for (int s2 = 0; s2 <= MAXD; ++s2) {
          if (s2 > q) break;
          const T nvs = sNv[ey][s2];
          acc[c][0] += nvs * sR[ex][4 * (rb + s2) + 0];
          acc[c][1] += nvs * sR[ex][4 * (rb + s2) + 1];
          acc[c][2] += nvs * sR[ex][4 * (rb + s2) + 2];
          if (RAT) acc[c][3] += nvs * sR[ex][4 * (rb + s2) + 3];
        }
      } else {
#pragma unroll UR
        for (int s2 = 0; s2 <= MAXD; ++s2) {
          if (s2 > q) break;
          const int rr = sv - q + s2 - v0 - r0;
          if (rr < 0 || rr >= nr) continue;
          const T nvs = sNv[ey][s2];
          acc[c][0] += nvs * sR[ex][4 * rr + 0];
          acc[c][1] += nvs * sR[ex][4 * rr + 1];
          acc[c][2] += nvs * sR[ex][4 * rr + 2];
          if (RAT) acc[c][3] += nvs * sR[ex][4 * rr + 3];
        }
      }
    }
    __syncthreads();  // sR is rewritten by the next chunk (or becomes sP)
  }
  // phase 4: corner points into sP (reuses the knot storage)
#pragma unroll
  for (int c = 0; c < kCorners; ++c) {
    if (64 * c >= ncorn) break;  // uniform
    const int x = lane + 64 * c;
    if (x >= ncorn) continue;
    T S[3] = {T(0), T(0), T(0)};
    if (!sd.zero) {
      if (RAT) {
        const T F = T(1) / acc[c][3];
        S[0] = acc[c][0] * F;
        S[1] = acc[c][1] * F;
        S[2] = acc[c][2] * F;
      } else {
        S[0] = acc[c][0];
        S[1] = acc[c][1];
        S[2] = acc[c][2];
      }
    }
    sP[x] = S[0];
    sP[kPts / 3 + x] = S[1];
    sP[2 * (kPts / 3) + x] = S[2];
  }
  __syncthreads();
  // phase 5: cell centres (and the tile's box of them)
  T am = T(0);
  const T inf = infinity_v<T>();
  T bmn[3] = {inf, inf, inf}, bmx[3] = {-inf, -inf, -inf};
  float fmn[3] = {__uint_as_float(0x7f800000u), __uint_as_float(0x7f800000u), __uint_as_float(0x7f800000u)};
  float fmx[3] = {-fmn[0], -fmn[0], -fmn[0]};
  for (int cidx = lane; cidx < kTileCells; cidx += 64) {
    const int ly = cidx / kTile, lx = cidx % kTile;
    if (ly >= ht || lx >= wd) continue;
    const int p0 = ly * cw + lx, p1 = p0 + 1, p2 = p0 + cw, p3 = p2 + 1;
    vec4_t<T> ctr;
    T* c = &ctr.x;
    const T k = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const T* pa = sP + a * (kPts / 3);
      const T a0 = pa[p0], a1 = pa[p1], a2 = pa[p2], a3 = pa[p3];
      const T lo = smin(a0, smin(a1, smin(a2, a3)));
      const T hi = smax(a0, smax(a1, smax(a2, a3)));
      c[a] = (lo - k + hi + k) / T(2);
    }
    ctr.w = T(0);
    const size_t o = (size_t)unit * w * w + (row0 + ly) * w + col0 + lx;
    C[o] = ctr;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      bmn[a] = fmin(bmn[a], c[a]);
      bmx[a] = fmax(bmx[a], c[a]);
    }
    if constexpr (std::is_same<T, double>::value) {
      if (C32) {
        float4 f;
        f.x = __double2float_rn(ctr.x);
        f.y = __double2float_rn(ctr.y);
        f.z = __double2float_rn(ctr.z);
        f.w = 0.f;
        C32[o] = f;
        fmn[0] = fminf(fmn[0], f.x);
        fmn[1] = fminf(fmn[1], f.y);
        fmn[2] = fminf(fmn[2], f.z);
        fmx[0] = fmaxf(fmx[0], f.x);
        fmx[1] = fmaxf(fmx[1], f.y);
        fmx[2] = fmaxf(fmx[2], f.z);
      }
      am = fmax(am, fmax(fabs(ctr.x), fmax(fabs(ctr.y), fabs(ctr.z))));
    }
  }
  if (out.lo) {  // uniform: a kernel argument
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        bmn[a] = fmin(bmn[a], __shfl_xor(bmn[a], off));
        bmx[a] = fmax(bmx[a], __shfl_xor(bmx[a], off));
      }
    if (lane == 0) {
      vec4_t<T> l, h;
      l.x = bmn[0];
      l.y = bmn[1];
      l.z = bmn[2];
      l.w = T(0);
      h.x = bmx[0];
      h.y = bmx[1];
      h.z = bmx[2];
      h.w = T(0);
      out.lo[bid] = l;  // bid = unit * tiles per unit + tile, the tile box index
      out.hi[bid] = h;
    }
  }
  if (out.lo32) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        fmn[a] = fminf(fmn[a], __shfl_xor(fmn[a], off));
        fmx[a] = fmaxf(fmx[a], __shfl_xor(fmx[a], off));
      }
    if (lane == 0) {
      out.lo32[bid] = make_float4(fmn[0], fmn[1], fmn[2], 0.f);
      out.hi32[bid] = make_float4(fmx[0], fmx[1], fmx[2], 0.f);
    }
  }
  if constexpr (std::is_same<T, double>::value) {
    if (absmax) {  // uniform: a kernel argument
      unsigned long long b = (unsigned long long)__double_as_longlong(am);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off);
        b = o > b ? o : b;
      }
      if (lane == 0 && b) atomicMax(&absmax[bid % kAmaxSlots], b);
    }
  }
}

// Both surfaces of a query in one launch when they share a variant: blocks
// [0, nb0) are tiles of j0, the rest tiles of j1.  Twice the wavefronts of
// one surface's launch keep more loads in flight per CU.
template <class T, int MAXD, bool RAT, bool KL>
__global__ __launch_bounds__(64) void k_eval_centres(EcJob<T> j0, EcJob<T> j1, int nb0) {
  const int b = (int)blockIdx.x;
  const bool second = b >= nb0;  // uniform: one body, its arguments selected
  const SurfDev<T> sd = second ? j1.sd : j0.sd;
  const EcOut<T> o = second ? j1.o : j0.o;
  ec_tile<T, MAXD, RAT, KL>(sd, second ? j1.reg : j0.reg, second ? j1.nunits : j0.nunits, second ? j1.nuv : j0.nuv, o,
                            second ? b - nb0 : b);
}

// Cell AABBs (construct_aabbs_distance_host, distance_analysis.cu:45-86;
// construct_aabbs_host + compute_K padding, intersection.cu:106-148,187-206,
// common.cu:170-171).  Cell id = iv*(nuv-1) + iu, corners (iv,iu), (iv,iu+1),
// (iv+1,iu), (iv+1,iu+1).
template <class T, bool ISECT>
__global__ void k_aabb(const T* pts, int nunits, int nuv, int use_k, const unsigned long long* kmax,
                       const vec4_t<T>* regions, T init_u, T init_v, vec4_t<T>* C, vec4_t<T>* R) {
  const int w = nuv - 1, cells = w * w, npts = nuv * nuv;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)nunits * cells) return;
  const int unit = (int)(gid / cells), cell = (int)(gid % cells);
  const int iv = cell / w, iu = cell % w;
  const T* base = pts + (size_t)unit * 3 * npts;
  const int p0 = iv * nuv + iu, p1 = p0 + 1, p2 = p0 + nuv, p3 = p2 + 1;
  T lo[3], hi[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T* b = base + c * npts;
    const T a0 = b[p0], a1 = b[p1], a2 = b[p2], a3 = b[p3];
    lo[c] = smin(a0, smin(a1, smin(a2, a3)));
    hi[c] = smax(a0, smax(a1, smax(a2, a3)));
  }
  vec4_t<T> ctr, rad;
  if (!ISECT) {
    const T k = T(0);
    ctr.x = (lo[0] - k + hi[0] + k) / T(2);
    ctr.y = (lo[1] - k + hi[1] + k) / T(2);
    ctr.z = (lo[2] - k + hi[2] + k) / T(2);
    ctr.w = T(0);
    C[gid] = ctr;
    return;
  }
  T pad = T(0);
  if (use_k) {
    const T muu = from_bits<T>(kmax[unit * 3 + 0]);
    const T muv = from_bits<T>(kmax[unit * 3 + 1]);
    const T mvv = from_bits<T>(kmax[unit * 3 + 2]);
    const int side = nuv > 1 ? nuv - 1 : 1;
    const T s2 = T(side * side);
    const T raw = (muu / s2 + T(2) * muv / s2 + mvv / s2) / T(8);
    const vec4_t<T> r = regions[unit];
    const T us = fabs(r.z - r.x) / init_u;
    const T vs = fabs(r.w - r.y) / init_v;
    const T ps = smax(us, vs);
    pad = raw * ps * ps;
    if (!isfinite(pad) || pad < T(0)) pad = T(0);
  }
  ctr.x = (lo[0] + hi[0]) / T(2);
  ctr.y = (lo[1] + hi[1]) / T(2);
  ctr.z = (lo[2] + hi[2]) / T(2);
  ctr.w = T(0);
  rad.x = (hi[0] - lo[0]) / T(2) + pad;
  rad.y = (hi[1] - lo[1]) / T(2) + pad;
  rad.z = (hi[2] - lo[2]) / T(2) + pad;
  rad.w = T(0);
  C[gid] = ctr;
  R[gid] = rad;
}

// ---------------------------------------------------------------------------
// Tiles: 16x16 blocks of cells of one unit's (nuv-1)^2 cell grid.
// ---------------------------------------------------------------------------
struct TileGeo {
  int row0, col0, ht, wd;
};
__device__ __forceinline__ TileGeo tile_geo(int t, int w) {
  const int tw = (w + kTile - 1) / kTile;
  TileGeo g;
  g.row0 = (t / tw) * kTile;
  g.col0 = (t % tw) * kTile;
  g.ht = min(kTile, w - g.row0);
  g.wd = min(kTile, w - g.col0);
  return g;
}
__device__ __forceinline__ int tiles_per_unit(int w) {
  const int tw = (w + kTile - 1) / kTile;
  return tw * tw;
}

// One wavefront per tile: min/max over the tile's cells.
template <class T, bool ISECT>
__global__ void k_tilebox(const vec4_t<T>* C, const vec4_t<T>* R, int nunits, int w, vec4_t<T>* lo,
                          vec4_t<T>* hi) {
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = lane_id();
  const int tpu = tiles_per_unit(w);
  if (wave >= (long long)nunits * tpu) return;
  const int unit = (int)(wave / tpu), t = (int)(wave % tpu);
  const TileGeo g = tile_geo(t, w);
  const T inf = infinity_v<T>();
  T mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
  T amax = T(0);
  const size_t ubase = (size_t)unit * w * w;
  for (int e = lane; e < g.ht * g.wd; e += 64) {
    const int cell = (g.row0 + e / g.wd) * w + g.col0 + e % g.wd;
    const vec4_t<T> c = C[ubase + cell];
    T v[3] = {c.x, c.y, c.z};
    if (ISECT) {
      const vec4_t<T> r = R[ubase + cell];
      const T rr[3] = {r.x, r.y, r.z};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = fmin(mn[a], v[a] - rr[a]);
        mx[a] = fmax(mx[a], v[a] + rr[a]);
        amax = fmax(amax, fabs(v[a]) + fabs(rr[a]));
      }
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = fmin(mn[a], v[a]);
        mx[a] = fmax(mx[a], v[a]);
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = fmin(mn[a], __shfl_xor(mn[a], off));
      mx[a] = fmax(mx[a], __shfl_xor(mx[a], off));
    }
    amax = fmax(amax, __shfl_xor(amax, off));
  }
  if (lane == 0) {
    vec4_t<T> l, h;
    T delta = T(0);
    if (ISECT) {
      // Outward margin covering the rounding of c-r, c+r and of the cell test
      // |c1-c2| < r1+r2 (a few units in the last place); 2^-20 relative is far
      // above that for float and double.
      delta = amax * T(9.5367431640625e-07) + T(1e-30);
    }
    l.x = mn[0] - delta;
    l.y = mn[1] - delta;
    l.z = mn[2] - delta;
    l.w = T(0);
    h.x = mx[0] + delta;
    h.y = mx[1] + delta;
    h.z = mx[2] + delta;
    h.w = T(0);
    lo[wave] = l;
    hi[wave] = h;
  }
}

struct TPIndex {
  int slot, t1, t2;
};
__device__ __forceinline__ TPIndex tp_decode(long long gid, int tpu1, int tpu2) {
  TPIndex r;
  const long long per = (long long)tpu1 * tpu2;
  r.slot = (int)(gid / per);
  const int rem = (int)(gid % per);
  r.t1 = rem / tpu2;
  r.t2 = rem % tpu2;
  return r;
}
__device__ __forceinline__ uint64_t tp_encode(int slot, int t1, int t2) {
  return ((uint64_t)(uint32_t)slot << 32) | ((uint64_t)t1 << 16) | (uint64_t)t2;
}

// Seed tile pairs of the distance screen.  For every (slot, side-1 tile) the
// side-2 tile with the smallest upper bound on centre d2 is kept (packed
// key: upper-bound bits << 32 | side-2 tile), then the nseed best of those
// (at most kSeedMaxTiles) form the seed list.  The exact k-th d2 over the
// seed's cell pairs bounds the global k-th d2 from above; any choice is
// valid, a good one is tight.
template <class T, bool FMA>
__device__ __forceinline__ T pair_d2(T ax, T ay, T az, T bx, T by, T bz);

constexpr int kSeedMaxTiles = 256;  // seed tile pairs at most
constexpr int kSeedSortMax = 8192;

// One workgroup per (slot, side-1 tile): the side-2 tile with the smallest
// upper bound, packed as (upper-bound bits << 32 | side-2 tile).
template <class T>
__device__ __forceinline__ void k_dist_seed_body(int nslots, const int* su1, const int* su2, int w1, int w2,
                                                   const vec4_t<T>* lo1, const vec4_t<T>* hi1,
                                                   const vec4_t<T>* lo2, const vec4_t<T>* hi2, int k,
                                                   unsigned long long* best) {
  __shared__ unsigned long long red[4];
  const int tpu1 = tiles_per_unit(w1), tpu2 = tiles_per_unit(w2);
  const int pos = blockIdx.x;  // slot * tpu1 + t1
  if (pos >= nslots * tpu1) return;
  (void)k;
  const int slot = pos / tpu1, t1 = pos % tpu1;
  const int a = su1[slot] * tpu1 + t1;
  const vec4_t<T> l1 = lo1[a], h1 = hi1[a];
  unsigned long long key = ~0ull;
  for (int t2 = threadIdx.x; t2 < tpu2; t2 += blockDim.x) {
    const int b = su2[slot] * tpu2 + t2;
    const vec4_t<T> l2 = lo2[b], h2 = hi2[b];
    const T ux = smax(h1.x - l2.x, h2.x - l1.x);
    const T uy = smax(h1.y - l2.y, h2.y - l1.y);
    const T uz = smax(h1.z - l2.z, h2.z - l1.z);
    const T ub2 = ux * ux + uy * uy + uz * uz;
    const unsigned long long kk = ((unsigned long long)ord32_up(ub2) << 32) | (unsigned long long)t2;
    key = kk < key ? kk : key;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(key, off);
    key = o < key ? o : key;
  }
  if (lane_id() == 0) red[threadIdx.x >> 6] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] < m ? red[w] : m;
    best[pos] = m;
  }
}
template <class T>
__global__ __launch_bounds__(256) void k_dist_seed(int nslots, const int* su1, const int* su2, int w1, int w2,
                                                   const vec4_t<T>* lo1, const vec4_t<T>* hi1,
                                                   const vec4_t<T>* lo2, const vec4_t<T>* hi2, int k,
                                                   unsigned long long* best) {
  k_dist_seed_body<T>(nslots, su1, su2, w1, w2, lo1, hi1, lo2, hi2, k, best);
}

// Single workgroup: the seed tile pairs, the nseed positions with the
// smallest upper bound (ordered = 1, one tile pair per slot and slots in rank
// order of their parent pairs: the first nseed slots).
template <class T>
__device__ __forceinline__ void k_seed_list_body(const unsigned long long* best, int npos, int tpu1, int nseed,
                                                    int ordered, uint64_t* list, unsigned* list_count) {
  __shared__ unsigned long long s[kSeedSortMax];
  __shared__ unsigned umax;
  const int m = nseed < npos ? nseed : npos;
  if (threadIdx.x == 0) umax = 0;
  if (!ordered) {
    int n2 = 1;
    while (n2 < npos) n2 <<= 1;
    for (int e = threadIdx.x; e < n2; e += blockDim.x)
      s[e] = e < npos ? ((best[e] & 0xffffffff00000000ull) | (unsigned long long)e) : ~0ull;
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < n2 / 2; t += blockDim.x) {
          const int i = 2 * t - (t & (j - 1)), l = i + j;
          const bool asc = (i & k) == 0;
          const unsigned long long x = s[i], y = s[l];
          if (asc ? (y < x) : (x < y)) {
            s[i] = y;
            s[l] = x;
          }
        }
        __syncthreads();
      }
  } else {
    for (int e = threadIdx.x; e < m; e += blockDim.x)
      s[e] = (best[e] & 0xffffffff00000000ull) | (unsigned long long)e;
    __syncthreads();
  }
  if ((int)threadIdx.x < m) {
    const int pos = (int)(s[threadIdx.x] & 0xffffffffull);
    const unsigned long long b = best[pos];
    list[threadIdx.x] = tp_encode(pos / tpu1, pos % tpu1, (int)(b & 0xffffffffull));
    atomicMax(&umax, (unsigned)(b >> 32));
  }
  __syncthreads();
  if (threadIdx.x == 0) *list_count = (unsigned)m;
}
template <class T>
__global__ __launch_bounds__(1024) void k_seed_list(const unsigned long long* best, int npos, int tpu1, int nseed,
                                                    int ordered, uint64_t* list, unsigned* list_count) {
  k_seed_list_body<T>(best, npos, tpu1, nseed, ordered, list, list_count);
}

constexpr int kSeedSplitMax = 16;    // workgroups per seed tile pair at most
constexpr int kSeedHistBins = 1024;  // bins of the seed histogram
constexpr int kSeedSample = 4;       // the range pass reads every 4th side-2 cell

// Bin width of the seed histogram over the key range [lo, hi]: 2^shift, the
// least power of two that puts hi in bin kSeedHistBins - 1 or below (a shift
// instead of a 64-bit division per key; 512-1024 bins span the range).
__device__ __forceinline__ int seed_bin_shift(unsigned long long lo, unsigned long long hi) {
  const unsigned long long r = (hi > lo ? hi - lo : 0ull) >> 10;
  static_assert(kSeedHistBins == 1024, "the shift assumes 1024 bins");
  return r ? 64 - __clzll(r) : 0;
}

// Two passes over the d2 of the seed tile pairs (no candidate records are
// written).  Each workgroup handles one seed tile pair against a 1/split
// slice of its side-2 cells; seed tile pair `it` is list[it], or with
// ordered (one tile per unit, slots in parent rank order) slot `it`.  MINMAX:
// per-workgroup smallest and largest key of a sample (every kSeedSample-th
// side-2 cell) -> mm[2*block].  Otherwise: a histogram in LDS, bins linear in
// the key bits over the sampled [min, max] (its resolution follows the spread
// of the values: refined cells have d2 within 1e-5 of each other), keys
// outside it counted in the end bins, whose non-zero bins are added to the
// global histogram gh.
template <class T, bool FMA, bool MINMAX>
__device__ __forceinline__ void k_seed_hist_body(const uint64_t* list, int n, int ordered, const int* su1,
                                                   const int* su2, int w1, int w2, const vec4_t<T>* C1,
                                                   const vec4_t<T>* C2, int split, const unsigned long long* range,
                                                   unsigned long long* mm, unsigned* gh) {
  __shared__ vec4_t<T> sB[kTileCells];
  __shared__ unsigned h[MINMAX ? 1 : kSeedHistBins];
  __shared__ unsigned long long rmin[4], rmax[4];
  unsigned long long lo = 0;
  int shift = 0;
  if (!MINMAX) {
    lo = range[0];
    shift = seed_bin_shift(lo, range[1]);
    for (int b = threadIdx.x; b < kSeedHistBins; b += blockDim.x) h[b] = 0;
  }
  unsigned long long kmin = ~0ull, kmax = 0ull;
  const int it = (int)(blockIdx.x / split);
  const int part = (int)(blockIdx.x % split);
  if (it < n) {
    const uint64_t e = ordered ? tp_encode(it, 0, 0) : list[it];
    const int slot = (int)(e >> 32), t1 = (int)((e >> 16) & 0xffff), t2 = (int)(e & 0xffff);
    const int u1 = su1[slot], u2 = su2[slot];
    const TileGeo g1 = tile_geo(t1, w1), g2 = tile_geo(t2, w2);
    const int nA = g1.ht * g1.wd, nB = g2.ht * g2.wd;
    const int j0 = part * nB / split, j1 = (part + 1) * nB / split;
    if ((int)threadIdx.x < j1 - j0) {
      const int jj = j0 + threadIdx.x;
      const int j = (g2.row0 + jj / g2.wd) * w2 + g2.col0 + jj % g2.wd;
      sB[threadIdx.x] = C2[(size_t)u2 * w2 * w2 + j];
    }
    __syncthreads();
    if ((int)threadIdx.x < nA) {
      const int i = (g1.row0 + threadIdx.x / g1.wd) * w1 + g1.col0 + threadIdx.x % g1.wd;
      const vec4_t<T> a = C1[(size_t)u1 * w1 * w1 + i];
      for (int jj = 0; jj < j1 - j0; jj += MINMAX ? kSeedSample : 1) {
        const vec4_t<T> b = sB[jj];
        const unsigned long long key = ord_bits(pair_d2<T, FMA>(a.x, a.y, a.z, b.x, b.y, b.z));
        if (MINMAX) {
          kmin = key < kmin ? key : kmin;
          kmax = key > kmax ? key : kmax;
        } else {
          unsigned long long bin = key <= lo ? 0ull : (key - lo) >> shift;
          if (bin >= (unsigned long long)kSeedHistBins) bin = kSeedHistBins - 1;
          atomicAdd(&h[bin], 1u);
        }
      }
    }
  }
  if (MINMAX) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long a = __shfl_xor(kmin, off), b = __shfl_xor(kmax, off);
      kmin = a < kmin ? a : kmin;
      kmax = b > kmax ? b : kmax;
    }
    if (lane_id() == 0) {
      rmin[threadIdx.x >> 6] = kmin;
      rmax[threadIdx.x >> 6] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < 4; ++w) {
        kmin = rmin[w] < kmin ? rmin[w] : kmin;
        kmax = rmax[w] > kmax ? rmax[w] : kmax;
      }
      mm[2 * blockIdx.x] = kmin;
      mm[2 * blockIdx.x + 1] = kmax;
    }
    return;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kSeedHistBins; b += blockDim.x)
    if (h[b]) atomicAdd(&gh[b], h[b]);
}
template <class T, bool FMA, bool MINMAX>
__global__ __launch_bounds__(256) void k_seed_hist(const uint64_t* list, int n, int ordered, const int* su1,
                                                   const int* su2, int w1, int w2, const vec4_t<T>* C1,
                                                   const vec4_t<T>* C2, int split, const unsigned long long* range,
                                                   unsigned long long* mm, unsigned* gh) {
  k_seed_hist_body<T, FMA, MINMAX>(list, n, ordered, su1, su2, w1, w2, C1, C2, split, range, mm, gh);
}

// Single workgroup: reduce the per-workgroup min/max into range[0..1] and
// clear the global histogram for the second pass.
__device__ __forceinline__ void k_seed_range_body(const unsigned long long* mm, int nblocks,
                                                     unsigned long long* range, unsigned* gh) {
  __shared__ unsigned long long smin[1024], smax[1024];
  unsigned long long a = ~0ull, b = 0ull;
  for (int x = threadIdx.x; x < nblocks; x += blockDim.x) {
    a = mm[2 * x] < a ? mm[2 * x] : a;
    b = mm[2 * x + 1] > b ? mm[2 * x + 1] : b;
  }
  smin[threadIdx.x] = a;
  smax[threadIdx.x] = b;
  for (int x = threadIdx.x; x < kSeedHistBins; x += blockDim.x) gh[x] = 0u;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      smin[threadIdx.x] = smin[threadIdx.x + off] < smin[threadIdx.x] ? smin[threadIdx.x + off] : smin[threadIdx.x];
      smax[threadIdx.x] = smax[threadIdx.x + off] > smax[threadIdx.x] ? smax[threadIdx.x + off] : smax[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    range[0] = smin[0];
    range[1] = smax[0];
  }
}
__global__ __launch_bounds__(1024) void k_seed_range(const unsigned long long* mm, int nblocks,
                                                     unsigned long long* range, unsigned* gh) {
  k_seed_range_body(mm, nblocks, range, gh);
}

// Single workgroup: threshold = upper edge of the histogram bin holding rank
// k (at least k seed pairs lie at or below it); +inf when the seeds hold
// fewer than k pairs.
template <class T>
__device__ __forceinline__ void k_seed_thresh_body(const unsigned* gh, const unsigned long long* range, int k,
                                                      T* thresh, unsigned* clear, int nclear) {
  if ((int)threadIdx.x < nclear) clear[threadIdx.x] = 0u;  // the screen's counters, free again after the seed
  __shared__ unsigned long long cum[kSeedHistBins];
  __shared__ int found;
  const int b = threadIdx.x;  // blockDim.x == kSeedHistBins
  const unsigned long long c = gh[b];
  cum[b] = c;
  if (b == 0) found = -1;
  __syncthreads();
  for (int off = 1; off < kSeedHistBins; off <<= 1) {
    const unsigned long long v = b >= off ? cum[b - off] : 0ull;
    __syncthreads();
    cum[b] += v;
    __syncthreads();
  }
  const unsigned long long before = cum[b] - c;
  if (before < (unsigned long long)k && cum[b] >= (unsigned long long)k) found = b;
  __syncthreads();
  if (b == 0) {
    // Bin `found` < last holds keys up to its upper edge only, so the k
    // smallest keys are at most that edge.  The last bin also counts every
    // key above the sampled range: then there is no threshold.
    const unsigned long long inf_bits = ord_bits(infinity_v<T>());
    unsigned long long edge = inf_bits;
    if (found >= 0 && found < kSeedHistBins - 1) {
      const unsigned long long lo = range[0];
      const int shift = seed_bin_shift(lo, range[1]);
      edge = lo + ((unsigned long long)(found + 1) << shift) - 1;
    }
    *thresh = edge >= inf_bits ? infinity_v<T>() : from_bits<T>(edge);
  }
}
template <class T>
__global__ __launch_bounds__(1024) void k_seed_thresh(const unsigned* gh, const unsigned long long* range, int k,
                                                      T* thresh, unsigned* clear, int nclear) {
  k_seed_thresh_body<T>(gh, range, k, thresh, clear, nclear);
}

// Active tile pairs of the distance screen.  Lower bound on the centre d2 of
// any cell pair of a tile pair: per axis the gap between the two centre boxes.
// Correctly rounded subtraction, squaring and addition are monotone, so the
// bound computed in T never exceeds any computed cell-pair d2; a tile pair is
// skipped only when the bound is strictly above the threshold.
template <class T, bool FMA>
__device__ __forceinline__ void k_dist_list_body(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                            const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, const T* thresh,
                            int cull, uint64_t* list, unsigned* list_count) {
  const int tpu1 = tiles_per_unit(w1), tpu2 = tiles_per_unit(w2);
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool keep = false;
  TPIndex x{0, 0, 0};
  if (gid < (long long)nslots * tpu1 * tpu2) {
    x = tp_decode(gid, tpu1, tpu2);
    keep = true;
    if (cull) {
      const int a = su1[x.slot] * tpu1 + x.t1, b = su2[x.slot] * tpu2 + x.t2;
      const vec4_t<T> l1 = lo1[a], h1 = hi1[a], l2 = lo2[b], h2 = hi2[b];
      const T gx = smax(T(0), smax(l1.x - h2.x, l2.x - h1.x));
      const T gy = smax(T(0), smax(l1.y - h2.y, l2.y - h1.y));
      const T gz = smax(T(0), smax(l1.z - h2.z, l2.z - h1.z));
      const T lb2 = FMA ? fma(gx, gx, fma(gy, gy, gz * gz)) : gx * gx + gy * gy + gz * gz;
      keep = !(lb2 > *thresh);
    }
  }
  const unsigned long long m = __ballot(keep);
  if (m) {
    unsigned base = 0;
    const int leader = __ffsll((long long)m) - 1;
    if (lane_id() == leader) base = atomicAdd(list_count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (keep) list[base + __popcll(m & ((1ull << lane_id()) - 1))] = tp_encode(x.slot, x.t1, x.t2);
  }
}
template <class T, bool FMA>
__global__ void k_dist_list(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                            const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, const T* thresh,
                            int cull, uint64_t* list, unsigned* list_count) {
  k_dist_list_body<T, FMA>(nslots, su1, su2, w1, w2, lo1, hi1, lo2, hi2, thresh, cull, list, list_count);
}

// The same test with one workgroup per (slot, side-1 tile) and its threads
// over the side-2 tiles (many tiles per unit, round 1): the side-1 box and
// the unit ids are uniform loads, the side-2 boxes are read coalesced, and
// each workgroup appends its kept tile pairs with one atomic per 256 tiles.
template <class T, bool FMA>
__device__ __forceinline__ void k_dist_list_rows_body(int nslots, const int* su1, const int* su2, int w1, int w2,
                                                        const vec4_t<T>* lo1, const vec4_t<T>* hi1,
                                                        const vec4_t<T>* lo2, const vec4_t<T>* hi2,
                                                        const T* thresh, int cull, uint64_t* list,
                                                        unsigned* list_count) {
  // side-2 tiles in groups of four runs of 256 (a 511 x 511-cell unit has
  // 1024 tiles: one group): the boxes of a group are loaded together and its
  // kept tile pairs reserved with one atomic, so a workgroup waits for one
  // load and one reservation per group, not per run
  constexpr int kRuns = 4;
  __shared__ unsigned wsum[kRuns][4], sbase;
  const int tpu1 = tiles_per_unit(w1), tpu2 = tiles_per_unit(w2);
  const int pos = blockIdx.x;  // slot * tpu1 + t1
  if (pos >= nslots * tpu1) return;  // whole workgroup
  const int slot = pos / tpu1, t1 = pos - slot * tpu1;
  const int a = su1[slot] * tpu1 + t1, ub = su2[slot] * tpu2;
  const vec4_t<T> l1 = lo1[a], h1 = hi1[a];
  const T th = *thresh;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (int g0 = 0; g0 < tpu2; g0 += kRuns * 256) {
    bool keep[kRuns];
    unsigned long long m[kRuns];
#pragma unroll
    for (int r = 0; r < kRuns; ++r) {
      const int t2 = g0 + 256 * r + (int)threadIdx.x;
      keep[r] = t2 < tpu2;
      if (keep[r] && cull) {
        const vec4_t<T> l2 = lo2[ub + t2], h2 = hi2[ub + t2];
        const T gx = smax(T(0), smax(l1.x - h2.x, l2.x - h1.x));
        const T gy = smax(T(0), smax(l1.y - h2.y, l2.y - h1.y));
        const T gz = smax(T(0), smax(l1.z - h2.z, l2.z - h1.z));
        const T lb2 = FMA ? fma(gx, gx, fma(gy, gy, gz * gz)) : gx * gx + gy * gy + gz * gz;
        keep[r] = !(lb2 > th);
      }
    }
#pragma unroll
    for (int r = 0; r < kRuns; ++r) {
      m[r] = __ballot(keep[r]);
      if (lane == 0) wsum[r][wv] = (unsigned)__popcll(m[r]);
    }
    __syncthreads();
    unsigned total = 0;
#pragma unroll
    for (int r = 0; r < kRuns; ++r) total += wsum[r][0] + wsum[r][1] + wsum[r][2] + wsum[r][3];
    if (total) {  // uniform
      if (threadIdx.x == 0) sbase = atomicAdd(list_count, total);
      __syncthreads();
      unsigned off = sbase;  // runs in order, waves in order within a run
#pragma unroll
      for (int r = 0; r < kRuns; ++r) {
        unsigned o = off;
        for (int v = 0; v < wv; ++v) o += wsum[r][v];
        if (keep[r])
          list[o + __popcll(m[r] & ((1ull << lane) - 1))] = tp_encode(slot, t1, g0 + 256 * r + (int)threadIdx.x);
        off += wsum[r][0] + wsum[r][1] + wsum[r][2] + wsum[r][3];
      }
    }
    __syncthreads();  // wsum / sbase are rewritten by the next group
  }
}
template <class T, bool FMA>
__global__ __launch_bounds__(256) void k_dist_list_rows(int nslots, const int* su1, const int* su2, int w1, int w2,
                                                        const vec4_t<T>* lo1, const vec4_t<T>* hi1,
                                                        const vec4_t<T>* lo2, const vec4_t<T>* hi2,
                                                        const T* thresh, int cull, uint64_t* list,
                                                        unsigned* list_count) {
  k_dist_list_rows_body<T, FMA>(nslots, su1, su2, w1, w2, lo1, hi1, lo2, hi2, thresh, cull, list, list_count);
}

// d2 of two centres.  FMA = false: center_distance_sq's order
// (distance_analysis.cu:202-207), no contraction (EXACT32 / FP64).  FMA = true:
// fma(dx, dx, fma(dy, dy, dz*dz)) for the FP32SCREEN screen, whose threshold
// margin covers its rounding (three roundings instead of five).
template <class T, bool FMA>
__device__ __forceinline__ T pair_d2(T ax, T ay, T az, T bx, T by, T bz) {
  const T dx = ax - bx;
  const T dy = ay - by;
  const T dz = az - bz;
  if (FMA) return fma(dx, dx, fma(dy, dy, dz * dz));
  return dx * dx + dy * dy + dz * dz;
}

template <class T>
__device__ __forceinline__ T quiet_nan();
template <>
__device__ __forceinline__ float quiet_nan<float>() {
  return __uint_as_float(0x7fc00000u);
}
template <>
__device__ __forceinline__ double quiet_nan<double>() {
  return __longlong_as_double(0x7ff8000000000000ll);
}

// Distance filter: every cell pair (i, j) of every listed tile pair with
// centre d2 <= threshold is appended to `out` (order irrelevant: records are
// sorted by the full key afterwards).  One tile pair per workgroup iteration;
// side-2 cells are staged in LDS (padded to a multiple of 8 with NaN, which
// never passes the test) and read as broadcasts, each lane owns one side-1
// cell.  The j loop is unrolled by 8 and tests the minimum of 8 values once;
// only a passing group re-tests its members and appends them.
//
// EXP (double): the test runs on the expanded form.  Side-2 cells are staged
// as (-2b, |b|^2), so s = a.(-2b) + |b|^2 = d2 - |a|^2 costs three FMAs, and
// a pair passes when s <= T - |a|^2 + delta, delta = 32 u (T + 4 m^2), u =
// 2^-53, m = largest |centre coordinate| of the round (*absmax).  delta
// bounds the rounding of s, of |a|^2, |b|^2, of T - |a|^2 and of the
// reference formula dx*dx + dy*dy + dz*dz (each at most a few u (T + 3 m^2)),
// so every pair whose reference d2 is <= T passes.  A passing pair's key is
// the reference d2 recomputed from a and b = -(1/2)(-2b) (exact), so records
// are identical to the direct test's; pairs that pass only by the margin carry
// their true key and cannot displace a smaller one.
// Profiling count of evaluated cell pairs: the workgroup's wavefront counts
// (held by lane 0 of each wavefront) are summed in LDS and added to one of
// kPairsSlots counters picked by workgroup index (thousands of atomics on a
// single address serialise and cost tens of microseconds per launch).  Every
// thread of the workgroup calls it.
__device__ __forceinline__ void add_pairs_done(unsigned long long* pairs_done, unsigned long long v) {
  __shared__ unsigned long long s_done[8];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_done[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sum = 0;
    for (int q = 0; q < (int)((blockDim.x + 63) >> 6); ++q) sum += s_done[q];
    if (sum) atomicAdd(&pairs_done[blockIdx.x & (kPairsSlots - 1)], sum);
  }
}

template <class T, bool FMA, bool EXP>
__device__ __forceinline__ void k_dist_filter_body(const uint64_t* list, const unsigned* list_count,
                                                     const int* su1, const int* su2, int w1, int w2,
                                                     const vec4_t<T>* C1, const vec4_t<T>* C2, const T* thresh,
                                                     Cand* out, unsigned long long* out_count,
                                                     unsigned long long cap, unsigned long long* pairs_done,
                                                     int cull, const unsigned long long* absmax) {
  static_assert(!EXP || std::is_same<T, double>::value, "expanded test: double only");
  constexpr int U = 8;
  constexpr int kStage = 128;  // candidates staged per wavefront before one global reservation
  __shared__ vec4_t<T> sB[kTileCells];
  __shared__ int sJ[kTileCells];
  __shared__ Cand stage[4][kStage];
  __shared__ vec4_t<T> sLo[kTileCells / 32], sHi[kTileCells / 32];
  unsigned long long done = 0;
  const unsigned n = *list_count;
  const T tv = *thresh;
  T delta = T(0);
  if (EXP) {
    unsigned long long mb = absmax[lane_id()];  // kAmaxSlots == wavefront size
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(mb, off);
      mb = o > mb ? o : mb;
    }
    const T m = __longlong_as_double((long long)mb);
    delta = T(32) * T(1.1102230246251565e-16) * (tv + T(4) * m * m);
  }
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  int wcount = 0;  // staged candidates of this wavefront (wave-uniform)
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(out_count, (unsigned long long)wcount);
    base = __shfl(base, 0);
    for (int e = lane; e < wcount; e += 64)
      if (base + e < cap) out[base + e] = stage[wv][e];
    __builtin_amdgcn_wave_barrier();
    wcount = 0;
  };
  for (unsigned it = blockIdx.x; it < n; it += gridDim.x) {
    const uint64_t e = list[it];
    const int slot = (int)(e >> 32), t1 = (int)((e >> 16) & 0xffff), t2 = (int)(e & 0xffff);
    const int u1 = su1[slot], u2 = su2[slot];
    const TileGeo g1 = tile_geo(t1, w1), g2 = tile_geo(t2, w2);
    const int nA = g1.ht * g1.wd, nB = g2.ht * g2.wd;
    const int nBp = (nB + U - 1) / U * U;
    __syncthreads();
    // side-2 cell t of the tile pair: staged in LDS (EXP: as -2b, |b|^2) and
    // kept in registers for the chunk boxes
    const bool bin = (int)threadIdx.x < nB;
    vec4_t<T> bme;
    bme.x = bme.y = bme.z = bme.w = T(0);
    if (bin) {
      const int j = (g2.row0 + threadIdx.x / g2.wd) * w2 + g2.col0 + threadIdx.x % g2.wd;
      bme = C2[(size_t)u2 * w2 * w2 + j];
      if (EXP) {
        vec4_t<T> x;
        x.x = T(-2) * bme.x;
        x.y = T(-2) * bme.y;
        x.z = T(-2) * bme.z;
        x.w = bme.x * bme.x + bme.y * bme.y + bme.z * bme.z;
        sB[threadIdx.x] = x;
      } else {
        sB[threadIdx.x] = bme;
      }
      sJ[threadIdx.x] = j;
    } else if ((int)threadIdx.x < nBp) {
      vec4_t<T> b;
      b.x = b.y = b.z = b.w = quiet_nan<T>();
      sB[threadIdx.x] = b;
      sJ[threadIdx.x] = 0;
    }
    const bool va = (int)threadIdx.x < nA;
    int i = 0;
    vec4_t<T> a;
    a.x = a.y = a.z = a.w = quiet_nan<T>();  // inactive lanes never pass
    if (va) {
      i = (g1.row0 + threadIdx.x / g1.wd) * w1 + g1.col0 + threadIdx.x % g1.wd;
      a = C1[(size_t)u1 * w1 * w1 + i];
    }
    // bounding box of every 32-cell chunk of the side-2 tile: a reduction
    // over the 32 lanes that hold the chunk's cells (min and max do not
    // depend on the order)
    constexpr int kChunk = 32;
    const int nchunk = (nBp + kChunk - 1) / kChunk;
    {
      const T inf = infinity_v<T>();
      T bl[3] = {bin ? bme.x : inf, bin ? bme.y : inf, bin ? bme.z : inf};
      T bh[3] = {bin ? bme.x : -inf, bin ? bme.y : -inf, bin ? bme.z : -inf};
#pragma unroll
      for (int off = kChunk / 2; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bl[c] = fmin(bl[c], __shfl_xor(bl[c], off));
          bh[c] = fmax(bh[c], __shfl_xor(bh[c], off));
        }
      if ((threadIdx.x & (kChunk - 1)) == 0 && (int)(threadIdx.x / kChunk) < nchunk) {
        vec4_t<T> lo, hi;
        lo.x = bl[0];
        lo.y = bl[1];
        lo.z = bl[2];
        hi.x = bh[0];
        hi.y = bh[1];
        hi.z = bh[2];
        lo.w = hi.w = T(0);
        sLo[threadIdx.x / kChunk] = lo;
        sHi[threadIdx.x / kChunk] = hi;
      }
    }
    T thr = tv;
    if (EXP) thr = (tv - (a.x * a.x + a.y * a.y + a.z * a.z)) + delta;  // NaN for inactive lanes
    __syncthreads();
    const uint64_t key_si = pack_idx((uint32_t)slot, (uint32_t)i, 0u);
    const int nvalid = __popcll(__ballot(va));
    for (int j0 = 0; j0 < nBp; j0 += U) {
      if ((j0 & (kChunk - 1)) == 0) {
        if (cull) {
          // point-to-box gap: a lower bound of every d2 of the chunk (monotone
          // rounding), so a chunk no lane can reach is skipped
          const vec4_t<T> lo = sLo[j0 / kChunk], hi = sHi[j0 / kChunk];
          const T gx = smax(T(0), smax(lo.x - a.x, a.x - hi.x));
          const T gy = smax(T(0), smax(lo.y - a.y, a.y - hi.y));
          const T gz = smax(T(0), smax(lo.z - a.z, a.z - hi.z));
          const T lb2 = FMA ? fma(gx, gx, fma(gy, gy, gz * gz)) : gx * gx + gy * gy + gz * gz;
          if (__ballot(lb2 <= tv) == 0ull) {
            j0 += kChunk - U;
            continue;
          }
        }
        // algorithmic work: this wavefront evaluates its lanes' pairs with the chunk
        if (lane == 0) done += (unsigned long long)nvalid * (unsigned long long)max(0, min(kChunk, nB - j0));
      }
      T d[U];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const vec4_t<T> b = sB[j0 + q];
        if (EXP)
          d[q] = fma(a.x, b.x, fma(a.y, b.y, fma(a.z, b.z, b.w)));
        else
          d[q] = pair_d2<T, FMA>(a.x, a.y, a.z, b.x, b.y, b.z);
      }
      T m = fmin(fmin(fmin(d[0], d[1]), fmin(d[2], d[3])), fmin(fmin(d[4], d[5]), fmin(d[6], d[7])));
      if (__ballot(m <= thr) == 0ull) continue;
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const bool hit = d[q] <= thr;
        const unsigned long long mk = __ballot(hit);
        if (mk) {
          const int cnt = __popcll(mk);
          if (wcount + cnt > kStage) flush();
          if (hit) {
            T dq = d[q];
            if (EXP) {
              const vec4_t<T> b = sB[j0 + q];
              dq = pair_d2<T, false>(a.x, a.y, a.z, T(-0.5) * b.x, T(-0.5) * b.y, T(-0.5) * b.z);
            }
            Cand c;
            c.hi = ord_bits(dq);
            c.lo = key_si | (uint64_t)(uint32_t)sJ[j0 + q];
            stage[wv][wcount + __popcll(mk & ((1ull << lane) - 1))] = c;
          }
          wcount += cnt;
        }
      }
    }
  }
  if (wcount) flush();
  if (pairs_done) add_pairs_done(pairs_done, done);
}
template <class T, bool FMA, bool EXP>
__global__ __launch_bounds__(256) void k_dist_filter(const uint64_t* list, const unsigned* list_count,
                                                     const int* su1, const int* su2, int w1, int w2,
                                                     const vec4_t<T>* C1, const vec4_t<T>* C2, const T* thresh,
                                                     Cand* out, unsigned long long* out_count,
                                                     unsigned long long cap, unsigned long long* pairs_done,
                                                     int cull, const unsigned long long* absmax) {
  k_dist_filter_body<T, FMA, EXP>(list, list_count, su1, su2, w1, w2, C1, C2, thresh, out, out_count, cap, pairs_done, cull, absmax);
}

// Distance filter for slots whose units are single tiles (w1, w2 <= 16: the
// refinement rounds).  Same records as k_dist_filter (including the EXP
// test), different work split: each lane keeps R = 4 side-1 cells in
// registers and each wavefront takes a quarter of the side-2 cells, so one
// broadcast LDS read of a side-2 cell serves four pairs per lane (the tile
// kernel is bound by the LDS return bandwidth of those broadcasts).  Within a
// group of U side-2 cells each of the R cells tests the minimum of its U
// values first.  No chunk culling: refined units are near each other
// everywhere, so it would prune nothing.
template <class T, bool FMA, bool EXP>
__device__ __forceinline__ void k_dist_filter_units_body(const uint64_t* list, const unsigned* list_count,
                                                           const int* su1, const int* su2, int w1, int w2,
                                                           const vec4_t<T>* C1, const vec4_t<T>* C2,
                                                           const T* thresh, Cand* out, unsigned long long* out_count,
                                                           unsigned long long cap, unsigned long long* pairs_done,
                                                           const unsigned long long* absmax) {
  static_assert(!EXP || std::is_same<T, double>::value, "expanded test: double only");
  constexpr int R = 4;  // side-1 cells per lane: 4 x 64 covers a 16x16 unit
  constexpr int U = 4;  // side-2 cells per unrolled group
  constexpr int kStage = 128;
  __shared__ vec4_t<T> sB[kTileCells];
  __shared__ Cand stage[4][kStage];
  unsigned long long done = 0;
  const unsigned n = *list_count;
  const T tv = *thresh;
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  T delta = T(0);
  if (EXP) {
    unsigned long long mb = absmax[lane];  // kAmaxSlots == wavefront size
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(mb, off);
      mb = o > mb ? o : mb;
    }
    const T m = __longlong_as_double((long long)mb);
    delta = T(32) * T(1.1102230246251565e-16) * (tv + T(4) * m * m);
  }
  const int nA = w1 * w1, nB = w2 * w2;
  const int nBp = (nB + 4 * U - 1) / (4 * U) * (4 * U);
  const int per = nBp / 4;  // side-2 cells per wavefront, a multiple of U
  int wcount = 0;
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(out_count, (unsigned long long)wcount);
    base = __shfl(base, 0);
    for (int e = lane; e < wcount; e += 64)
      if (base + e < cap) out[base + e] = stage[wv][e];
    __builtin_amdgcn_wave_barrier();
    wcount = 0;
  };
  for (unsigned it = blockIdx.x; it < n; it += gridDim.x) {
    const int slot = (int)(list[it] >> 32);
    const int u1 = su1[slot], u2 = su2[slot];
    __syncthreads();  // sB of the previous slot is no longer read
    if ((int)threadIdx.x < nB) {
      vec4_t<T> b = C2[(size_t)u2 * nB + threadIdx.x];
      if (EXP) {
        vec4_t<T> x;
        x.x = T(-2) * b.x;
        x.y = T(-2) * b.y;
        x.z = T(-2) * b.z;
        x.w = b.x * b.x + b.y * b.y + b.z * b.z;
        b = x;
      }
      sB[threadIdx.x] = b;
    } else if ((int)threadIdx.x < nBp) {
      vec4_t<T> b;
      b.x = b.y = b.z = b.w = quiet_nan<T>();  // never passes
      sB[threadIdx.x] = b;
    }
    vec4_t<T> a[R];
    T thr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane + 64 * r;
      a[r].x = a[r].y = a[r].z = a[r].w = quiet_nan<T>();  // absent cells never pass
      if (i < nA) a[r] = C1[(size_t)u1 * nA + i];
      thr[r] = EXP ? (tv - (a[r].x * a[r].x + a[r].y * a[r].y + a[r].z * a[r].z)) + delta : tv;
    }
    __syncthreads();
    if (threadIdx.x == 0) done += (unsigned long long)nA * (unsigned long long)nB;
    const uint64_t key_s = pack_idx((uint32_t)slot, 0u, 0u);
    for (int j0 = wv * per; j0 < (wv + 1) * per; j0 += U) {
      T d[U][R];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const vec4_t<T> b = sB[j0 + q];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (EXP)
            d[q][r] = fma(a[r].x, b.x, fma(a[r].y, b.y, fma(a[r].z, b.z, b.w)));
          else
            d[q][r] = pair_d2<T, FMA>(a[r].x, a[r].y, a[r].z, b.x, b.y, b.z);
        }
      }
      bool any = false;
#pragma unroll
      for (int r = 0; r < R; ++r) any |= fmin(fmin(d[0][r], d[1][r]), fmin(d[2][r], d[3][r])) <= thr[r];
      if (__ballot(any) == 0ull) continue;
#pragma unroll
      for (int q = 0; q < U; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const bool hit = d[q][r] <= thr[r];
          const unsigned long long mk = __ballot(hit);
          if (mk) {
            const int cnt = __popcll(mk);
            if (wcount + cnt > kStage) flush();
            if (hit) {
              T dq = d[q][r];
              if (EXP) {
                const vec4_t<T> b = sB[j0 + q];
                dq = pair_d2<T, false>(a[r].x, a[r].y, a[r].z, T(-0.5) * b.x, T(-0.5) * b.y, T(-0.5) * b.z);
              }
              Cand c;
              c.hi = ord_bits(dq);
              c.lo = key_s | ((uint64_t)(uint32_t)(lane + 64 * r) << 20) | (uint64_t)(uint32_t)(j0 + q);
              stage[wv][wcount + __popcll(mk & ((1ull << lane) - 1))] = c;
            }
            wcount += cnt;
          }
        }
    }
  }
  if (wcount) flush();
  if (pairs_done) add_pairs_done(pairs_done, done);
}
template <class T, bool FMA, bool EXP>
__global__ __launch_bounds__(256) void k_dist_filter_units(const uint64_t* list, const unsigned* list_count,
                                                           const int* su1, const int* su2, int w1, int w2,
                                                           const vec4_t<T>* C1, const vec4_t<T>* C2,
                                                           const T* thresh, Cand* out, unsigned long long* out_count,
                                                           unsigned long long cap, unsigned long long* pairs_done,
                                                           const unsigned long long* absmax) {
  k_dist_filter_units_body<T, FMA, EXP>(list, list_count, su1, su2, w1, w2, C1, C2, thresh, out, out_count, cap, pairs_done, absmax);
}

// Distance filter of the FP64 refinement rounds on the matrix cores.  The
// expanded test value s = a.(-2b) + |b|^2 (see k_dist_filter) of a cell pair is
// the product of the side-2 row [-2bx -2by -2bz |b|^2] and the side-1 column
// [ax ay az 1], so one v_mfma_f64_16x16x4f64 forms the values of a 16 x 16
// block of cell pairs (8 FLOP per pair, all on the matrix core); the vector
// ALU only compares them with the thresholds T - |a|^2 + delta, which depend on
// the column.  Operand layouts of the instruction: A lane l = (row l%16, k
// l/16), B lane l = (k l/16, column l%16), result register r of lane l = (row
// 4r + l/16, column l%16): a lane's results share its column, hence one
// threshold per lane and side-1 block, and "any pass" is min(results) <= thr.
// One workgroup per slot: the side-1 B operands and thresholds are staged in
// LDS (blocks of 16 cells).  Only full blocks go to the matrix cores (a unit of
// 225 cells has 14 full blocks and one cell): wavefront w keeps the A operands
// of side-2 blocks w, w+4, .. in registers and the products of the leftover
// full side-2 blocks are shared out so that the four wavefronts do equal work
// (49 products each for 14 x 14 blocks, where padding both sides to 16-cell
// blocks and splitting by side-2 block gave one wavefront 60); for each side-1
// block a wavefront issues its products before testing them.  Pairs with a
// cell of a partial block take the reference test d2 <= T on the vector ALU.
// delta covers the rounding of a 4-term dot product in any order (at most
// 36 u m^2 here) and the other terms of the expanded test (the bound is
// checked by the screen-bound tests).  A passing pair's key is the reference
// d2 recomputed from a and b, so the records equal the direct test's.
typedef double mfma_v4d __attribute__((ext_vector_type(4)));

// v_min_f64 on finite values (products of full blocks): written as the
// instruction because fmin() first canonicalises both operands (two extra
// instructions per minimum).  The compiler does not see inline asm as a reader
// of a matrix-core result, so callers insert the wait states themselves
// (k_dist_filter_mfma: s_nop 15; s_nop 2 between scheduling barriers).  The
// instruction order this relies on was read back from the code object of the
// toolchain below; another compiler must be checked the same way before the
// guard is widened.
#if !(HIP_VERSION_MAJOR == 7 && HIP_VERSION_MINOR == 2 && __clang_major__ == 22)
#error "k_dist_filter_mfma's hand-inserted wait states are verified for HIP 7.2 / clang 22 only"
#endif
__device__ __forceinline__ double min_nan(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

template <int NQ>  // NQ = (side-2 cells / 16) / 4: full side-2 blocks each wavefront holds in registers
__device__ __forceinline__ void k_dist_filter_mfma_body(const uint64_t* list, const unsigned* list_count,
                                                          int nslots, const int* su1, const int* su2, int w1, int w2,
                                                          const double4* C1, const double4* C2, const double* thresh,
                                                          Cand* out, unsigned long long* out_count,
                                                          unsigned long long cap, unsigned long long* pairs_done,
                                                          const unsigned long long* absmax) {
  constexpr int kStage = 128;
  __shared__ double sB[17][64];  // one padding block: the prefetch past the last block stays in bounds
  __shared__ double sT[17 * 16];
  __shared__ double sA[16][64];  // side-2 A operands (read by the rare hit path)
  __shared__ Cand stage[4][kStage];
  unsigned long long done = 0;
  const unsigned n = *list_count;
  const double tv = *thresh;
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  unsigned long long mb = absmax[lane];  // kAmaxSlots == wavefront size
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mb, off);
    mb = o > mb ? o : mb;
  }
  const double mm = __longlong_as_double((long long)mb);
  const double delta = 32.0 * 1.1102230246251565e-16 * (tv + 4.0 * mm * mm);
  const int nA = w1 * w1, nB = w2 * w2;
  const int nAf = nA >> 4, nBf = nB >> 4;  // full blocks of 16 cells (side-1: columns, side-2: rows)
  const int nL = nBf - 4 * NQ;              // full side-2 blocks beyond the wavefronts' NQ each (0..3)
  const double qnan = quiet_nan<double>();
  int wcount = 0;
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(out_count, (unsigned long long)wcount);
    base = __shfl(base, 0);
    for (int e = lane; e < wcount; e += 64)
      if (base + e < cap) out[base + e] = stage[wv][e];
    __builtin_amdgcn_wave_barrier();
    wcount = 0;
  };
  // The workgroup walks slots it = blockIdx.x + k gridDim.x (at most kMaxWalk of
  // them; grids hold at least one workgroup per kMaxWalk slots): their unit
  // ids are read up front, and the centres of the next slot are loaded into
  // registers while the current one computes.  When every slot is listed, the
  // list is the identity (records are sorted later, so the order of the slots
  // does not matter).
  constexpr int kMaxWalk = 16;
  const bool all = n == (unsigned)nslots;
  struct WalkSlot {
    int slot, u1, u2;  // slot -1: past the end of the walk
  };
  auto walk_slot = [&](int k) {
    WalkSlot w{-1, 0, 0};
    const unsigned i = blockIdx.x + (unsigned)k * gridDim.x;
    if (k < kMaxWalk && i < n) {
      w.slot = all ? (int)i : (int)(list[i] >> 32);
      w.u1 = su1[w.slot];
      w.u2 = su2[w.slot];
    }
    return w;
  };
  const int t = threadIdx.x;
  auto load_a = [&](int v1) {  // side-1 cell t
    double4 a;
    a.x = a.y = a.z = a.w = qnan;
    if (t < nA) a = C1[(size_t)v1 * nA + t];
    return a;
  };
  auto load_b = [&](int v2) {  // side-2 cell t (its A operand column is formed at staging)
    double4 b;
    b.x = b.y = b.z = b.w = qnan;
    if (t < nB) b = C2[(size_t)v2 * nB + t];
    return b;
  };
  // The slot ids are read two slots ahead and the next slot's cells are
  // loaded as they are stored (no arithmetic on them until the staging), so
  // the loads stay in flight across the products.  The walk is a rolled loop
  // (its body holds the product loops of every segment: unrolled copies would
  // outgrow the instruction cache).
  WalkSlot cur = walk_slot(0), nxt = walk_slot(1);
  double4 a_cur = load_a(cur.u1);
  double4 b_cur = load_b(cur.u2);
#pragma unroll 1
  for (int k = 0; cur.slot >= 0; ++k) {
    const int slot = cur.slot;
    double4 a_nxt, b_nxt;
    if (nxt.slot >= 0) {
      a_nxt = load_a(nxt.u1);
      b_nxt = load_b(nxt.u2);
    }
    const WalkSlot nxt2 = walk_slot(k + 2);
    __syncthreads();  // the previous slot's operands are no longer read
    {
      const double4 av = a_cur;
      const bool in = t < nA;
      sB[t >> 4][t & 15] = av.x;
      sB[t >> 4][16 + (t & 15)] = av.y;
      sB[t >> 4][32 + (t & 15)] = av.z;
      sB[t >> 4][48 + (t & 15)] = in ? 1.0 : qnan;
      sT[t] = in ? (tv - (av.x * av.x + av.y * av.y + av.z * av.z)) + delta : -infinity_v<double>();
    }
    {  // A operand column of side-2 cell t: lane (row t%16, k) of block t/16 holds [-2bx -2by -2bz |b|^2][k]
      const double4 bv = b_cur;
      sA[t >> 4][t & 15] = -2.0 * bv.x;
      sA[t >> 4][16 + (t & 15)] = -2.0 * bv.y;
      sA[t >> 4][32 + (t & 15)] = -2.0 * bv.z;
      sA[t >> 4][48 + (t & 15)] = bv.x * bv.x + bv.y * bv.y + bv.z * bv.z;
    }
    __syncthreads();
    if (threadIdx.x == 0) done += (unsigned long long)nA * (unsigned long long)nB;
    const uint64_t key_s = pack_idx((uint32_t)slot, 0u, 0u);
    // Appends the passing pairs of up to four products: product q covers side-1
    // block jq[q] and side-2 block rq[q]; bit 4q + r of `mask` = result r of q.
    auto append = [&](const mfma_v4d* d, const double* th, const int* jq, const int* rq, int cnt) {
      unsigned mask = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q < cnt)
#pragma unroll
          for (int r = 0; r < 4; ++r) mask |= (d[q][r] <= th[q] ? 1u : 0u) << (4 * q + r);
      for (;;) {
        const bool has = mask != 0u;
        const unsigned long long mk = __ballot(has);
        if (!mk) break;
        const int cnt2 = __popcll(mk);
        if (wcount + cnt2 > kStage) flush();
        if (has) {
          const int bit = __builtin_ctz(mask);
          mask &= mask - 1u;
          const int q = bit >> 2;
          const int jb = q == 0 ? jq[0] : q == 1 ? jq[1] : q == 2 ? jq[2] : jq[3];
          const int rb = q == 0 ? rq[0] : q == 1 ? rq[1] : q == 2 ? rq[2] : rq[3];
          const int ci = 16 * jb + col;        // side-1 cell (column)
          const int rr = 4 * (bit & 3) + kq;   // row in the side-2 block
          const int cj = 16 * rb + rr;         // side-2 cell
          // a from the B operands, b = -(1/2)(-2b) from the A operands (exact)
          const double ax = sB[jb][col], ay = sB[jb][16 + col], az = sB[jb][32 + col];
          const double bx = -0.5 * sA[rb][rr], by = -0.5 * sA[rb][16 + rr], bz = -0.5 * sA[rb][32 + rr];
          const double dq = pair_d2<double, false>(ax, ay, az, bx, by, bz);
          Cand c;
          c.hi = ord_bits(dq);
          c.lo = key_s | ((uint64_t)(uint32_t)ci << 20) | (uint64_t)(uint32_t)cj;
          stage[wv][wcount + __popcll(mk & ((1ull << lane) - 1))] = c;
        }
        wcount += cnt2;
      }
    };
    const mfma_v4d z = {0.0, 0.0, 0.0, 0.0};
    // Products of full blocks.  Wavefront w holds side-2 blocks w, w+4, ..
    // (NQ of them) and multiplies them with every full side-1 block.  The
    // products of the nL leftover side-2 blocks 4 NQ + i with the side-1
    // blocks, taken in the order (i, jb), are cut into four contiguous runs,
    // one per wavefront: a run covers at most two leftover blocks over disjoint
    // ranges of jb, so the wavefront walks jb in up to four segments, each a
    // loop with a fixed number of products (NQ, or NQ + 1 with that leftover
    // block), and every wavefront does the same number of products to within one.
    if (nBf > 0) {
      double aop[NQ > 0 ? NQ : 1];
#pragma unroll
      for (int q = 0; q < NQ; ++q) aop[q] = sA[wv + 4 * q][lane];
      auto run = [&](auto extra_c, int jb0, int jb1, int rbx) {
        constexpr bool EXTRA = decltype(extra_c)::value;
        if (NQ + (EXTRA ? 1 : 0) == 0 || jb0 >= jb1) return;
        const double ax = EXTRA ? sA[rbx][lane] : 0.0;
        double bop = sB[jb0][lane], thr = sT[16 * jb0 + col];
        for (int jb = jb0; jb < jb1; ++jb) {
          const double bop_n = sB[jb + 1][lane];  // next block (row 16 of sB is never written: read, unused)
          const double thr_n = sT[16 * (jb + 1) + col];
          mfma_v4d d[4];  // NQ + EXTRA <= 4 (append reads up to four)
#pragma unroll
          for (int q = 0; q < NQ; ++q) d[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[q], bop, z, 0, 0, 0);
          if (EXTRA) d[NQ < 4 ? NQ : 3] = __builtin_amdgcn_mfma_f64_16x16x4f64(ax, bop, z, 0, 0, 0);
          // all products issue before the minima; the minima are written as the
          // instruction, so the compiler does not count them as readers of the
          // products: 19 wait states (a DGEMM 16x16 result read by the vector
          // ALU) are inserted by hand after the last product
          __builtin_amdgcn_sched_barrier(0);
          asm volatile("s_nop 15\n\ts_nop 2" ::: "memory");
          __builtin_amdgcn_sched_barrier(0);
          double m = infinity_v<double>();
#pragma unroll
          for (int q = 0; q < NQ + (EXTRA ? 1 : 0); ++q)
            m = min_nan(m, min_nan(min_nan(d[q][0], d[q][1]), min_nan(d[q][2], d[q][3])));
          if (__ballot(m <= thr) != 0ull) {  // rare
            const double th[4] = {thr, thr, thr, thr};
            const int jq[4] = {jb, jb, jb, jb};
            int rq[4] = {wv, wv + 4, wv + 8, wv + 12};
            if (EXTRA && NQ < 4) rq[NQ < 4 ? NQ : 0] = rbx;
            append(d, th, jq, rq, NQ + (EXTRA ? 1 : 0));
          }
          bop = bop_n;
          thr = thr_n;
        }
      };
      using plain_t = std::integral_constant<bool, false>;
      using extra_t = std::integral_constant<bool, true>;
      const int nE = nL * nAf;
      const int e0 = (nE * wv) >> 2, e1 = (nE * (wv + 1)) >> 2;
      if (e1 <= e0) {
        run(plain_t(), 0, nAf, 0);
      } else {
        // run [e0, e1): leftover block i0 over jb [j0, min(nAf, j0 + len)), then block i0 + 1 over jb [0, r)
        const int i0 = e0 / nAf, j0 = e0 - i0 * nAf;
        const int j1 = min(nAf, j0 + (e1 - e0));
        const int r = (e1 - e0) - (j1 - j0);
        run(extra_t(), 0, r, 4 * NQ + i0 + 1);
        run(plain_t(), r, j0, 0);
        run(extra_t(), j0, j1, 4 * NQ + i0);
        run(plain_t(), j1, nAf, 0);
      }
    }
    // Pairs with a cell of a partial block (side-1 cells 16 nAf.., or side-2
    // cells 16 nBf..): the reference test d2 <= T on the vector ALU, one pair
    // per thread: each partial side-1 cell against side-2 cell t, then each
    // partial side-2 cell against full-block side-1 cell t.
    {
      const int rA = nA - 16 * nAf, rB = nB - 16 * nBf;
      for (int e = 0; e < rA + rB; ++e) {
        const bool s1p = e < rA;
        const int ci = s1p ? 16 * nAf + e : t;
        const int cj = s1p ? t : 16 * nBf + (e - rA);
        const bool valid = s1p ? t < nB : t < 16 * nAf;
        bool hit = false;
        double dq = 0.0;
        if (valid) {
          const double ax = sB[ci >> 4][ci & 15], ay = sB[ci >> 4][16 + (ci & 15)], az = sB[ci >> 4][32 + (ci & 15)];
          const double bx = -0.5 * sA[cj >> 4][cj & 15], by = -0.5 * sA[cj >> 4][16 + (cj & 15)],
                       bz = -0.5 * sA[cj >> 4][32 + (cj & 15)];
          dq = pair_d2<double, false>(ax, ay, az, bx, by, bz);
          hit = dq <= tv;
        }
        const unsigned long long mk = __ballot(hit);
        if (mk) {
          const int cnt = __popcll(mk);
          if (wcount + cnt > kStage) flush();
          if (hit) {
            Cand c;
            c.hi = ord_bits(dq);
            c.lo = key_s | ((uint64_t)(uint32_t)ci << 20) | (uint64_t)(uint32_t)cj;
            stage[wv][wcount + __popcll(mk & ((1ull << lane) - 1))] = c;
          }
          wcount += cnt;
        }
      }
    }
    if (nxt.slot >= 0) {
      a_cur = a_nxt;
      b_cur = b_nxt;
    }
    cur = nxt;
    nxt = nxt2;
  }
  if (wcount) flush();
  if (pairs_done) add_pairs_done(pairs_done, done);
}
template <int NQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NQ < 4 ? 4 : 3))) void k_dist_filter_mfma(const uint64_t* list, const unsigned* list_count,
                                                          int nslots, const int* su1, const int* su2, int w1, int w2,
                                                          const double4* C1, const double4* C2, const double* thresh,
                                                          Cand* out, unsigned long long* out_count,
                                                          unsigned long long cap, unsigned long long* pairs_done,
                                                          const unsigned long long* absmax) {
  k_dist_filter_mfma_body<NQ>(list, list_count, nslots, su1, su2, w1, w2, C1, C2, thresh, out, out_count, cap, pairs_done, absmax);
}

// FP32 screen of the refinement rounds (FP32SCREEN) on the matrix cores: the
// work split of k_dist_filter_mfma (same units, same records) with the
// products formed by v_mfma_f32_16x16x4f32, twice the FP64 rate, and the
// whole pass test folded into the product.  Each side is taken relative to an
// origin of its own (o1: side-1 cell 0, o2: side-2 cell 0; any points would
// do): a' = fl32(a - o1), b' = fl32(b - o2), c = fl64(o1 - o2), so that
//   |a' - b' + c|^2 - |c|^2 = P(a') + Q(b') - 2 a'.b',
//   P(a') = 2 c.a' + |a'|^2,  Q(b') = |b'|^2 - 2 c.b':
// only P, Q and the cross term, of the size of the units times |c| (not
// |c|^2), are rounded to float.  The instruction forms
//   D = [-2b' 1] . [a' fl32(P(a') - kd)] + fl32(Q(b'))
// (the side-2 row's Q in the accumulator: each lane holds four rows of one
// side-2 block, so its four accumulator values change only with that block),
// and a pair passes when D <= 0.  With u = 2^-24, u64 = 2^-53, the slot's
// bounds A >= |a'|, B >= |b'| (diagonals of the units' centre boxes: o1 and
// every side-1 centre lie in unit 1's box; L1 lengths),
//   R = sqrt(T (1 + 1e-12)) (1 + 1e-14) + 1.01 u (A + B) + 1.01 u64 |c|,
//   K = R^2 - |c|^2,  kd = fl32 rounded up (K + delta),
//   delta = (16 u (2 A B + B^2 + 2 |c| B + 2 |c| A + A^2 + |K|)
//            + 8 u64 (R^2 + |c|^2 + 2 |c| A + A^2)) (1 + 1e-6) + 1e-30.
// Every pair whose FP64 d2 is at most T (1 + 1e-13) passes: the roundings of
// a - o1, b - o2 and o1 - o2 move |a' - b' + c| from |a - b| by at most
// 1.01 u (|a'| + |b'|) + 1.01 u64 |c|, so |a' - b' + c| <= R and
// P + Q - 2 a'.b' <= K; P, Q and P - kd in float (c rounded to float, 3-term
// sums) and the 5-term sum of the instruction in any order and rounding add at
// most 16 u (2 A B + B^2 + 2 |c| B + 2 |c| A + A^2 + |K| + delta) (|P - kd| <=
// 2 |c| A + A^2 + |K| + delta), the FP64 parts a few u64 of their terms
// (1e-30: flushed denormals).  "Some result <= 0" is read from the integer
// minimum of the results' bits (negative floats are negative integers; a
// NaN's sign is not used, so a NaN may send a product to the exact test).  A
// passing pair's record holds its FP64 d2 from the FP64 centres (kept in
// LDS), so the records are those of k_dist_filter_mfma except near-threshold
// extras above T, which never reach the k smallest (at least k seed pairs
// have d2 <= T).  Result register r of lane l of this instruction is
// (row 4 (l/16) + r, column l%16).
template <int NQ>
__device__ __forceinline__ void k_dist_filter_mfma32_body(const uint64_t* list, const unsigned* list_count,
                                                            int nslots, const int* su1, const int* su2, int w1,
                                                            int w2, const double4* C1, const double4* C2,
                                                            const double4* lo1, const double4* hi1,
                                                            const double4* lo2, const double4* hi2,
                                                            const double* thresh, Cand* out,
                                                            unsigned long long* out_count, unsigned long long cap,
                                                            unsigned long long* pairs_done) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  constexpr int kStage = 128;
  __shared__ float sB[17][64];  // side-1 B operands [16 k + column]; one padding block for the prefetch
  __shared__ float sA[16][64];  // side-2 A operands [16 k + row]
  __shared__ float sQ[256];     // side-2 fl32(Q(b'))
  __shared__ double sCa[3][256], sCb[3][256];  // FP64 centres (records and partial blocks)
  __shared__ Cand stage[4][kStage];
  unsigned long long done = 0;
  const unsigned n = *list_count;
  const double tv = *thresh;
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  constexpr double u = 5.9604644775390625e-08;    // 2^-24
  constexpr double u64 = 1.1102230246251565e-16;  // 2^-53
  const double sqt = sqrt(tv * (1.0 + 1e-12)) * (1.0 + 1e-14);
  const int nA = w1 * w1, nB = w2 * w2;
  const int nAf = nA >> 4, nBf = nB >> 4;  // full blocks of 16 cells (side-1: columns, side-2: rows)
  const int nL = nBf - 4 * NQ;              // full side-2 blocks beyond the wavefronts' NQ each (0..3)
  const double qnan = quiet_nan<double>();
  const float qnanf = quiet_nan<float>();
  int wcount = 0;
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(out_count, (unsigned long long)wcount);
    base = __shfl(base, 0);
    for (int e = lane; e < wcount; e += 64)
      if (base + e < cap) out[base + e] = stage[wv][e];
    __builtin_amdgcn_wave_barrier();
    wcount = 0;
  };
  // the slot walk of k_dist_filter_mfma (slot ids two ahead, centres one ahead)
  constexpr int kMaxWalk = 16;
  const bool all = n == (unsigned)nslots;
  struct WalkSlot {
    int slot, u1, u2;  // slot -1: past the end of the walk
  };
  auto walk_slot = [&](int k) {
    WalkSlot w{-1, 0, 0};
    const unsigned i = blockIdx.x + (unsigned)k * gridDim.x;
    if (k < kMaxWalk && i < n) {
      w.slot = all ? (int)i : (int)(list[i] >> 32);
      w.u1 = su1[w.slot];
      w.u2 = su2[w.slot];
    }
    return w;
  };
  const int t = threadIdx.x;
  auto load_a = [&](int v1) {
    double4 a;
    a.x = a.y = a.z = a.w = qnan;
    if (t < nA) a = C1[(size_t)v1 * nA + t];
    return a;
  };
  auto load_b = [&](int v2) {
    double4 b;
    b.x = b.y = b.z = b.w = qnan;
    if (t < nB) b = C2[(size_t)v2 * nB + t];
    return b;
  };
  // bound on the diagonal of a unit's centre box (its L1 length, rounded up)
  auto diag = [](const double4& l, const double4& h) {
    return ((h.x - l.x) + (h.y - l.y) + (h.z - l.z)) * (1.0 + 1e-15);
  };
  WalkSlot cur = walk_slot(0), nxt = walk_slot(1);
  double4 a_cur = load_a(cur.u1), b_cur = load_b(cur.u2);
#pragma unroll 1
  for (int k = 0; cur.slot >= 0; ++k) {
    const int slot = cur.slot, u1 = cur.u1, u2 = cur.u2;
    double4 a_nxt, b_nxt;
    if (nxt.slot >= 0) {
      a_nxt = load_a(nxt.u1);
      b_nxt = load_b(nxt.u2);
    }
    const WalkSlot nxt2 = walk_slot(k + 2);
    // the slot's constants (one address per wavefront)
    const double4 o1 = C1[(size_t)u1 * nA], o2 = C2[(size_t)u2 * nB];
    const double A = diag(lo1[u1], hi1[u1]), B = diag(lo2[u2], hi2[u2]);
    const double cx = o1.x - o2.x, cy = o1.y - o2.y, cz = o1.z - o2.z;
    float kd;  // K + delta, rounded up to float
    {
      const double c2 = cx * cx + cy * cy + cz * cz;
      const double cn = (fabs(cx) + fabs(cy) + fabs(cz)) * (1.0 + 1e-15);  // >= |c|
      const double R = sqt + 1.01 * u * (A + B) + 1.01 * u64 * cn;
      const double R2 = R * R, K = R2 - c2;
      const double delta = (16.0 * u * (2.0 * A * B + B * B + 2.0 * cn * B + 2.0 * cn * A + A * A + fabs(K)) +
                            8.0 * u64 * (R2 + c2 + 2.0 * cn * A + A * A)) *
                               (1.0 + 1e-6) +
                           1e-30;
      kd = __double2float_ru(K + delta);
    }
    const float cxf = (float)cx, cyf = (float)cy, czf = (float)cz;
    __syncthreads();  // the previous slot's operands are no longer read
    {
      const double4 av = a_cur;
      const bool in = t < nA;
      const float ax = __double2float_rn(av.x - o1.x), ay = __double2float_rn(av.y - o1.y),
                  az = __double2float_rn(av.z - o1.z);
      const float a2 = fmaf(ax, ax, fmaf(ay, ay, az * az));
      const float pa = fmaf(2.f * cxf, ax, fmaf(2.f * cyf, ay, fmaf(2.f * czf, az, a2)));
      sB[t >> 4][t & 15] = ax;
      sB[t >> 4][16 + (t & 15)] = ay;
      sB[t >> 4][32 + (t & 15)] = az;
      sB[t >> 4][48 + (t & 15)] = in ? pa - kd : qnanf;
      sCa[0][t] = av.x;
      sCa[1][t] = av.y;
      sCa[2][t] = av.z;
    }
    {  // A operand column of side-2 cell t: lane (row t%16, k) of block t/16 holds [-2bx' -2by' -2bz' 1][k]
      const double4 bv = b_cur;
      const bool in = t < nB;
      const float bx = __double2float_rn(bv.x - o2.x), by = __double2float_rn(bv.y - o2.y),
                  bz = __double2float_rn(bv.z - o2.z);
      const float b2 = fmaf(bx, bx, fmaf(by, by, bz * bz));
      sA[t >> 4][t & 15] = -2.f * bx;
      sA[t >> 4][16 + (t & 15)] = -2.f * by;
      sA[t >> 4][32 + (t & 15)] = -2.f * bz;
      sA[t >> 4][48 + (t & 15)] = in ? 1.f : qnanf;
      sQ[t] = fmaf(-2.f * cxf, bx, fmaf(-2.f * cyf, by, fmaf(-2.f * czf, bz, b2)));
      sCb[0][t] = bv.x;
      sCb[1][t] = bv.y;
      sCb[2][t] = bv.z;
    }
    __syncthreads();
    if (threadIdx.x == 0) done += (unsigned long long)nA * (unsigned long long)nB;
    const uint64_t key_s = pack_idx((uint32_t)slot, 0u, 0u);
    auto record = [&](int ci, int cj, const unsigned long long mk, bool hit) {
      // FP64 d2 of cells ci, cj from the FP64 centres
      double dq = 0.0;
      if (hit) dq = pair_d2<double, false>(sCa[0][ci], sCa[1][ci], sCa[2][ci], sCb[0][cj], sCb[1][cj], sCb[2][cj]);
      const int cnt = __popcll(mk);
      if (wcount + cnt > kStage) flush();
      if (hit) {
        Cand c;
        c.hi = ord_bits(dq);
        c.lo = key_s | ((uint64_t)(uint32_t)ci << 20) | (uint64_t)(uint32_t)cj;
        stage[wv][wcount + __popcll(mk & ((1ull << lane) - 1))] = c;
      }
      wcount += cnt;
    };
    // Appends the passing pairs of up to four products (product q: side-1
    // block jb, side-2 block rq[q]; bit 4q + r of `mask` = result r of q).
    auto append = [&](const v4f* d, int jb, const int* rq, int cnt) {
      unsigned mask = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q < cnt)
#pragma unroll
          for (int r = 0; r < 4; ++r) mask |= (d[q][r] <= 0.f ? 1u : 0u) << (4 * q + r);
      for (;;) {
        const bool has = mask != 0u;
        const unsigned long long mk = __ballot(has);
        if (!mk) break;
        int ci = 0, cj = 0;
        if (has) {
          const int bit = __builtin_ctz(mask);
          mask &= mask - 1u;
          const int q = bit >> 2;
          const int rb = q == 0 ? rq[0] : q == 1 ? rq[1] : q == 2 ? rq[2] : rq[3];
          ci = 16 * jb + col;                 // side-1 cell (column)
          cj = 16 * rb + 4 * kq + (bit & 3);  // side-2 cell (row 4 (l/16) + r)
        }
        record(ci, cj, mk, has);
      }
    };
    // this lane's four rows of side-2 block rb: their Q, the accumulator
    auto acc_of = [&](int rb) {
      const float* q = &sQ[16 * rb + 4 * kq];
      const v4f c = {q[0], q[1], q[2], q[3]};
      return c;
    };
    if (nBf > 0) {
      float aop[NQ > 0 ? NQ : 1];
      v4f cq[NQ > 0 ? NQ : 1];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        aop[q] = sA[wv + 4 * q][lane];
        cq[q] = acc_of(wv + 4 * q);
      }
      auto run = [&](auto extra_c, int jb0, int jb1, int rbx) {
        // (NQ = 4 leaves no leftover side-2 blocks: that form never runs with EXTRA)
        constexpr bool EXTRA = decltype(extra_c)::value && NQ < 4;
        constexpr int NP = NQ + (EXTRA ? 1 : 0);  // products per side-1 block, <= 4
        if (NP == 0 || jb0 >= jb1) return;
        const float ax = EXTRA ? sA[rbx][lane] : 0.f;
        const v4f cx4 = EXTRA ? acc_of(rbx) : v4f{0.f, 0.f, 0.f, 0.f};
        float bop = sB[jb0][lane];
        for (int jb = jb0; jb < jb1; ++jb) {
          const float bop_n = sB[jb + 1][lane];  // next block (row 16 of sB: read, unused)
          v4f d[4];
#pragma unroll
          for (int q = 0; q < NQ; ++q) d[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop[q], bop, cq[q], 0, 0, 0);
          if (EXTRA) d[EXTRA ? NP - 1 : 0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax, bop, cx4, 0, 0, 0);
          int mn = INT_MAX;
#pragma unroll
          for (int q = 0; q < NP; ++q)
            mn = min(mn, min(min(__float_as_int(d[q][0]), __float_as_int(d[q][1])),
                             min(__float_as_int(d[q][2]), __float_as_int(d[q][3]))));
          if (__ballot(mn <= 0) != 0ull) {  // rare
            int rq[4] = {wv, wv + 4, wv + 8, wv + 12};
            if (EXTRA) rq[EXTRA ? NP - 1 : 0] = rbx;
            append(d, jb, rq, NP);
          }
          bop = bop_n;
        }
      };
      using plain_t = std::integral_constant<bool, false>;
      using extra_t = std::integral_constant<bool, true>;
      const int nE = nL * nAf;
      const int e0 = (nE * wv) >> 2, e1 = (nE * (wv + 1)) >> 2;
      if (e1 <= e0) {
        run(plain_t(), 0, nAf, 0);
      } else {
        const int i0 = e0 / nAf, j0 = e0 - i0 * nAf;
        const int j1 = min(nAf, j0 + (e1 - e0));
        const int r = (e1 - e0) - (j1 - j0);
        run(extra_t(), 0, r, 4 * NQ + i0 + 1);
        run(plain_t(), r, j0, 0);
        run(extra_t(), j0, j1, 4 * NQ + i0);
        run(plain_t(), j1, nAf, 0);
      }
    }
    // pairs with a cell of a partial block: the reference test d2 <= T in FP64
    {
      const int rA = nA - 16 * nAf, rB = nB - 16 * nBf;
      for (int e = 0; e < rA + rB; ++e) {
        const bool s1p = e < rA;
        const int ci = s1p ? 16 * nAf + e : t;
        const int cj = s1p ? t : 16 * nBf + (e - rA);
        const bool valid = s1p ? t < nB : t < 16 * nAf;
        bool hit = false;
        if (valid)
          hit = pair_d2<double, false>(sCa[0][ci], sCa[1][ci], sCa[2][ci], sCb[0][cj], sCb[1][cj], sCb[2][cj]) <= tv;
        const unsigned long long mk = __ballot(hit);
        if (mk) record(ci, cj, mk, hit);
      }
    }
    if (nxt.slot >= 0) {
      a_cur = a_nxt;
      b_cur = b_nxt;
    }
    cur = nxt;
    nxt = nxt2;
  }
  if (wcount) flush();
  if (pairs_done) add_pairs_done(pairs_done, done);
}
template <int NQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_dist_filter_mfma32(
    const uint64_t* list, const unsigned* list_count, int nslots, const int* su1, const int* su2, int w1, int w2,
    const double4* C1, const double4* C2, const double4* lo1, const double4* hi1, const double4* lo2,
    const double4* hi2, const double* thresh, Cand* out, unsigned long long* out_count, unsigned long long cap,
    unsigned long long* pairs_done) {
  k_dist_filter_mfma32_body<NQ>(list, list_count, nslots, su1, su2, w1, w2, C1, C2, lo1, hi1, lo2, hi2, thresh,
                                out, out_count, cap, pairs_done);
}

// ---------------------------------------------------------------------------
// FP32 screening with an FP64 result (FP32SCREEN).  Centres are evaluated in
// FP64 and rounded to float for the screen; the screen's threshold is widened
// by a bound on every rounding step so that its candidate set contains every
// pair the FP64 ranking can select, and the candidates are re-ranked with
// their FP64 d2.
// ---------------------------------------------------------------------------
// T <- (sqrt(T)(1+16u) + 4 sqrt(3) E (1+16u))^2 (1+16u) + 1e-36, u = 2^-24,
// E = 2^-24 max|centre coordinate| (the largest of the kAmaxSlots partial
// maxima): one widening covers (a) the float
// rounding of both centres (<= E per coordinate, <= 2 sqrt(3) E in distance),
// (b) the float evaluation of d2 (<= 5u relative) and (c) the double
// evaluation of d2 (<= 5 * 2^-53 relative), applied once to bound the FP64
// k-th d2 from k float values and once more to reach every pair below it.
__device__ __forceinline__ void k_inflate_body(float* thresh, const unsigned long long* absmax_bits) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const float t = *thresh;
  if (!(t < __uint_as_float(0x7f800000u))) return;  // +inf stays +inf
  const double u = 5.9604644775390625e-08;         // 2^-24
  unsigned long long mb = 0;
  for (int x = 0; x < kAmaxSlots; ++x) mb = absmax_bits[x] > mb ? absmax_bits[x] : mb;
  const double E = __longlong_as_double((long long)mb) * u;
  const double g = 1.0 + 16.0 * u;
  const double r = sqrt((double)t) * g + 4.0 * 1.7320508075688772 * E * g;
  *thresh = __double2float_ru(r * r * g + 1e-36);
}
__global__ void k_inflate(float* thresh, const unsigned long long* absmax_bits) {
  k_inflate_body(thresh, absmax_bits);
}

__device__ __forceinline__ void rerank_one(Cand* cand, unsigned long long e, const int* su1, const int* su2,
                                           int cells1, int cells2, const double4* C1, const double4* C2);

// Candidate records get their FP64 d2 (same recipe as the FP64 screen).  The
// record count is min(*n_dev, cap).
__device__ __forceinline__ void k_rerank_body(Cand* cand, const unsigned long long* n_dev, unsigned long long cap, const int* su1,
                         const int* su2, int cells1, int cells2, const double4* C1, const double4* C2) {
  const unsigned long long n = min(*n_dev, cap);
  for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (unsigned long long)gridDim.x * blockDim.x)
    rerank_one(cand, e, su1, su2, cells1, cells2, C1, C2);
}
__global__ void k_rerank(Cand* cand, const unsigned long long* n_dev, unsigned long long cap, const int* su1,
                         const int* su2, int cells1, int cells2, const double4* C1, const double4* C2) {
  k_rerank_body(cand, n_dev, cap, su1, su2, cells1, cells2, C1, C2);
}

__device__ __forceinline__ void rerank_one(Cand* cand, unsigned long long e, const int* su1, const int* su2,
                                           int cells1, int cells2, const double4* C1, const double4* C2) {
  Cand c = cand[e];
  const int slot = (int)(c.lo >> 40), i = (int)((c.lo >> 20) & 0xfffff), j = (int)(c.lo & 0xfffff);
  const double4 a = C1[(size_t)su1[slot] * cells1 + i];
  const double4 b = C2[(size_t)su2[slot] * cells2 + j];
  const double dx = a.x - b.x;
  const double dy = a.y - b.y;
  const double dz = a.z - b.z;
  const double d2 = dx * dx + dy * dy + dz * dz;
  c.hi = (unsigned long long)__double_as_longlong(d2);
  cand[e] = c;
}

// ---------------------------------------------------------------------------
// Operator ABI (mask semantics of jsi_ascend_distance_topk_mask /
// jsi_ascend_intersection_mask, hotspots_910b/src/jsi_ascend_hotspots.cpp:33-113):
// flat AoS centre arrays are laid out on a w x w cell grid padded with +inf
// centres (radius 0), which never pass a finite threshold or an overlap test.
// ---------------------------------------------------------------------------
__global__ void k_pack_cells(const float* __restrict__ c, const float* __restrict__ r, long long n, int w,
                             float4* __restrict__ C, float4* __restrict__ R) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (long long)w * w) return;
  const float inf = __uint_as_float(0x7f800000u);
  float4 a, b;
  if (x < n) {
    a.x = c[3 * x];
    a.y = c[3 * x + 1];
    a.z = c[3 * x + 2];
    if (r) {
      b.x = r[3 * x];
      b.y = r[3 * x + 1];
      b.z = r[3 * x + 2];
    } else {
      b.x = b.y = b.z = 0.f;
    }
  } else {
    a.x = a.y = a.z = inf;
    b.x = b.y = b.z = 0.f;
  }
  a.w = b.w = 0.f;
  C[x] = a;
  if (R) R[x] = b;
}

// MODE 0: distance, mark both endpoints of every pair with d2 <= *thresh.
// MODE 1: overlap, mark both endpoints of every overlapping pair.
template <int MODE>
__global__ __launch_bounds__(256) void k_pair_mask(const uint64_t* list, const unsigned long long* list_count,
                                                   const unsigned* list_count32, int w1, int w2, const float4* C1, const float4* R1,
                                                   const float4* C2, const float4* R2, const float* thresh,
                                                   long long n1, long long n2, int* m1, int* m2) {
  __shared__ float4 sC[kTileCells];
  __shared__ float4 sR[kTileCells];
  __shared__ int sJ[kTileCells];
  const unsigned long long n = list_count ? *list_count : (unsigned long long)*list_count32;
  const float tv = MODE == 0 ? *thresh : 0.f;
  for (unsigned long long it = blockIdx.x; it < n; it += gridDim.x) {
    const uint64_t e = list[it];
    const int t1 = (int)((e >> 16) & 0xffff), t2 = (int)(e & 0xffff);
    const TileGeo g1 = tile_geo(t1, w1), g2 = tile_geo(t2, w2);
    const int nA = g1.ht * g1.wd, nB = g2.ht * g2.wd;
    __syncthreads();
    if ((int)threadIdx.x < nB) {
      const int j = (g2.row0 + threadIdx.x / g2.wd) * w2 + g2.col0 + threadIdx.x % g2.wd;
      sC[threadIdx.x] = C2[j];
      if (MODE == 1) sR[threadIdx.x] = R2[j];
      sJ[threadIdx.x] = j;
    }
    const bool va = (int)threadIdx.x < nA;
    int i = 0;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), ra = a;
    if (va) {
      i = (g1.row0 + threadIdx.x / g1.wd) * w1 + g1.col0 + threadIdx.x % g1.wd;
      a = C1[i];
      if (MODE == 1) ra = R1[i];
    }
    __syncthreads();
    if (!va || i >= n1) continue;
    bool any = false;
    for (int jj = 0; jj < nB; ++jj) {
      const float4 b = sC[jj];
      bool hit;
      if (MODE == 0) {
        const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        hit = dx * dx + dy * dy + dz * dz <= tv;
      } else {
        const float4 rb = sR[jj];
        hit = fabsf(a.x - b.x) < (ra.x + rb.x) && fabsf(a.y - b.y) < (ra.y + rb.y) &&
              fabsf(a.z - b.z) < (ra.z + rb.z);
      }
      if (hit && sJ[jj] < n2) {
        any = true;
        m2[sJ[jj]] = 1;
      }
    }
    if (any) m1[i] = 1;
  }
}

template <class T>
__device__ __forceinline__ void k_set_thresh_body(T* thresh, const Cand* rec, int idx, int use_inf) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *thresh = use_inf ? infinity_v<T>() : from_bits<T>(rec[idx].hi);
}
template <class T>
__global__ void k_set_thresh(T* thresh, const Cand* rec, int idx, int use_inf) {
  k_set_thresh_body<T>(thresh, rec, idx, use_inf);
}

// ---------------------------------------------------------------------------
// Top-k selection without a full sort: histogram radix select on the d2 key.
// The key range starts at the smallest and largest key present (one reduction
// pass); each iteration histograms [lo, hi] into kSelBins linear bins and
// narrows it to the bin holding rank k, until at most small_cap records have
// key <= hi (or the bin is a single key).  Those records are compacted and
// sorted.  hi is always an upper bound of the k-th key with at least k
// records at or below it, which is what a screening threshold needs.
// ---------------------------------------------------------------------------
constexpr int kSelBins = 4096;

__device__ __forceinline__ unsigned long long sel_width(unsigned long long lo, unsigned long long hi) {
  return (hi > lo ? hi - lo : 0ull) / kSelBins + 1;
}

__device__ __forceinline__ unsigned long long sel_n(unsigned long long n, const unsigned long long* n_dev) {
  return n_dev ? min(*n_dev, n) : n;
}

__device__ __forceinline__ void k_sel_init_body(SelState* st, unsigned long long n,
                                                  const unsigned long long* n_dev, unsigned long long k,
                                                  unsigned long long maxkey, unsigned small_cap, unsigned* hist,
                                                  unsigned long long* out_count) {
  for (int b = threadIdx.x; b < kSelBins; b += blockDim.x) hist[b] = 0;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (out_count) *out_count = 0;
  n = sel_n(n, n_dev);
  SelState s;
  s.lo = ~0ull;  // min / max key, reduced by k_sel_minmax
  s.hi = 0;
  s.width = 0;
  s.need = k < n ? k : n;
  s.below = 0;
  s.final_hi = maxkey;
  s.done = (n <= small_cap || s.need == 0) ? 1 : 0;
  *st = s;
}
__global__ __launch_bounds__(256) void k_sel_init(SelState* st, unsigned long long n,
                                                  const unsigned long long* n_dev, unsigned long long k,
                                                  unsigned long long maxkey, unsigned small_cap, unsigned* hist,
                                                  unsigned long long* out_count) {
  k_sel_init_body(st, n, n_dev, k, maxkey, small_cap, hist, out_count);
}

__device__ __forceinline__ void k_sel_minmax_body(const Cand* cand, unsigned long long n,
                                                    const unsigned long long* n_dev, SelState* st) {
  __shared__ unsigned long long rmin[4], rmax[4];
  if (st->done) return;
  n = sel_n(n, n_dev);
  if ((unsigned long long)blockIdx.x * blockDim.x >= n) return;  // whole block: no records
  unsigned long long a = ~0ull, b = 0ull;
  for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long key = cand[e].hi;
    a = key < a ? key : a;
    b = key > b ? key : b;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long x = __shfl_xor(a, off), y = __shfl_xor(b, off);
    a = x < a ? x : a;
    b = y > b ? y : b;
  }
  if (lane_id() == 0) {
    rmin[threadIdx.x >> 6] = a;
    rmax[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      a = rmin[w] < a ? rmin[w] : a;
      b = rmax[w] > b ? rmax[w] : b;
    }
    atomicMin(&st->lo, a);
    atomicMax(&st->hi, b);
  }
}
__global__ __launch_bounds__(256) void k_sel_minmax(const Cand* cand, unsigned long long n,
                                                    const unsigned long long* n_dev, SelState* st) {
  k_sel_minmax_body(cand, n, n_dev, st);
}

__device__ __forceinline__ void k_sel_hist_body(const Cand* cand, unsigned long long n,
                                                  const unsigned long long* n_dev, const SelState* st,
                                                  unsigned* hist) {
  __shared__ unsigned h[kSelBins];
  if (st->done) return;
  n = sel_n(n, n_dev);
  if ((unsigned long long)blockIdx.x * blockDim.x >= n) return;  // whole block: no records
  const unsigned long long lo = st->lo, hi = st->hi, width = sel_width(lo, hi);
  for (int b = threadIdx.x; b < kSelBins; b += blockDim.x) h[b] = 0;
  __syncthreads();
  for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long key = cand[e].hi;
    if (key >= lo && key <= hi) atomicAdd(&h[(key - lo) / width], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kSelBins; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}
__global__ __launch_bounds__(256) void k_sel_hist(const Cand* cand, unsigned long long n,
                                                  const unsigned long long* n_dev, const SelState* st,
                                                  unsigned* hist) {
  k_sel_hist_body(cand, n, n_dev, st, hist);
}

__device__ __forceinline__ void k_sel_step_body(SelState* st, unsigned* hist, unsigned small_cap) {
  __shared__ unsigned long long cum[1024];
  __shared__ int found;
  __shared__ unsigned long long found_before;
  SelState s = *st;
  if (s.done) return;
  const unsigned long long width = sel_width(s.lo, s.hi);
  constexpr int per = kSelBins / 1024;
  unsigned long long loc = 0;
  for (int q = 0; q < per; ++q) loc += hist[threadIdx.x * per + q];
  cum[threadIdx.x] = loc;
  if (threadIdx.x == 0) found = kSelBins;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned long long v = threadIdx.x >= (unsigned)off ? cum[threadIdx.x - off] : 0ull;
    __syncthreads();
    cum[threadIdx.x] += v;
    __syncthreads();
  }
  // first bin whose inclusive count reaches `need`
  unsigned long long run = cum[threadIdx.x] - loc;
  for (int q = 0; q < per; ++q) {
    const unsigned c = hist[threadIdx.x * per + q];
    if (run < s.need && run + c >= s.need) {
      found = threadIdx.x * per + q;
      found_before = run;
    }
    run += c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int b = found < kSelBins ? found : kSelBins - 1;
    const unsigned long long before = found < kSelBins ? found_before : 0ull;
    const unsigned long long nlo = s.lo + (unsigned long long)b * width;
    const unsigned long long span_hi = nlo + width - 1;
    const unsigned long long nhi = span_hi < s.hi ? span_hi : s.hi;
    s.below += before;
    s.need -= before;
    s.lo = nlo;
    s.hi = nhi;
    s.final_hi = nhi;
    if (s.below + hist[b] <= small_cap || nlo == nhi) s.done = 1;
    *st = s;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kSelBins; b += 1024) hist[b] = 0;
}
__global__ __launch_bounds__(1024) void k_sel_step(SelState* st, unsigned* hist, unsigned small_cap) {
  k_sel_step_body(st, hist, small_cap);
}

__device__ __forceinline__ void k_sel_compact_body(const Cand* cand, unsigned long long n,
                                                     const unsigned long long* n_dev, const SelState* st, Cand* out,
                                                     unsigned long long* out_count, unsigned long long cap) {
  n = sel_n(n, n_dev);
  const unsigned long long fh = st->final_hi;
  for (unsigned long long e0 = (unsigned long long)blockIdx.x * blockDim.x; e0 < n;
       e0 += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long e = e0 + threadIdx.x;
    bool keep = false;
    Cand c;
    if (e < n) {
      c = cand[e];
      keep = c.hi <= fh;
    }
    const unsigned long long m = __ballot(keep);
    if (m) {
      unsigned long long base = 0;
      const int leader = __ffsll((long long)m) - 1;
      if (lane_id() == leader) base = atomicAdd(out_count, (unsigned long long)__popcll(m));
      base = __shfl(base, leader);
      const unsigned long long pos = base + __popcll(m & ((1ull << lane_id()) - 1));
      if (keep && pos < cap) out[pos] = c;
    }
  }
}
__global__ __launch_bounds__(256) void k_sel_compact(const Cand* cand, unsigned long long n,
                                                     const unsigned long long* n_dev, const SelState* st, Cand* out,
                                                     unsigned long long* out_count, unsigned long long cap) {
  k_sel_compact_body(cand, n, n_dev, st, out, out_count, cap);
}

template <class T>
__global__ void k_thresh_from_state(T* thresh, const SelState* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *thresh = from_bits<T>(st->final_hi);
}

// Float screening threshold from a bound on FP64 keys: rounded up, so every
// pair whose FP64 d2 is <= the bound stays below it before the widening.
__global__ void k_thresh_from_state_ru(float* thresh, const SelState* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *thresh = __double2float_ru(from_bits<double>(st->final_hi));
}
__global__ void k_set_thresh_ru(float* thresh, const Cand* rec, int idx) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *thresh = __double2float_ru(from_bits<double>(rec[idx].hi));
}

// Start of a distance query: both sides' round-1 regions, slot 0 -> unit 0,
// and the max |coordinate| slots cleared (one launch instead of five).
template <class T>
__device__ __forceinline__ void k_query_init_body(vec4_t<T>* reg1, vec4_t<T> g1, vec4_t<T>* reg2, vec4_t<T> g2, int* su1, int* su2,
                             unsigned long long* amax) {
  if (amax && threadIdx.x < kAmaxSlots) amax[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) {
    *reg1 = g1;
    *reg2 = g2;
    *su1 = 0;
    *su2 = 0;
  }
}
template <class T>
__global__ void k_query_init(vec4_t<T>* reg1, vec4_t<T> g1, vec4_t<T>* reg2, vec4_t<T> g2, int* su1, int* su2,
                             unsigned long long* amax) {
  k_query_init_body<T>(reg1, g1, reg2, g2, su1, su2, amax);
}

// Small device-side store (no host round trip): one record.
__global__ void k_copy_cand(Cand* dst, const Cand* src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}


// ---------------------------------------------------------------------------
// Bitonic sort of candidate records (power-of-two length, >= kSortBlock).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool cand_less(const Cand& a, const Cand& b) {
  return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}

__global__ void k_fill_sentinel(Cand* a, unsigned long long from, unsigned long long n2) {
  const unsigned long long i = from + (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) {
    Cand c;
    c.hi = ~0ull;
    c.lo = ~0ull;
    a[i] = c;
  }
}

__global__ __launch_bounds__(1024) void k_bsort_local(Cand* a, unsigned long long merge_k) {
  __shared__ Cand s[kSortBlock];
  const unsigned long long base = (unsigned long long)blockIdx.x * kSortBlock;
  for (int t = threadIdx.x; t < kSortBlock; t += blockDim.x) s[t] = a[base + t];
  __syncthreads();
  auto step = [&](unsigned long long k, int j) {
    for (int t = threadIdx.x; t < kSortBlock / 2; t += blockDim.x) {
      const int i = 2 * t - (t & (j - 1));
      const int l = i + j;
      const bool asc = ((base + i) & k) == 0;
      const Cand x = s[i], y = s[l];
      const bool swap = asc ? cand_less(y, x) : cand_less(x, y);
      if (swap) {
        s[i] = y;
        s[l] = x;
      }
    }
    __syncthreads();
  };
  if (merge_k == 0) {
    for (int k = 2; k <= kSortBlock; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) step((unsigned long long)k, j);
  } else {
    for (int j = kSortBlock >> 1; j > 0; j >>= 1) step(merge_k, j);
  }
  for (int t = threadIdx.x; t < kSortBlock; t += blockDim.x) a[base + t] = s[t];
}

// Whole sort of n2 <= kBlockSortMax records in one workgroup's LDS (128 KiB).
constexpr int kBlockSortMax = 8192;
__global__ __launch_bounds__(1024) void k_bsort_single(Cand* a, int n2) {
  extern __shared__ Cand sh[];
  for (int t = threadIdx.x; t < n2; t += blockDim.x) sh[t] = a[t];
  __syncthreads();
  for (int k = 2; k <= n2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < n2 / 2; t += blockDim.x) {
        const int i = 2 * t - (t & (j - 1)), l = i + j;
        const bool asc = (i & k) == 0;
        const Cand x = sh[i], y = sh[l];
        if (asc ? cand_less(y, x) : cand_less(x, y)) {
          sh[i] = y;
          sh[l] = x;
        }
      }
      __syncthreads();
    }
  for (int t = threadIdx.x; t < n2; t += blockDim.x) a[t] = sh[t];
}

// Rank sort of m distinct records (the keys (d2, slot, i, j) are unique):
// each record's position is the number of records smaller than it.  Only
// positions below `keep` are written.  A workgroup ranks 64 records (one per
// lane); each of its 16 wavefronts counts against 1/16 of the set, staged 64
// records at a time in LDS (broadcast reads), and the partial ranks are
// summed in LDS.  O(m^2) compares spread over m/64 * 16 wavefronts.
constexpr int kRankWaves = 16;
__device__ __forceinline__ void k_rank_sort_body(const Cand* in, const unsigned long long* m_dev, unsigned m_cap,
                                                    unsigned keep, Cand* out, Cand* last_out) {
  __shared__ Cand stage[kRankWaves][64];
  __shared__ unsigned part[kRankWaves][64];
  const unsigned m = (unsigned)min((unsigned long long)m_cap, *m_dev);
  if (blockIdx.x * 64u >= m) return;  // whole workgroup
  const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const unsigned x = blockIdx.x * 64u + lane;
  Cand mine;
  mine.hi = mine.lo = ~0ull;
  if (x < m) mine = in[x];
  const unsigned slice = (m + kRankWaves - 1) / kRankWaves;
  const unsigned s0 = w * slice, s1 = min(m, s0 + slice);
  unsigned rank = 0;
  for (unsigned c0 = 0; c0 < slice; c0 += 64) {  // same trip count in every wavefront
    const unsigned t = s0 + c0 + lane;
    Cand r;
    r.hi = r.lo = ~0ull;  // never below a real record (real keys have d2 bits < ~0)
    if (t < s1) r = in[t];
    stage[w][lane] = r;
    __syncthreads();
    const unsigned cn = min(64u, slice - c0);
    for (unsigned q = 0; q < cn; ++q) rank += cand_less(stage[w][q], mine) ? 1u : 0u;
    __syncthreads();
  }
  part[w][lane] = rank;
  __syncthreads();
  if (w == 0 && x < m) {
    unsigned r = 0;
#pragma unroll
    for (int v = 0; v < kRankWaves; ++v) r += part[v][lane];
    if (r < keep) out[r] = mine;
    if (last_out && r == min(keep, m) - 1) *last_out = mine;  // the last kept record
  }
}
// The same ordering for the batched form, where 64 queries' rank sorts would
// each do their m^2 compares: runs of kRun records are sorted in LDS (a
// bitonic network, one workgroup per run), then each record's position is
// its index in its own run plus, for every other run, the number of that
// run's records below it (a binary search; the keys are unique, so the
// positions are a permutation and equal the rank sort's).  O(m log m) work.
// `runs` holds the sorted runs (m_cap rounded up to kRun records); positions
// past m in the last run hold sentinels above every real key.
constexpr int kRun = 1024;
__device__ __forceinline__ void k_run_sort_body(const Cand* in, const unsigned long long* m_dev, unsigned m_cap,
                                                Cand* runs) {
  __shared__ Cand s[kRun];
  const unsigned m = (unsigned)min((unsigned long long)m_cap, *m_dev);
  const unsigned base = blockIdx.x * kRun;
  if (base >= m) return;  // whole workgroup
  const int t = threadIdx.x;  // kRun / 2 threads, two records each
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    Cand c;
    c.hi = c.lo = ~0ull;
    if (base + t + h * (kRun / 2) < m) c = in[base + t + h * (kRun / 2)];
    s[t + h * (kRun / 2)] = c;
  }
  __syncthreads();
  for (int k = 2; k <= kRun; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int i = 2 * t - (t & (j - 1)), l = i + j;
      const bool asc = (i & k) == 0;
      const Cand x = s[i], y = s[l];
      if (asc ? cand_less(y, x) : cand_less(x, y)) {
        s[i] = y;
        s[l] = x;
      }
      __syncthreads();
    }
  runs[base + t] = s[t];
  runs[base + t + kRun / 2] = s[t + kRun / 2];
}

__device__ __forceinline__ void k_run_rank_body(const Cand* runs, const unsigned long long* m_dev, unsigned m_cap,
                                                unsigned keep, Cand* out, Cand* last_out) {
  const unsigned m = (unsigned)min((unsigned long long)m_cap, *m_dev);
  const unsigned x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= m) return;
  const Cand c = runs[x];
  const unsigned r = x / kRun, nruns = (m + kRun - 1) / kRun;
  unsigned rank = x - r * kRun;
  for (unsigned q = 0; q < nruns; ++q) {
    if (q == r) continue;
    const Cand* R = runs + (size_t)q * kRun;
    unsigned lo = 0, hi = min((unsigned)kRun, m - q * kRun);
    while (lo < hi) {
      const unsigned mid = (lo + hi) >> 1;
      if (cand_less(R[mid], c))
        lo = mid + 1;
      else
        hi = mid;
    }
    rank += lo;
  }
  if (rank < keep) out[rank] = c;
  if (last_out && rank == min(keep, m) - 1) *last_out = c;  // the last kept record
}

__global__ __launch_bounds__(1024) void k_rank_sort(const Cand* in, const unsigned long long* m_dev, unsigned m_cap,
                                                    unsigned keep, Cand* out, Cand* last_out) {
  k_rank_sort_body(in, m_dev, m_cap, keep, out, last_out);
}

__global__ void k_bsort_global(Cand* a, unsigned long long n2, unsigned long long k, unsigned long long j) {
  const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n2 / 2) return;
  const unsigned long long i = 2 * t - (t & (j - 1));
  const unsigned long long l = i + j;
  const bool asc = (i & k) == 0;
  const Cand x = a[i], y = a[l];
  const bool swap = asc ? cand_less(y, x) : cand_less(x, y);
  if (swap) {
    a[i] = y;
    a[l] = x;
  }
}

// ---------------------------------------------------------------------------
// Distance refinement bookkeeping (minimum_distance, distance_analysis.cu:400-491,
// JSI_CAD_PRESERVE_PAIR_RELATION=1): jobs are the sorted top-k pairs; each side
// keeps one unit per distinct (slot, cell) in order of first appearance; slot k
// of the next round is (unit of job k's side-1 endpoint, unit of its side-2
// endpoint).  Child region: child_region (jsi_device.hpp).
// ---------------------------------------------------------------------------

constexpr int kDedupMax = 8192;
constexpr int kDedupPer = kDedupMax / 1024;  // records per thread of k_dedup_units

__device__ __forceinline__ void dedup_decode(const Cand* rec, int e, int* s, int* i, int* j) {
  const uint64_t lo = rec[e].lo;
  *s = (int)(lo >> 40);
  *i = (int)((lo >> 20) & 0xfffff);
  *j = (int)(lo & 0xfffff);
}

// Pass 1: firstpos[slot * cells + cell] = smallest record index with that
// (slot, cell), per side (tables hold INT_MAX between uses).
__device__ __forceinline__ void k_dedup_mark_body(const Cand* rec, int n, int cells1, int cells2, int* fp1,
                                                    int* fp2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int s, i, j;
  dedup_decode(rec, e, &s, &i, &j);
  atomicMin(&fp1[(long long)s * cells1 + i], e);
  atomicMin(&fp2[(long long)s * cells2 + j], e);
}
__global__ __launch_bounds__(256) void k_dedup_mark(const Cand* rec, int n, int cells1, int cells2, int* fp1,
                                                    int* fp2) {
  k_dedup_mark_body(rec, n, cells1, cells2, fp1, fp2);
}

// Pass 2 (single workgroup, contiguous chunks of <= 8 records per thread, all
// loads issued up front): a record whose firstpos is itself opens a new unit;
// unit ids follow record order (first appearance); every record's next-round
// slot map entry is the unit of its first appearance.
__device__ __forceinline__ void k_dedup_units_body(const Cand* rec, int n, int cells1, int cells2,
                                                      const int* fp1, const int* fp2, int* nsu1, int* nsu2,
                                                      unsigned char* isfirst, int* nunits, unsigned long long* zero,
                                                      int nzero) {
  if ((int)threadIdx.x < nzero) zero[threadIdx.x] = 0ull;  // next round's max |coordinate| slots
  __shared__ int sUnit[kDedupMax];
  __shared__ int sScan[1024];
  const int per = (n + 1023) / 1024;
  const int e0 = threadIdx.x * per;
  int f[2][kDedupPer];
#pragma unroll
  for (int q = 0; q < kDedupPer; ++q) {
    f[0][q] = f[1][q] = -1;
    const int e = e0 + q;
    if (q < per && e < n) {
      int s, i, j;
      dedup_decode(rec, e, &s, &i, &j);
      f[0][q] = fp1[(long long)s * cells1 + i];
      f[1][q] = fp2[(long long)s * cells2 + j];
    }
  }
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    int local = 0;
#pragma unroll
    for (int q = 0; q < kDedupPer; ++q) local += f[side][q] == e0 + q ? 1 : 0;
    sScan[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int v = threadIdx.x >= (unsigned)off ? sScan[threadIdx.x - off] : 0;
      __syncthreads();
      sScan[threadIdx.x] += v;
      __syncthreads();
    }
    int u = sScan[threadIdx.x] - local;
#pragma unroll
    for (int q = 0; q < kDedupPer; ++q) {
      const int e = e0 + q;
      if (f[side][q] < 0) continue;
      const bool first = f[side][q] == e;
      if (first) sUnit[e] = u++;
      isfirst[side * kDedupMax + e] = first ? 1 : 0;
    }
    if (threadIdx.x == 1023) nunits[side] = sScan[1023];
    __syncthreads();
    int* nsu = side == 0 ? nsu1 : nsu2;
#pragma unroll
    for (int q = 0; q < kDedupPer; ++q)
      if (f[side][q] >= 0) nsu[e0 + q] = sUnit[f[side][q]];
    __syncthreads();  // sUnit is reused by the next side
  }
}
__global__ __launch_bounds__(1024) void k_dedup_units(const Cand* rec, int n, int cells1, int cells2,
                                                      const int* fp1, const int* fp2, int* nsu1, int* nsu2,
                                                      unsigned char* isfirst, int* nunits, unsigned long long* zero,
                                                      int nzero) {
  k_dedup_units_body(rec, n, cells1, cells2, fp1, fp2, nsu1, nsu2, isfirst, nunits, zero, nzero);
}

// Pass 3: child regions of the new units, and firstpos back to INT_MAX.
template <class T>
__device__ __forceinline__ void k_dedup_apply_body(const Cand* rec, int n, int cells1, int cells2, const int* su1,
                                                     const int* su2, const vec4_t<T>* reg1, const vec4_t<T>* reg2,
                                                     int pnuv1, int pnuv2, const unsigned char* isfirst,
                                                     const int* nsu1, const int* nsu2, vec4_t<T>* nreg1,
                                                     vec4_t<T>* nreg2, int* fp1, int* fp2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int s, i, j;
  dedup_decode(rec, e, &s, &i, &j);
  if (isfirst[e]) nreg1[nsu1[e]] = child_region<T>(reg1[su1[s]], pnuv1, i);
  if (isfirst[kDedupMax + e]) nreg2[nsu2[e]] = child_region<T>(reg2[su2[s]], pnuv2, j);
  fp1[(long long)s * cells1 + i] = INT_MAX;
  fp2[(long long)s * cells2 + j] = INT_MAX;
}
template <class T>
__global__ __launch_bounds__(256) void k_dedup_apply(const Cand* rec, int n, int cells1, int cells2, const int* su1,
                                                     const int* su2, const vec4_t<T>* reg1, const vec4_t<T>* reg2,
                                                     int pnuv1, int pnuv2, const unsigned char* isfirst,
                                                     const int* nsu1, const int* nsu2, vec4_t<T>* nreg1,
                                                     vec4_t<T>* nreg2, int* fp1, int* fp2) {
  k_dedup_apply_body<T>(rec, n, cells1, cells2, su1, su2, reg1, reg2, pnuv1, pnuv2, isfirst, nsu1, nsu2, nreg1, nreg2, fp1, fp2);
}

// ---------------------------------------------------------------------------
// Intersection screen (AabbOverlap / intersect_collect_pairs,
// intersection.cu:46-63,259-309).
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_isect_list(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                             const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, int cull,
                             uint64_t* list, unsigned long long* list_count, unsigned long long list_cap) {
  const int tpu1 = tiles_per_unit(w1), tpu2 = tiles_per_unit(w2);
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool keep = false;
  TPIndex x{0, 0, 0};
  if (gid < (long long)nslots * tpu1 * tpu2) {
    x = tp_decode(gid, tpu1, tpu2);
    keep = true;
    if (cull) {
      const int a = su1[x.slot] * tpu1 + x.t1, b = su2[x.slot] * tpu2 + x.t2;
      const vec4_t<T> l1 = lo1[a], h1 = hi1[a], l2 = lo2[b], h2 = hi2[b];
      keep = l1.x <= h2.x && l2.x <= h1.x && l1.y <= h2.y && l2.y <= h1.y && l1.z <= h2.z && l2.z <= h1.z;
    }
  }
  const unsigned long long m = __ballot(keep);
  if (m) {
    unsigned long long base = 0;
    const int leader = __ffsll((long long)m) - 1;
    if (lane_id() == leader) base = atomicAdd(list_count, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    const unsigned long long pos = base + __popcll(m & ((1ull << lane_id()) - 1));
    if (keep && pos < list_cap) list[pos] = tp_encode(x.slot, x.t1, x.t2);
  }
}

// WRITE = false: per-row hit counts (row = slot*cells1 + i).  WRITE = true:
// each hit goes to rowoff[row] + (arrival order within the row); rows are
// sorted by j afterwards, which restores the (slot, i, j) order.
template <class T, bool WRITE>
__global__ __launch_bounds__(256) void k_isect_pass(const uint64_t* list, const unsigned long long* list_count,
                                                    const int* su1, const int* su2, int w1, int w2,
                                                    const vec4_t<T>* C1, const vec4_t<T>* R1,
                                                    const vec4_t<T>* C2, const vec4_t<T>* R2, unsigned* rowcount,
                                                    const unsigned long long* rowoff, unsigned* rowcursor,
                                                    IPair* out, unsigned long long* masks,
                                                    unsigned long long mask_tp) {
  // masks (optional): the count pass keeps each lane's hits of the first
  // mask_tp listed tile pairs (four words: bit jj of word jj / 64), and the
  // write pass reads them instead of repeating the tests
  __shared__ vec4_t<T> sC[kTileCells];
  __shared__ vec4_t<T> sR[kTileCells];
  __shared__ int sJ[kTileCells];
  const unsigned long long n = *list_count;
  const int cells1 = w1 * w1;
  for (unsigned long long it = blockIdx.x; it < n; it += gridDim.x) {
    const uint64_t e = list[it];
    const int slot = (int)(e >> 32), t1 = (int)((e >> 16) & 0xffff), t2 = (int)(e & 0xffff);
    const int u1 = su1[slot], u2 = su2[slot];
    const TileGeo g1 = tile_geo(t1, w1), g2 = tile_geo(t2, w2);
    const int nA = g1.ht * g1.wd, nB = g2.ht * g2.wd;
    __syncthreads();
    if ((int)threadIdx.x < nB) {
      const int j = (g2.row0 + threadIdx.x / g2.wd) * w2 + g2.col0 + threadIdx.x % g2.wd;
      sC[threadIdx.x] = C2[(size_t)u2 * w2 * w2 + j];
      sR[threadIdx.x] = R2[(size_t)u2 * w2 * w2 + j];
      sJ[threadIdx.x] = j;
    }
    const bool va = (int)threadIdx.x < nA;
    int i = 0;
    vec4_t<T> c1, r1;
    c1.x = c1.y = c1.z = c1.w = T(0);
    r1 = c1;
    if (va) {
      i = (g1.row0 + threadIdx.x / g1.wd) * w1 + g1.col0 + threadIdx.x % g1.wd;
      c1 = C1[(size_t)u1 * cells1 + i];
      r1 = R1[(size_t)u1 * cells1 + i];
    }
    __syncthreads();
    if (!va) continue;
    const size_t row = (size_t)slot * cells1 + i;
    unsigned long long* const mw = masks && it < mask_tp ? masks + ((size_t)it * kTileCells + threadIdx.x) * 4 : nullptr;
    if (WRITE && mw) {
      for (int w = 0; w < 4; ++w)
        for (unsigned long long r = mw[w]; r; r &= r - 1) {
          const int jj = 64 * w + __ffsll((long long)r) - 1;
          const unsigned k = atomicAdd(&rowcursor[row], 1u);
          IPair pr;
          pr.slot = slot;
          pr.i = i;
          pr.j = sJ[jj];
          out[rowoff[row] + k] = pr;
        }
      continue;
    }
    unsigned cnt = 0;
    unsigned long long word = 0;
    for (int jj = 0; jj < nB; ++jj) {
      const vec4_t<T> c2 = sC[jj], r2 = sR[jj];
      const bool hit = (fabs(c1.x - c2.x) < (r1.x + r2.x)) & (fabs(c1.y - c2.y) < (r1.y + r2.y)) &
                       (fabs(c1.z - c2.z) < (r1.z + r2.z));
      if (!WRITE) {
        word |= (unsigned long long)hit << (jj & 63);
        if ((jj & 63) == 63 || jj == nB - 1) {
          if (mw) mw[jj >> 6] = word;
          cnt += (unsigned)__popcll(word);
          word = 0;
        }
      } else if (hit) {
        const unsigned k = atomicAdd(&rowcursor[row], 1u);
        IPair pr;
        pr.slot = slot;
        pr.i = i;
        pr.j = sJ[jj];
        out[rowoff[row] + k] = pr;
      }
    }
    if (!WRITE && mw)
      for (int w = (nB + 63) >> 6; w < 4; ++w) mw[w] = 0;  // words past the tile's cells
    if (!WRITE && cnt) atomicAdd(&rowcount[row], cnt);
  }
}

// FP32SCREEN intersection screen.  Float (centre, radius) records, one per
// cell, whose strict float overlap test (k_isect_pass<float>) accepts every
// cell pair that the FP64 test accepts and, given the VERSION5 float values of
// the same grid (round 1), every pair that VERSION5's own float test accepts.
// Per axis: fl64(|c1 - c2|) < fl64(r1 + r2) implies |c1 - c2| < (r1 + r2)(1 +
// 2.0001 u64), so the FP64 test accepts only pairs whose closed intervals
// [c - r (1 + 3 u64), c + r (1 + 3 u64)] meet; the float test likewise with
// [c32 - r32 (1 + 3u), c32 + r32 (1 + 3u)], u = 2^-24.  The record covers the
// hull [lo, hi] of those intervals (widened by 4 u64 (|c| + r) for the FP64
// arithmetic): c = fl32(mid), r = ru32((half + |c - mid|)(1 + 3u) + 4 u64
// (|mid| + half) + 1e-30).  Two records whose hulls meet satisfy |c1 - c2| <=
// r1 + r2 - m with m >= 2.99u (r1 + r2), and then fl32(|c1 - c2|) <=
// |c1 - c2|(1 + u) < (r1 + r2)(1 - u) <= fl32(r1 + r2): the float test accepts.
__global__ void k_isect_boxes32(const double4* C, const double4* R, const float4* Cv, const float4* Rv,
                                unsigned long long n, float4* X, float4* Y) {
  const unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double4 c = C[e], r = R[e];
  float4 cv = make_float4(0.f, 0.f, 0.f, 0.f), rv = cv;
  if (Cv) {
    cv = Cv[e];
    rv = Rv[e];
  }
  const bool has32 = Cv != nullptr;
  float4 x, y;
  record32_axis(c.x, r.x, has32, cv.x, rv.x, &x.x, &y.x);
  record32_axis(c.y, r.y, has32, cv.y, rv.y, &x.y, &y.y);
  record32_axis(c.z, r.z, has32, cv.z, rv.z, &x.z, &y.z);
  x.w = y.w = 0.f;
  X[e] = x;
  Y[e] = y;
}

// FP64 test of the float screen's candidates (the expression of
// k_isect_pass<double>, so the decisions are the FP64 screen's), then the
// order-keeping compaction of the accepted ones.
__global__ void k_isect_retest(const IPair* cand, unsigned long long n, const int* su1, const int* su2, int cells1,
                               int cells2, const double4* C1, const double4* R1, const double4* C2,
                               const double4* R2, unsigned* flag) {
  const unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const IPair p = cand[e];
  const size_t a = (size_t)su1[p.slot] * cells1 + p.i, b = (size_t)su2[p.slot] * cells2 + p.j;
  const double4 c1 = C1[a], r1 = R1[a], c2 = C2[b], r2 = R2[b];
  const bool hit = fabs(c1.x - c2.x) < (r1.x + r2.x) && fabs(c1.y - c2.y) < (r1.y + r2.y) &&
                   fabs(c1.z - c2.z) < (r1.z + r2.z);
  flag[e] = hit ? 1u : 0u;
}
__global__ void k_isect_scatter(const IPair* cand, const unsigned* flag, const unsigned long long* idx,
                                unsigned long long n, IPair* out) {
  const unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n && flag[e]) out[idx[e]] = cand[e];
}

// Sort of each row's pairs by j.  A row's pairs share (slot, i), so only
// their j values move: rows of up to 64 pairs are sorted in registers by a
// bitonic network of 16, 32 or 64 (indices fixed at compile time, padding
// above every j), longer rows by insertion in place.
template <int N>
__device__ __forceinline__ void row_sort_net(IPair* row, int len) {
  int v[N];
#pragma unroll
  for (int x = 0; x < N; ++x) v[x] = x < len ? row[x].j : INT_MAX;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int x = 0; x < N; ++x) {
        const int l = x ^ j;
        if (l > x) {
          const int a = v[x], b = v[l];
          const bool asc = (x & k) == 0;
          v[x] = asc ? min(a, b) : max(a, b);
          v[l] = asc ? max(a, b) : min(a, b);
        }
      }
#pragma unroll
  for (int x = 0; x < N; ++x)
    if (x < len) row[x].j = v[x];
}

__global__ void k_row_sort(const unsigned long long* rowoff, unsigned long long nrows, IPair* out) {
  const unsigned long long r = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const unsigned long long b = rowoff[r], e = rowoff[r + 1];
  const unsigned long long len = e - b;
  if (len <= 1) return;
  if (len <= 16) {
    row_sort_net<16>(out + b, (int)len);
  } else if (len <= 32) {
    row_sort_net<32>(out + b, (int)len);
  } else if (len <= 64) {
    row_sort_net<64>(out + b, (int)len);
  } else {
    for (unsigned long long x = b + 1; x < e; ++x) {
      const IPair v = out[x];
      unsigned long long y = x;
      while (y > b && out[y - 1].j > v.j) {
        out[y] = out[y - 1];
        --y;
      }
      out[y] = v;
    }
  }
}

// Exclusive scan (u32 or u64 input -> u64 output with the total at out[n]);
// blockIdx.y selects one of several arrays of the same length, `*_stride`
// elements apart, scanned in the same launches.
constexpr int kScanBlock = 4096;  // 1024 threads x 4

template <class IN>
__global__ __launch_bounds__(1024) void k_scan_reduce(const IN* in, unsigned long long n,
                                                      unsigned long long* sums, size_t in_stride,
                                                      size_t sums_stride) {
  __shared__ unsigned long long s[1024];
  in += blockIdx.y * in_stride;
  sums += blockIdx.y * sums_stride;
  const unsigned long long base = (unsigned long long)blockIdx.x * kScanBlock;
  unsigned long long v = 0;
  for (int e = 0; e < 4; ++e) {
    const unsigned long long i = base + threadIdx.x * 4 + e;
    if (i < n) v += (unsigned long long)in[i];
  }
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

template <class IN>
// block_off: each block's offset; or, with block_off null and sums given,
// the block adds up sums[0 .. blockIdx.x) itself (at most 1024 blocks), so a
// scan of up to 4M entries takes two launches.  total (optional): total[y]
// receives the array's total as well as out[n].
__global__ __launch_bounds__(1024) void k_scan_apply(const IN* in, unsigned long long n,
                                                     const unsigned long long* block_off, unsigned long long* out,
                                                     size_t in_stride, size_t off_stride, size_t out_stride,
                                                     const unsigned long long* sums, unsigned long long* total) {
  __shared__ unsigned long long s[1024];
  in += blockIdx.y * in_stride;
  out += blockIdx.y * out_stride;
  if (block_off) block_off += blockIdx.y * off_stride;
  unsigned long long boff = 0;
  if (!block_off && sums) {
    sums += blockIdx.y * off_stride;
    s[threadIdx.x] = threadIdx.x < blockIdx.x ? sums[threadIdx.x] : 0ull;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    boff = s[0];
    __syncthreads();
  } else if (block_off) {
    boff = block_off[blockIdx.x];
  }
  const unsigned long long base = (unsigned long long)blockIdx.x * kScanBlock;
  unsigned long long v[4];
  unsigned long long t = 0;
  for (int e = 0; e < 4; ++e) {
    const unsigned long long i = base + threadIdx.x * 4 + e;
    v[e] = i < n ? (unsigned long long)in[i] : 0ull;
    t += v[e];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned long long x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0ull;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  unsigned long long run = boff + s[threadIdx.x] - t;
  for (int e = 0; e < 4; ++e) {
    const unsigned long long i = base + threadIdx.x * 4 + e;
    if (i < n) out[i] = run;
    run += v[e];
    if (i == n - 1) {
      out[n] = run;
      if (total) total[blockIdx.y] = run;
    }
  }
}

__global__ void k_isect_flags1(const IPair* pairs, unsigned long long n, unsigned* flag1) {
  const unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const IPair a = pairs[p];
  bool f = true;
  if (p > 0) {
    const IPair b = pairs[p - 1];
    f = a.slot != b.slot || a.i != b.i;
  }
  flag1[p] = f ? 1u : 0u;
}

__global__ void k_isect_first2(const IPair* pairs, unsigned long long n, int cells2, int* firstpos2) {
  const unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const IPair a = pairs[p];
  atomicMin(&firstpos2[(size_t)a.slot * cells2 + a.j], (int)p);
}

__global__ void k_isect_flags2(const IPair* pairs, unsigned long long n, int cells2, const int* firstpos2,
                               unsigned* flag2) {
  const unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const IPair a = pairs[p];
  flag2[p] = firstpos2[(size_t)a.slot * cells2 + a.j] == (int)p ? 1u : 0u;
}

template <class T>
__global__ void k_isect_units(const IPair* pairs, unsigned long long n, int cells2, const int* su1, const int* su2,
                              const vec4_t<T>* reg1, const vec4_t<T>* reg2, int pnuv1, int pnuv2, const unsigned* flag1,
                              const unsigned* flag2, const unsigned long long* idx1,
                              const unsigned long long* idx2, const int* firstpos2, int* nsu1, int* nsu2,
                              vec4_t<T>* nreg1, vec4_t<T>* nreg2) {
  const unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const IPair a = pairs[p];
  nsu1[p] = (int)(idx1[p] + flag1[p] - 1);
  if (flag1[p]) nreg1[idx1[p]] = child_region<T>(reg1[su1[a.slot]], pnuv1, a.i);
  const int first = firstpos2[(size_t)a.slot * cells2 + a.j];
  nsu2[p] = (int)idx2[first];
  if (flag2[p]) nreg2[idx2[p]] = child_region<T>(reg2[su2[a.slot]], pnuv2, a.j);
}

__global__ void k_isect_reset2(const IPair* pairs, unsigned long long n, int cells2, int* firstpos2) {
  const unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const IPair a = pairs[p];
  firstpos2[(size_t)a.slot * cells2 + a.j] = INT_MAX;
}

template <class T>
__global__ void k_promote(const float* src, T* dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (T)src[i];
}

__global__ void k_fill_i32(int* p, int v, unsigned long long n) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__device__ __forceinline__ void k_fill_u32_body(unsigned* p, unsigned v, unsigned long long n) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(unsigned* p, unsigned v, unsigned long long n) {
  k_fill_u32_body(p, v, n);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <class T>
hipError_t launch_promote(const float* src, T* dst, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_promote<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

// Grid points, one workgroup per 16 x 16-point tile of a unit's nuv x nuv grid
// (EvaluateOneUV, evaluation.cu:259-356; all 18 values, or the positions and
// the compute_K maxima).  The tile's 16 columns and 16 rows get their spans
// and A2.3 basis rows (derivatives up to min(2, degree); knots in LDS when
// KL), the u-contractions R[k][s] = sum_j Nu[k][j] (w) P of each (column,
// control row) are formed once into LDS and shared by the column's points, and
// each point adds its q+1 rows in the reference order for every (k, l), then
// applies the quotient rule.  The operations on every value are those of
// k_points' single-point recipe, so the results are bit-identical; the
// contraction is shared, the control points are read once per column and row
// instead of once per point and derivative order, and no basis tables go
// through HBM.  Control rows are processed in chunks of PtRows.  BICUBIC
// (degree 3 in u and v) uses the fixed-degree basis recurrence, so the
// kernel needs no scratch memory.
template <int MAXD>
struct PtRows {
  static constexpr int value = MAXD <= 3 ? 8 : 4;
};

template <class T, int MAXD, bool RAT, bool KL, int MODE, bool BICUBIC>
__global__ __launch_bounds__(256) void k_points_tile(SurfDev<T> sd, const vec4_t<T>* regions, int nunits, int nuv,
                                                     int steps_given, T* out, unsigned long long* kmax) {
  constexpr int TP = kTile;  // points per tile edge
  constexpr int NB = MAXD + 1;
  constexpr int NC = RAT ? 4 : 3;
  constexpr int NK = MODE == kEvalPoint ? 1 : 3;  // derivative orders kept per direction
  constexpr int kRows = PtRows<MAXD>::value;
  constexpr int kKn = KL ? kEcKnots : 1;
  // one buffer, three lives: knots (phases 0-1), u-contractions (phases 2-3)
  // and, for the 18-value output, the tile's results; barriers separate them
  // u-contractions of a column: kRows x NK entries of NC values, read as
  // 16-byte words; the column stride is an odd number of such words, so the
  // 16 columns that the lanes of a wavefront read fall on distinct banks (an
  // even stride put them on 1-2 bank groups: 8- and 16-way conflicts)
  constexpr int kRw = kRows * NK * NC;
  static_assert(kRw * sizeof(T) % 16 == 0, "column of u-contractions: whole 16-byte words");
  constexpr int kRs = (kRw * (int)sizeof(T) / 16) % 2 == 0 ? kRw + 16 / (int)sizeof(T) : kRw;
  constexpr int kKsz = 2 * kKn, kRsz = TP * kRs, kOsz = MODE == kEvalFull18 ? TP * TP * 18 : 1;
  constexpr int kA = kKsz > kRsz ? (kKsz > kOsz ? kKsz : kOsz) : (kRsz > kOsz ? kRsz : kOsz);
  __shared__ T sA[kA];
  T* const sKu = sA;
  T* const sKv = sA + kKn;
  // entry (column, row, u order) of the u-contractions
  const auto sR = [&](int cx, int rr, int k) -> T* { return sA + cx * kRs + (rr * NK + k) * NC; };
  __shared__ T sNu[TP][NK][NB], sNv[TP][NK][NB];
  __shared__ int sSu[TP], sSv[TP];
  __shared__ unsigned long long sk[4][3];
  const int tid = threadIdx.x;
  const int tw = (nuv + TP - 1) / TP, tpu = tw * tw;
  const int unit = blockIdx.x / tpu, t = blockIdx.x % tpu;
  if (unit >= nunits) return;  // whole workgroup
  const int row0 = (t / tw) * TP, col0 = (t % tw) * TP;
  const int ht = min(TP, nuv - row0), wd = min(TP, nuv - col0);  // points of the tile
  const int p = sd.degu, q = sd.degv;
  const int du = p < 2 ? p : 2, dv = q < 2 ? q : 2;
  const int ku = MODE == kEvalPoint ? 0 : du;  // u orders formed
  const int lv = MODE == kEvalPoint ? 0 : dv;  // v orders formed
  const vec4_t<T> r = regions[unit];
  const int ex = tid % TP, ey = tid / TP;
  const bool active = ex < wd && ey < ht;
  const int npts = nuv * nuv;
  const int pt = (row0 + ey) * nuv + col0 + ex;
  // phase 0: knot vectors in LDS
  if (KL && !sd.zero) {
    for (int x = tid; x < sd.nku; x += 256) sKu[x] = sd.U[x];
    for (int x = tid; x < sd.nkv; x += 256) sKv[x] = sd.V[x];
  }
  __syncthreads();
  // phase 1: spans and basis rows of the tile's columns (u) and rows (v);
  // the parameter follows the grid driver (evaluation.cu:366)
  if (!sd.zero && tid < 2 * TP) {
    const int dir = tid >= TP;
    const int e = dir ? tid - TP : tid;
    if (e < (dir ? ht : wd)) {
      const T start = dir ? r.y : r.x;
      const T step = steps_given ? (dir ? r.w : r.z) : ((dir ? r.w : r.z) - start) / T(nuv);
      const int idx = (dir ? row0 : col0) + e;
      const T tt = step * T(idx) + start;
      const int deg = dir ? q : p, der = dir ? dv : du;
      const T* knots = KL ? (dir ? sKv : sKu) : (dir ? sd.V : sd.U);
      const int span = find_span(dir ? sd.nkv : sd.nku, deg, tt, knots);
      T(*const nb)[NK][NB] = dir ? sNv : sNu;
      if (BICUBIC)  // degree 3 both ways: one recurrence, no degree switch
        basis_row_fixed<T, 3, NK, NB>(span, tt, knots, der, nb[e]);
      else  // fixed-degree recurrences up to degree 7 (no scratch)
        basis_row_any<T, MAXD, NK, NB>(span, deg, tt, knots, der, nb[e]);
      (dir ? sSv : sSu)[e] = span;
    }
  }
  __syncthreads();
  // control rows the tile reads: [vlo - q, vhi] (every thread reads the
  // tile's row spans)
  int vlo = INT_MAX, vhi = INT_MIN;
  if (!sd.zero)
    for (int e = 0; e < ht; ++e) {
      vlo = min(vlo, sSv[e]);
      vhi = max(vhi, sSv[e]);
    }
  const int v0 = vlo - q, nrows = sd.zero ? 0 : vhi - vlo + q + 1;
  T h[6][4];
#pragma unroll
  for (int m = 0; m < 6; ++m)
#pragma unroll
    for (int c = 0; c < 4; ++c) h[m][c] = T(0);
  const int sv = active && !sd.zero ? sSv[ey] : 0;
  for (int r0 = 0; r0 < nrows; r0 += kRows) {
    const int nr = min(kRows, nrows - r0);
    // phase 2: u-contractions of rows [r0, r0 + nr) for every column, all u
    // orders from one read of each control point
    for (int task = tid; task < wd * nr; task += 256) {
      const int rr = task % nr, cx = task / nr;
      const int base = sd.ncpu * (v0 + r0 + rr) + sSu[cx] - p;
      T acc[NK][NC];
#pragma unroll
      for (int k = 0; k < NK; ++k)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[k][c] = T(0);
      constexpr int UR = MAXD <= 7 ? MAXD + 1 : 1;
#pragma unroll UR
      for (int j = 0; j <= MAXD; ++j) {
        if (j > p) break;
        const int id = base + j;
        const T wj = RAT ? sd.W[id] : T(1);
        const T px = sd.CP[3 * id + 0], py = sd.CP[3 * id + 1], pz = sd.CP[3 * id + 2];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          if (k > ku) break;
          const T n = sNu[cx][k][j];
          if (RAT) {
            const T nw = n * wj;
            acc[k][0] += nw * px;
            acc[k][1] += nw * py;
            acc[k][2] += nw * pz;
            acc[k][NC - 1] += nw;
          } else {
            acc[k][0] += n * px;
            acc[k][1] += n * py;
            acc[k][2] += n * pz;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NK; ++k)
        if (k <= ku)
#pragma unroll
          for (int c = 0; c < NC; ++c) sR(cx, rr, k)[c] = acc[k][c];
    }
    __syncthreads();
    // phase 3: each point adds this chunk's rows, in s order, for every (k, l)
    if (active && !sd.zero) {
      if (nrows <= kRows) {  // one chunk (r0 = 0) holds every row the tile
                             // reads: the point's rows are sv - vlo + s2
        const int rb = sv - vlo;
#pragma unroll
        for (int l = 0; l < NK; ++l)
#pragma unroll
          for (int k = 0; k < NK - l; ++k) {
            if (l > lv || k > ku) continue;
            const int m = out_row(k, l);
            constexpr int UR = MAXD <= 7 ? MAXD + 1 : 1;
#pragma unroll UR
            for (int s2 = 0; s2 <= MAXD; ++s2) {
              if (s2 > q) break;
              const T nv = sNv[ey][l][s2];
#pragma unroll
              for (int c = 0; c < NC; ++c) h[m][c] += nv * sR(ex, rb + s2, k)[c];
            }
          }
      } else {
#pragma unroll
        for (int l = 0; l < NK; ++l)
#pragma unroll
          for (int k = 0; k < NK - l; ++k) {  // unrolled: h stays in registers
            if (l > lv || k > ku) continue;
            const int m = out_row(k, l);
            constexpr int UR = MAXD <= 7 ? MAXD + 1 : 1;
#pragma unroll UR
            for (int s2 = 0; s2 <= MAXD; ++s2) {
              if (s2 > q) break;
              const int rr = sv - q + s2 - v0 - r0;
              if (rr < 0 || rr >= nr) continue;
              const T nv = sNv[ey][l][s2];
#pragma unroll
              for (int c = 0; c < NC; ++c) h[m][c] += nv * sR(ex, rr, k)[c];
            }
          }
      }
    }
    __syncthreads();  // sR is rewritten by the next chunk
  }
  if (RAT && active && !sd.zero) {
    // EG_EvaluateQuotientRule2 for dim 3, der 2, stride 4 (evaluation.cu:77-128).
    const T F = T(1) / h[0][3];
    if (MODE == kEvalPoint) {
#pragma unroll
      for (int c = 0; c < 3; ++c) h[0][c] *= F;
    } else {
#pragma unroll
      for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int c = 0; c < 4; ++c) h[m][c] *= F;
      const T ws = -h[1][3], wt = -h[2][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        h[1][c] += ws * h[0][c];
        h[2][c] += wt * h[0][c];
      }
      const T wss = -h[3][3], wst = -h[4][3], wtt = -h[5][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T S = h[0][c], Su = h[1][c], Sv = h[2][c];
        h[3][c] += wss * S + T(2) * ws * Su;
        h[4][c] += wst * S + wt * Su + ws * Sv;
        h[5][c] += wtt * S + T(2) * wt * Sv;
      }
    }
  }
  if (MODE == kEvalFull18) {
    // 18 values per point, point-major: a tile row is one contiguous run of
    // 18 wd values, so the tile goes through LDS (the u-contractions' storage,
    // free after the last barrier of the chunk loop) and is written with
    // consecutive lanes on consecutive addresses; no barrier follows the
    // stores, so the workgroup ends with them in flight
    T(*const sOut)[TP * 18] = reinterpret_cast<T(*)[TP * 18]>(sA);
    if (active)
#pragma unroll
      for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int c = 0; c < 3; ++c) sOut[ey][ex * 18 + 3 * m + c] = h[m][c];
    __syncthreads();
    // pairs of values: every point starts at an even offset (18 values), so
    // the pairs are aligned; a full-width row is TP * 9 pairs, a constant
    using T2 = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;
    T2* const o2 = reinterpret_cast<T2*>(out + ((size_t)unit * npts + (size_t)row0 * nuv + col0) * 18);
    const T2(*const s2)[TP * 9] = reinterpret_cast<const T2(*)[TP * 9]>(sA);
    const int stride = nuv * 9;  // pairs per grid row
    if (wd == TP) {
#pragma unroll
      for (int it = 0; it < (TP * TP * 9 + 255) / 256; ++it) {
        const int e = tid + 256 * it, ry = e / (TP * 9), off = e - ry * (TP * 9);
        if (ry < ht) o2[ry * stride + off] = s2[ry][off];
      }
    } else {
      const int run = wd * 9;
      for (int e = tid; e < ht * run; e += 256) {
        const int ry = e / run, off = e - ry * run;
        o2[ry * stride + off] = s2[ry][off];
      }
    }
    return;
  }
  if (active) {
    T* o = out + (size_t)unit * 3 * npts;
    o[pt] = h[0][0];
    o[npts + pt] = h[0][1];
    o[2 * npts + pt] = h[0][2];
  }
  if (MODE == kEvalPointK) {
    // compute_K maxima (common.cu:155-168); max is exact and order-free.
    unsigned long long mk[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      const T v = cmax(fabs(h[3 + g][0]), cmax(fabs(h[3 + g][1]), fabs(h[3 + g][2])));
      mk[g] = active ? ord_bits(v) : 0ull;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mk[g], off);
        mk[g] = o > mk[g] ? o : mk[g];
      }
    }
    const int wv = tid >> 6;
    if (lane_id() == 0)
      for (int g = 0; g < 3; ++g) sk[wv][g] = mk[g];
    __syncthreads();
    if (tid < 3) {
      unsigned long long m = 0;
      for (int w2 = 0; w2 < 4; ++w2) m = sk[w2][tid] > m ? sk[w2][tid] : m;
      atomicMax(&kmax[unit * 3 + tid], m);
    }
  }
}

template <class T, int MAXD, bool RAT, bool KL, bool BIC>
static void points_tile_mode(const SurfDev<T>& sd, const vec4_t<T>* reg, int nunits, int nuv, int mode,
                             int steps_given, T* out, unsigned long long* kmax, hipStream_t s) {
  const int tw = (nuv + kTile - 1) / kTile;
  const dim3 g((unsigned)((long long)nunits * tw * tw)), b(256);
  if (mode == kEvalFull18)
    hipLaunchKernelGGL((k_points_tile<T, MAXD, RAT, KL, kEvalFull18, BIC>), g, b, 0, s, sd, reg, nunits, nuv,
                       steps_given, out, kmax);
  else if (mode == kEvalPoint)
    hipLaunchKernelGGL((k_points_tile<T, MAXD, RAT, KL, kEvalPoint, BIC>), g, b, 0, s, sd, reg, nunits, nuv,
                       steps_given, out, kmax);
  else
    hipLaunchKernelGGL((k_points_tile<T, MAXD, RAT, KL, kEvalPointK, BIC>), g, b, 0, s, sd, reg, nunits, nuv,
                       steps_given, out, kmax);
}

template <class T, int MAXD, bool RAT, bool KL>
static void points_tile_bic(const SurfDev<T>& sd, const vec4_t<T>* reg, int nunits, int nuv, int mode,
                            int steps_given, T* out, unsigned long long* kmax, hipStream_t s) {
  if (MAXD == 3 && sd.degu == 3 && sd.degv == 3 && !sd.zero)
    points_tile_mode<T, MAXD, RAT, KL, true>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
  else
    points_tile_mode<T, MAXD, RAT, KL, false>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
}

template <class T, int MAXD>
static void points_tile_rat(const SurfDev<T>& sd, const vec4_t<T>* reg, int nunits, int nuv, int mode,
                            int steps_given, T* out, unsigned long long* kmax, hipStream_t s) {
  const bool kl = sd.nku <= kEcKnots && sd.nkv <= kEcKnots;
  if (sd.rational) {
    if (kl)
      points_tile_bic<T, MAXD, true, true>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
    else
      points_tile_bic<T, MAXD, true, false>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
  } else {
    if (kl)
      points_tile_bic<T, MAXD, false, true>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
    else
      points_tile_bic<T, MAXD, false, false>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
  }
}

template <class T>
hipError_t launch_points_tile(const SurfDev<T>& sd, const vec4_t<T>* reg, int nunits, int nuv, int mode,
                              int steps_given, T* out, unsigned long long* kmax, hipStream_t s) {
  if (nunits <= 0 || nuv <= 0) return hipSuccess;
  const long long tw = (nuv + kTile - 1) / kTile;
  if ((long long)nunits * tw * tw > 0x7fffffffll) return hipErrorInvalidValue;
  const int d = std::max(sd.degu, sd.degv);
  if (d <= 3)
    points_tile_rat<T, 3>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
  else if (d <= 7)
    points_tile_rat<T, 7>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
  else
    points_tile_rat<T, kMaxDeg - 1>(sd, reg, nunits, nuv, mode, steps_given, out, kmax, s);
  return hipGetLastError();
}

template <class T>
static int ec_blocks(const EcJob<T>& j) {
  if (j.nunits <= 0 || j.nuv < 2) return 0;
  const int tw = (j.nuv - 1 + kTile - 1) / kTile;
  return j.nunits * tw * tw;
}

// variant of a job: 0-2 = degree class (3, 7, kMaxDeg - 1), +3 rational, +6 knots in LDS
template <class T>
static int ec_variant(const EcJob<T>& j) {
  const int d = std::max(j.sd.degu, j.sd.degv);
  const int dc = d <= 3 ? 0 : d <= 7 ? 1 : 2;
  const bool kl = j.sd.nku <= kEcKnots && j.sd.nkv <= kEcKnots;
  return dc + (j.sd.rational ? 3 : 0) + (kl ? 6 : 0);
}

template <class T, int MAXD>
static void ec_launch_rat(int var, const EcJob<T>& j0, const EcJob<T>& j1, int nb0, int nb, hipStream_t s) {
  const dim3 g((unsigned)nb), b(64);
  const bool rat = (var / 3) % 2 == 1, kl = var >= 6;
  if (rat) {
    if (kl)
      hipLaunchKernelGGL((k_eval_centres<T, MAXD, true, true>), g, b, 0, s, j0, j1, nb0);
    else
      hipLaunchKernelGGL((k_eval_centres<T, MAXD, true, false>), g, b, 0, s, j0, j1, nb0);
  } else {
    if (kl)
      hipLaunchKernelGGL((k_eval_centres<T, MAXD, false, true>), g, b, 0, s, j0, j1, nb0);
    else
      hipLaunchKernelGGL((k_eval_centres<T, MAXD, false, false>), g, b, 0, s, j0, j1, nb0);
  }
}

template <class T>
static void ec_launch(int var, const EcJob<T>& j0, const EcJob<T>& j1, int nb0, int nb, hipStream_t s) {
  if (nb <= 0) return;
  switch (var % 3) {
    case 0: ec_launch_rat<T, 3>(var, j0, j1, nb0, nb, s); break;
    case 1: ec_launch_rat<T, 7>(var, j0, j1, nb0, nb, s); break;
    default: ec_launch_rat<T, kMaxDeg - 1>(var, j0, j1, nb0, nb, s); break;
  }
}

template <class T>
static bool ec_valid(const EcOut<T>& o) {
  if ((o.C32 || o.absmax || o.lo32) && !std::is_same<T, double>::value) return false;
  return !((o.C32 && !o.absmax) || (o.lo32 && !o.C32) || (!o.lo != !o.hi) || (!o.lo32 != !o.hi32));
}

template <class T>
hipError_t launch_eval_centres(const EcJob<T>& a, const EcJob<T>* b, hipStream_t s) {
  if (!ec_valid(a.o) || (b && !ec_valid(b->o))) return hipErrorInvalidValue;
  const int na = ec_blocks(a), nb = b ? ec_blocks(*b) : 0;
  const int va = ec_variant(a);
  if (b && nb > 0 && ec_variant(*b) == va) {
    if ((long long)na + nb > INT_MAX) return hipErrorInvalidValue;
    ec_launch(va, a, *b, na, na + nb, s);
  } else {
    ec_launch(va, a, a, na, na, s);
    if (b) ec_launch(ec_variant(*b), *b, *b, nb, nb, s);
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_aabb(const T* pts, int nunits, int nuv, int isect, int use_k, const unsigned long long* kmax,
                       const vec4_t<T>* regions, T init_u, T init_v, vec4_t<T>* C, vec4_t<T>* R, hipStream_t s) {
  const long long n = (long long)nunits * (nuv - 1) * (nuv - 1);
  if (n <= 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  if (isect)
    hipLaunchKernelGGL((k_aabb<T, true>), g, b, 0, s, pts, nunits, nuv, use_k, kmax, regions, init_u, init_v, C, R);
  else
    hipLaunchKernelGGL((k_aabb<T, false>), g, b, 0, s, pts, nunits, nuv, use_k, kmax, regions, init_u, init_v, C,
                       R);
  return hipGetLastError();
}

template <class T>
hipError_t launch_tilebox(const vec4_t<T>* C, const vec4_t<T>* R, int nunits, int w, int isect, vec4_t<T>* lo,
                          vec4_t<T>* hi, hipStream_t s) {
  const int tw = (w + kTile - 1) / kTile;
  const long long waves = (long long)nunits * tw * tw;
  if (waves <= 0) return hipSuccess;
  const dim3 g(blocks_for(waves * 64, 256)), b(256);
  if (isect)
    hipLaunchKernelGGL((k_tilebox<T, true>), g, b, 0, s, C, R, nunits, w, lo, hi);
  else
    hipLaunchKernelGGL((k_tilebox<T, false>), g, b, 0, s, C, R, nunits, w, lo, hi);
  return hipGetLastError();
}

static int tiles_per_unit_h(int w) {
  const int tw = (w + kTile - 1) / kTile;
  return tw * tw;
}

static long long n_tile_pairs(int nslots, int w1, int w2) {
  const long long t1 = (w1 + kTile - 1) / kTile, t2 = (w2 + kTile - 1) / kTile;
  return (long long)nslots * t1 * t1 * t2 * t2;
}

template <class T>
hipError_t launch_dist_seed(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                            const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, int k,
                            unsigned long long* best, hipStream_t s) {
  const long long npos = (long long)nslots * tiles_per_unit_h(w1);
  if (npos <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dist_seed<T>, dim3((unsigned)npos), dim3(256), 0, s, nslots, su1, su2, w1, w2, lo1, hi1, lo2,
                     hi2, k, best);
  return hipGetLastError();
}

template <class T>
hipError_t launch_seed_thresh(const unsigned long long* best, int npos, int tpu1, int nseed, int ordered,
                              uint64_t* list, unsigned* list_count, unsigned long long* range,
                              unsigned long long* work, const int* su1, const int* su2, int w1, int w2,
                              const vec4_t<T>* C1, const vec4_t<T>* C2, int k, int fma, T* thresh, hipStream_t s,
                              unsigned* clear, int nclear) {
  if (npos > kSeedSortMax || nseed > kSeedMaxTiles || nclear > kSeedHistBins) return hipErrorInvalidValue;
  if (ordered && tpu1 != 1) return hipErrorInvalidValue;
  // ordered: seed tile pair i is slot i, nothing to pick
  if (!ordered)
    hipLaunchKernelGGL(k_seed_list<T>, dim3(1), dim3(1024), 0, s, best, npos, tpu1, nseed, ordered, list,
                       list_count);
  const int m = nseed < npos ? nseed : npos;
  // about 1024 workgroups: fill the chip without splitting tiles finely
  int split = m > 0 ? 1024 / m : 1;
  split = split < 1 ? 1 : (split > kSeedSplitMax ? kSeedSplitMax : split);
  const int nb = m > 0 ? m * split : 1;
  unsigned* gh = reinterpret_cast<unsigned*>(work + 2 * (size_t)kSeedMaxTiles * kSeedSplitMax);
  if (fma) {
    hipLaunchKernelGGL((k_seed_hist<T, true, true>), dim3(nb), dim3(256), 0, s, list, m, ordered, su1, su2, w1, w2,
                       C1, C2, split, range, work, gh);
    hipLaunchKernelGGL(k_seed_range, dim3(1), dim3(1024), 0, s, work, nb, range, gh);
    hipLaunchKernelGGL((k_seed_hist<T, true, false>), dim3(nb), dim3(256), 0, s, list, m, ordered, su1, su2, w1, w2,
                       C1, C2, split, range, work, gh);
  } else {
    hipLaunchKernelGGL((k_seed_hist<T, false, true>), dim3(nb), dim3(256), 0, s, list, m, ordered, su1, su2, w1,
                       w2, C1, C2, split, range, work, gh);
    hipLaunchKernelGGL(k_seed_range, dim3(1), dim3(1024), 0, s, work, nb, range, gh);
    hipLaunchKernelGGL((k_seed_hist<T, false, false>), dim3(nb), dim3(256), 0, s, list, m, ordered, su1, su2, w1,
                       w2, C1, C2, split, range, work, gh);
  }
  hipLaunchKernelGGL(k_seed_thresh<T>, dim3(1), dim3(kSeedHistBins), 0, s, gh, range, k, thresh, clear, nclear);
  return hipGetLastError();
}

size_t seed_work_bytes() {
  return 2 * (size_t)kSeedMaxTiles * kSeedSplitMax * sizeof(unsigned long long) + kSeedHistBins * sizeof(unsigned);
}

template <class T>
hipError_t launch_dist_list(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                            const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, const T* thresh,
                            int cull, uint64_t* list, unsigned* list_count, hipStream_t s, int fma) {
  const long long n = n_tile_pairs(nslots, w1, w2);
  if (n <= 0) return hipSuccess;
  const long long rows = (long long)nslots * tiles_per_unit_h(w1);
  if (tiles_per_unit_h(w2) >= 64 && rows <= 0x7fffffffll) {  // many side-2 tiles: a workgroup per side-1 tile
    if (fma)
      hipLaunchKernelGGL((k_dist_list_rows<T, true>), dim3((unsigned)rows), dim3(256), 0, s, nslots, su1, su2, w1, w2,
                         lo1, hi1, lo2, hi2, thresh, cull, list, list_count);
    else
      hipLaunchKernelGGL((k_dist_list_rows<T, false>), dim3((unsigned)rows), dim3(256), 0, s, nslots, su1, su2, w1,
                         w2, lo1, hi1, lo2, hi2, thresh, cull, list, list_count);
    return hipGetLastError();
  }
  if (fma)
    hipLaunchKernelGGL((k_dist_list<T, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s, nslots, su1, su2, w1, w2,
                       lo1, hi1, lo2, hi2, thresh, cull, list, list_count);
  else
    hipLaunchKernelGGL((k_dist_list<T, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s, nslots, su1, su2, w1, w2,
                       lo1, hi1, lo2, hi2, thresh, cull, list, list_count);
  return hipGetLastError();
}

template <class T>
hipError_t launch_dist_filter(const uint64_t* list, const unsigned* list_count, unsigned list_cap, const int* su1,
                              const int* su2, int w1, int w2, const vec4_t<T>* C1, const vec4_t<T>* C2,
                              const T* thresh, Cand* out, unsigned long long* out_count, unsigned long long cap,
                              int grid, hipStream_t s, unsigned long long* pairs_done, int fma, int cull,
                              const unsigned long long* absmax, int screen32, const vec4_t<T>* blo1,
                              const vec4_t<T>* bhi1, const vec4_t<T>* blo2, const vec4_t<T>* bhi2, int walk) {
  if (list_cap == 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<long long>(grid, (long long)list_cap);
  if (w1 <= kTile && w2 <= kTile) {  // every unit is one tile: one slot per list entry
    const unsigned gu = list_cap;     // one workgroup per entry: setup latency overlaps across slots
    if (absmax) {
      if constexpr (std::is_same<T, double>::value) {
        // one resident round of workgroups (4 per CU: the kernel is held to 128 registers for four
        // wavefronts per SIMD) walking the slots: a workgroup lives a few microseconds per slot, too
        // short to launch one per slot (at most 8 slots per workgroup: the grid grows for more slots)
        // (at least one workgroup per `walk` slots, walk <= 16 = the kernels' kMaxWalk)
        const unsigned wk = (unsigned)std::min(16, std::max(1, walk));
        const unsigned gp = std::max(std::max(1u, (unsigned)grid / 2u), (gu + wk - 1u) / wk);
        const int nq = (w2 * w2 / 16) / 4;
#define JSI_MFMA_FILTER(NQ)                                                                                         \
  if (screen32)                                                                                                     \
    hipLaunchKernelGGL(k_dist_filter_mfma32<NQ>, dim3(std::min(gu, gp)), dim3(256), 0, s, list, list_count, (int)gu, \
                       su1, su2, w1, w2, C1, C2, blo1, bhi1, blo2, bhi2, thresh, out, out_count, cap, pairs_done);  \
  else                                                                                                              \
    hipLaunchKernelGGL(k_dist_filter_mfma<NQ>, dim3(std::min(gu, gp)), dim3(256), 0, s, list, list_count, (int)gu,  \
                       su1, su2, w1, w2, C1, C2, thresh, out, out_count, cap, pairs_done, absmax)
        switch (nq) {
          case 0: JSI_MFMA_FILTER(0); break;
          case 1: JSI_MFMA_FILTER(1); break;
          case 2: JSI_MFMA_FILTER(2); break;
          case 3: JSI_MFMA_FILTER(3); break;
          default: JSI_MFMA_FILTER(4); break;
        }
#undef JSI_MFMA_FILTER
      } else {
        return hipErrorInvalidValue;
      }
    } else if (fma) {
      hipLaunchKernelGGL((k_dist_filter_units<T, true, false>), dim3(gu), dim3(256), 0, s, list, list_count, su1, su2,
                         w1, w2, C1, C2, thresh, out, out_count, cap, pairs_done, absmax);
    } else {
      hipLaunchKernelGGL((k_dist_filter_units<T, false, false>), dim3(gu), dim3(256), 0, s, list, list_count, su1, su2,
                         w1, w2, C1, C2, thresh, out, out_count, cap, pairs_done, absmax);
    }
    return hipGetLastError();
  }
  if (absmax) {
    if constexpr (std::is_same<T, double>::value)
      hipLaunchKernelGGL((k_dist_filter<T, false, true>), dim3(g), dim3(256), 0, s, list, list_count, su1, su2, w1,
                         w2, C1, C2, thresh, out, out_count, cap, pairs_done, cull, absmax);
    else
      return hipErrorInvalidValue;
  } else if (fma) {
    hipLaunchKernelGGL((k_dist_filter<T, true, false>), dim3(g), dim3(256), 0, s, list, list_count, su1, su2, w1, w2,
                       C1, C2, thresh, out, out_count, cap, pairs_done, cull, absmax);
  } else {
    hipLaunchKernelGGL((k_dist_filter<T, false, false>), dim3(g), dim3(256), 0, s, list, list_count, su1, su2, w1,
                       w2, C1, C2, thresh, out, out_count, cap, pairs_done, cull, absmax);
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_set_thresh(T* thresh, const Cand* rec, int idx, int use_inf, hipStream_t s) {
  hipLaunchKernelGGL(k_set_thresh<T>, dim3(1), dim3(64), 0, s, thresh, rec, idx, use_inf);
  return hipGetLastError();
}

hipError_t launch_inflate(float* thresh, const unsigned long long* absmax_bits, hipStream_t s) {
  hipLaunchKernelGGL(k_inflate, dim3(1), dim3(64), 0, s, thresh, absmax_bits);
  return hipGetLastError();
}

hipError_t launch_rerank(Cand* cand, const unsigned long long* n_dev, unsigned long long cap, const int* su1,
                         const int* su2, int cells1, int cells2, const double4* C1, const double4* C2, hipStream_t s) {
  if (cap == 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<unsigned long long>(1024, (cap + 255) / 256);
  hipLaunchKernelGGL(k_rerank, dim3(g), dim3(256), 0, s, cand, n_dev, cap, su1, su2, cells1, cells2, C1, C2);
  return hipGetLastError();
}

hipError_t launch_pack_cells(const float* c, const float* r, long long n, int w, float4* C, float4* R,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_pack_cells, dim3(blocks_for((long long)w * w, 256)), dim3(256), 0, s, c, r, n, w, C, R);
  return hipGetLastError();
}

hipError_t launch_pair_mask(int mode, const uint64_t* list, const unsigned long long* list_count,
                            const unsigned* list_count32, int w1, int w2, const float4* C1, const float4* R1,
                            const float4* C2, const float4* R2, const float* thresh, long long n1, long long n2,
                            int* m1, int* m2, int grid, hipStream_t s) {
  if (mode == 0)
    hipLaunchKernelGGL(k_pair_mask<0>, dim3(grid), dim3(256), 0, s, list, list_count, list_count32, w1, w2, C1, R1,
                       C2, R2, thresh, n1, n2, m1, m2);
  else
    hipLaunchKernelGGL(k_pair_mask<1>, dim3(grid), dim3(256), 0, s, list, list_count, list_count32, w1, w2, C1, R1,
                       C2, R2, thresh, n1, n2, m1, m2);
  return hipGetLastError();
}

hipError_t select_topk(const Cand* cand, unsigned long long n, unsigned long long k, unsigned long long maxkey,
                       SelState* st, unsigned* hist, Cand* out, unsigned long long* out_count, unsigned small_cap,
                       hipStream_t s, int iters, bool resume, const unsigned long long* n_dev) {
  const unsigned g = (unsigned)std::max(1ull, std::min<unsigned long long>(1024, (n + 255) / 256));
  if (!resume) {
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, s, st, n, n_dev, k, maxkey, small_cap, hist,
                       out ? out_count : nullptr);
    hipLaunchKernelGGL(k_sel_minmax, dim3(g), dim3(256), 0, s, cand, n, n_dev, st);
  } else if (out) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(out_count), 0u, 2ull);
  }
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_sel_hist, dim3(g), dim3(256), 0, s, cand, n, n_dev, st, hist);
    hipLaunchKernelGGL(k_sel_step, dim3(1), dim3(1024), 0, s, st, hist, small_cap);
  }
  if (out) {
    if (n > 0)
      hipLaunchKernelGGL(k_sel_compact, dim3(g), dim3(256), 0, s, cand, n, n_dev, st, out, out_count,
                         (unsigned long long)small_cap);
  }
  return hipGetLastError();
}

hipError_t launch_thresh_from_state_ru(float* thresh, const SelState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_thresh_from_state_ru, dim3(1), dim3(64), 0, s, thresh, st);
  return hipGetLastError();
}

hipError_t launch_set_thresh_ru(float* thresh, const Cand* rec, int idx, hipStream_t s) {
  hipLaunchKernelGGL(k_set_thresh_ru, dim3(1), dim3(64), 0, s, thresh, rec, idx);
  return hipGetLastError();
}

hipError_t launch_copy_cand(Cand* dst, const Cand* src, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_cand, dim3(1), dim3(64), 0, s, dst, src);
  return hipGetLastError();
}


template <class T>
hipError_t launch_query_init(vec4_t<T>* reg1, vec4_t<T> g1, vec4_t<T>* reg2, vec4_t<T> g2, int* su1, int* su2,
                             unsigned long long* amax, hipStream_t s) {
  hipLaunchKernelGGL(k_query_init<T>, dim3(1), dim3(64), 0, s, reg1, g1, reg2, g2, su1, su2, amax);
  return hipGetLastError();
}
template hipError_t launch_query_init<float>(float4*, float4, float4*, float4, int*, int*, unsigned long long*,
                                             hipStream_t);
template hipError_t launch_query_init<double>(double4*, double4, double4*, double4, int*, int*, unsigned long long*,
                                              hipStream_t);

template <class T>
hipError_t launch_thresh_from_state(T* thresh, const SelState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_thresh_from_state<T>, dim3(1), dim3(64), 0, s, thresh, st);
  return hipGetLastError();
}
template hipError_t launch_thresh_from_state<float>(float*, const SelState*, hipStream_t);
template hipError_t launch_thresh_from_state<double>(double*, const SelState*, hipStream_t);

hipError_t launch_fill_sentinel(Cand* a, unsigned long long from, unsigned long long n2, hipStream_t s) {
  if (n2 <= from) return hipSuccess;
  hipLaunchKernelGGL(k_fill_sentinel, dim3(blocks_for(n2 - from, 256)), dim3(256), 0, s, a, from, n2);
  return hipGetLastError();
}

hipError_t rank_sort(const Cand* in, const unsigned long long* m_dev, unsigned m_cap, unsigned keep, Cand* out,
                     hipStream_t s, Cand* last_out) {
  if (m_cap == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rank_sort, dim3(blocks_for(m_cap, 64)), dim3(64 * kRankWaves), 0, s, in, m_dev, m_cap, keep, out,
                     last_out);
  return hipGetLastError();
}

hipError_t sort_cands(Cand* a, unsigned long long n2, hipStream_t s) {
  if ((n2 & (n2 - 1)) != 0) return hipErrorInvalidValue;
  static const bool big_lds =
      hipFuncSetAttribute((const void*)k_bsort_single, hipFuncAttributeMaxDynamicSharedMemorySize,
                          kBlockSortMax * (int)sizeof(Cand)) == hipSuccess;
  if (n2 <= (unsigned long long)(big_lds ? kBlockSortMax : 4096)) {
    hipLaunchKernelGGL(k_bsort_single, dim3(1), dim3(1024), n2 * sizeof(Cand), s, a, (int)n2);
    return hipGetLastError();
  }
  if (n2 < (unsigned long long)kSortBlock) return hipErrorInvalidValue;
  const unsigned nb = (unsigned)(n2 / kSortBlock);
  hipLaunchKernelGGL(k_bsort_local, dim3(nb), dim3(1024), 0, s, a, 0ull);
  for (unsigned long long k = 2ull * kSortBlock; k <= n2; k <<= 1) {
    for (unsigned long long j = k >> 1; j >= (unsigned long long)kSortBlock; j >>= 1)
      hipLaunchKernelGGL(k_bsort_global, dim3(blocks_for(n2 / 2, 256)), dim3(256), 0, s, a, n2, k, j);
    hipLaunchKernelGGL(k_bsort_local, dim3(nb), dim3(1024), 0, s, a, k);
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_dist_dedup(const Cand* rec, int n, int cells1, int cells2, const int* su1, const int* su2,
                             const vec4_t<T>* reg1, const vec4_t<T>* reg2, int parent_nuv1, int parent_nuv2,
                             int* firstpos1, int* firstpos2, int* nsu1, int* nsu2, vec4_t<T>* nreg1,
                             vec4_t<T>* nreg2, int* nunits, unsigned char* work, hipStream_t s,
                             unsigned long long* zero, int nzero) {
  if (n <= 0 || n > kDedupMax || nzero > 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_dedup_mark, dim3(blocks_for(n, 256)), dim3(256), 0, s, rec, n, cells1, cells2, firstpos1,
                     firstpos2);
  hipLaunchKernelGGL(k_dedup_units, dim3(1), dim3(1024), 0, s, rec, n, cells1, cells2, firstpos1, firstpos2, nsu1,
                     nsu2, work, nunits, zero, nzero);
  hipLaunchKernelGGL(k_dedup_apply<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, rec, n, cells1, cells2, su1, su2,
                     reg1, reg2, parent_nuv1, parent_nuv2, work, nsu1, nsu2, nreg1, nreg2, firstpos1, firstpos2);
  return hipGetLastError();
}

size_t dedup_work_bytes() { return 2 * (size_t)kDedupMax; }

template <class T>
hipError_t launch_isect_list(int nslots, const int* su1, const int* su2, int w1, int w2, const vec4_t<T>* lo1,
                             const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, int cull,
                             uint64_t* list, unsigned long long* list_count, unsigned long long list_cap,
                             hipStream_t s) {
  const long long n = n_tile_pairs(nslots, w1, w2);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_isect_list<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, nslots, su1, su2, w1, w2, lo1,
                     hi1, lo2, hi2, cull, list, list_count, list_cap);
  return hipGetLastError();
}

template <class T>
hipError_t launch_isect_pass(int write, const uint64_t* list, const unsigned long long* list_count, const int* su1,
                             const int* su2, int w1, int w2, const vec4_t<T>* C1, const vec4_t<T>* R1,
                             const vec4_t<T>* C2, const vec4_t<T>* R2, unsigned* rowcount,
                             const unsigned long long* rowoff, unsigned* rowcursor, IPair* out, int grid,
                             hipStream_t s, unsigned long long* masks, unsigned long long mask_tp) {
  if (!masks) mask_tp = 0;
  if (write)
    hipLaunchKernelGGL((k_isect_pass<T, true>), dim3(grid), dim3(256), 0, s, list, list_count, su1, su2, w1, w2, C1,
                       R1, C2, R2, rowcount, rowoff, rowcursor, out, masks, mask_tp);
  else
    hipLaunchKernelGGL((k_isect_pass<T, false>), dim3(grid), dim3(256), 0, s, list, list_count, su1, su2, w1, w2,
                       C1, R1, C2, R2, rowcount, rowoff, rowcursor, out, masks, mask_tp);
  return hipGetLastError();
}

hipError_t launch_row_sort(const unsigned long long* rowoff, unsigned long long nrows, IPair* out, hipStream_t s) {
  if (nrows == 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_sort, dim3(blocks_for(nrows, 256)), dim3(256), 0, s, rowoff, nrows, out);
  return hipGetLastError();
}

hipError_t launch_isect_boxes32(const double4* C, const double4* R, const float4* Cv, const float4* Rv,
                                unsigned long long n, float4* X, float4* Y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_isect_boxes32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, C, R, Cv, Rv, n, X, Y);
  return hipGetLastError();
}
hipError_t launch_isect_retest(const IPair* cand, unsigned long long n, const int* su1, const int* su2, int cells1,
                               int cells2, const double4* C1, const double4* R1, const double4* C2, const double4* R2,
                               unsigned* flag, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_isect_retest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cand, n, su1, su2, cells1,
                     cells2, C1, R1, C2, R2, flag);
  return hipGetLastError();
}
hipError_t launch_isect_scatter(const IPair* cand, const unsigned* flag, const unsigned long long* idx,
                                unsigned long long n, IPair* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_isect_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cand, flag, idx, n, out);
  return hipGetLastError();
}

size_t scan_tmp_bytes(unsigned long long n, int arrays) {
  size_t total = 0;
  unsigned long long m = n;
  while (true) {
    const unsigned long long nb = (m + kScanBlock - 1) / kScanBlock;
    total += (nb + 1) * 2 * sizeof(unsigned long long) * (size_t)arrays;
    if (nb <= 1) break;
    m = nb;
  }
  return total + 256;
}

// K arrays: in + y in_stride -> out + y out_stride, y < K (total: see
// k_scan_apply)
template <class IN>
static hipError_t scan_rec(const IN* in, size_t in_stride, unsigned long long* out, size_t out_stride,
                           unsigned long long n, int K, unsigned long long* tmp, hipStream_t s,
                           unsigned long long* total = nullptr) {
  const unsigned long long nb = (n + kScanBlock - 1) / kScanBlock;
  const unsigned long long* none = nullptr;
  if (nb == 1) {  // one block per array: the scan and the total in one launch
    hipLaunchKernelGGL(k_scan_apply<IN>, dim3(1, (unsigned)K), dim3(1024), 0, s, in, n, none, out, in_stride,
                       (size_t)0, out_stride, none, total);
    return hipGetLastError();
  }
  const size_t st = nb + 1;  // per-array stride of the block sums and offsets
  unsigned long long* sums = tmp;
  unsigned long long* offs = tmp + (size_t)K * st;
  hipLaunchKernelGGL(k_scan_reduce<IN>, dim3((unsigned)nb, (unsigned)K), dim3(1024), 0, s, in, n, sums, in_stride,
                     st);
  if (nb <= 1024) {  // each block adds up the sums before it
    hipLaunchKernelGGL(k_scan_apply<IN>, dim3((unsigned)nb, (unsigned)K), dim3(1024), 0, s, in, n, none, out,
                       in_stride, st, out_stride, (const unsigned long long*)sums, total);
    return hipGetLastError();
  }
  hipError_t e = scan_rec<unsigned long long>(sums, st, offs, st, nb, K, tmp + 2 * (size_t)K * st, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_apply<IN>, dim3((unsigned)nb, (unsigned)K), dim3(1024), 0, s, in, n,
                     (const unsigned long long*)offs, out, in_stride, st, out_stride, none, total);
  return hipGetLastError();
}

hipError_t scan_u32_to_u64(const unsigned* in, unsigned long long* out, unsigned long long n, void* tmp,
                           size_t tmp_bytes, hipStream_t s) {
  if (n == 0) return launch_fill_u32(reinterpret_cast<unsigned*>(out), 0u, 2, s);
  if (tmp_bytes < scan_tmp_bytes(n)) return hipErrorInvalidValue;
  return scan_rec<unsigned>(in, 0, out, 0, n, 1, static_cast<unsigned long long*>(tmp), s);
}

hipError_t scan3_u32_to_u64(const unsigned* in, size_t in_stride, unsigned long long* out, size_t out_stride,
                            unsigned long long n, void* tmp, size_t tmp_bytes, hipStream_t s,
                            unsigned long long* totals) {
  if (n == 0) {
    for (int y = 0; y < 3; ++y) {
      const hipError_t e = launch_fill_u32(reinterpret_cast<unsigned*>(out + y * out_stride), 0u, 2, s);
      if (e != hipSuccess) return e;
    }
    return totals ? launch_fill_u32(reinterpret_cast<unsigned*>(totals), 0u, 6, s) : hipSuccess;
  }
  if (tmp_bytes < scan_tmp_bytes(n, 3) || in_stride < n || out_stride < n + 1) return hipErrorInvalidValue;
  return scan_rec<unsigned>(in, in_stride, out, out_stride, n, 3, static_cast<unsigned long long*>(tmp), s, totals);
}

hipError_t launch_isect_dedup_flags(const IPair* pairs, unsigned long long n, int cells2, int* firstpos2,
                                    unsigned* flag1, unsigned* flag2, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  hipLaunchKernelGGL(k_isect_flags1, g, b, 0, s, pairs, n, flag1);
  hipLaunchKernelGGL(k_isect_first2, g, b, 0, s, pairs, n, cells2, firstpos2);
  hipLaunchKernelGGL(k_isect_flags2, g, b, 0, s, pairs, n, cells2, firstpos2, flag2);
  return hipGetLastError();
}

template <class T>
hipError_t launch_isect_units(const IPair* pairs, unsigned long long n, int cells2, const int* su1, const int* su2,
                              const vec4_t<T>* reg1, const vec4_t<T>* reg2, int parent_nuv1, int parent_nuv2,
                              const unsigned* flag1, const unsigned* flag2, const unsigned long long* idx1,
                              const unsigned long long* idx2, int* firstpos2, int* nsu1, int* nsu2,
                              vec4_t<T>* nreg1, vec4_t<T>* nreg2, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  hipLaunchKernelGGL(k_isect_units<T>, g, b, 0, s, pairs, n, cells2, su1, su2, reg1, reg2, parent_nuv1, parent_nuv2,
                     flag1, flag2, idx1, idx2, firstpos2, nsu1, nsu2, nreg1, nreg2);
  hipLaunchKernelGGL(k_isect_reset2, g, b, 0, s, pairs, n, cells2, firstpos2);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, int value, unsigned long long n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, value, n);
  return hipGetLastError();
}
hipError_t launch_fill_u32(unsigned* p, unsigned value, unsigned long long n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, value, n);
  return hipGetLastError();
}

// ---- explicit instantiations ----------------------------------------------
#define JSI_INST(T)                                                                                                 \
  template hipError_t launch_promote<T>(const float*, T*, int, hipStream_t);                                       \
  template hipError_t launch_aabb<T>(const T*, int, int, int, int, const unsigned long long*, const vec4_t<T>*, T, \
                                     T, vec4_t<T>*, vec4_t<T>*, hipStream_t);                                      \
  template hipError_t launch_tilebox<T>(const vec4_t<T>*, const vec4_t<T>*, int, int, int, vec4_t<T>*,             \
                                        vec4_t<T>*, hipStream_t);                                                  \
  template hipError_t launch_dist_seed<T>(int, const int*, const int*, int, int, const vec4_t<T>*,                 \
                                          const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, int,               \
                                          unsigned long long*, hipStream_t);                                       \
  template hipError_t launch_dist_list<T>(int, const int*, const int*, int, int, const vec4_t<T>*,                 \
                                          const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, const T*, int,     \
                                          uint64_t*, unsigned*, hipStream_t, int);                                      \
  template hipError_t launch_dist_filter<T>(const uint64_t*, const unsigned*, unsigned, const int*, const int*,    \
                                            int, int, const vec4_t<T>*, const vec4_t<T>*, const T*, Cand*,         \
                                            unsigned long long*, unsigned long long, int, hipStream_t,             \
                                            unsigned long long*, int, int, const unsigned long long*, int,         \
                                            const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, \
                                            int);                                                                  \
  template hipError_t launch_set_thresh<T>(T*, const Cand*, int, int, hipStream_t);                                \
  template hipError_t launch_eval_centres<T>(const EcJob<T>&, const EcJob<T>*, hipStream_t);                      \
  template hipError_t launch_points_tile<T>(const SurfDev<T>&, const vec4_t<T>*, int, int, int, int, T*,           \
                                            unsigned long long*, hipStream_t);                                     \
  template hipError_t launch_seed_thresh<T>(const unsigned long long*, int, int, int, int, uint64_t*, unsigned*,   \
                                            unsigned long long*, unsigned long long*, const int*, const int*, int, \
                                            int, const vec4_t<T>*, const vec4_t<T>*, int, int, T*, hipStream_t,    \
                                            unsigned*, int);                                                       \
  template hipError_t launch_dist_dedup<T>(const Cand*, int, int, int, const int*, const int*, const vec4_t<T>*,   \
                                           const vec4_t<T>*, int, int, int*, int*, int*, int*, vec4_t<T>*,         \
                                           vec4_t<T>*, int*, unsigned char*, hipStream_t, unsigned long long*,     \
                                           int);                                                                   \
  template hipError_t launch_isect_list<T>(int, const int*, const int*, int, int, const vec4_t<T>*,                \
                                           const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, int, uint64_t*,   \
                                           unsigned long long*, unsigned long long, hipStream_t);                  \
  template hipError_t launch_isect_pass<T>(int, const uint64_t*, const unsigned long long*, const int*, const int*, \
                                           int, int, const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*,         \
                                           const vec4_t<T>*, unsigned*, const unsigned long long*, unsigned*,      \
                                           IPair*, int, hipStream_t, unsigned long long*, unsigned long long);                                              \
  template hipError_t launch_isect_units<T>(const IPair*, unsigned long long, int, const int*, const int*,         \
                                            const vec4_t<T>*, const vec4_t<T>*, int, int, const unsigned*,         \
                                            const unsigned*, const unsigned long long*, const unsigned long long*, \
                                            int*, int*, int*, vec4_t<T>*, vec4_t<T>*, hipStream_t);

JSI_INST(float)
JSI_INST(double)
#undef JSI_INST


// ---------------------------------------------------------------------------
// Batched distance queries (jsi_internal.hpp).  Each kernel of the query runs
// once for the whole batch: blockIdx.y selects the query, whose arguments are
// one tuple of the body's parameters in a device array; a query whose failure
// flag is set returns at once.  The bodies are the single-query kernels'
// bodies, so every query computes exactly what it computes alone.
// ---------------------------------------------------------------------------
template <auto F>
struct BodyArgs;
template <class... P, void (*F)(P...)>
struct BodyArgs<F> {
  using tuple = std::tuple<std::decay_t<P>...>;
  using type = BArgs<std::decay_t<P>...>;
};

template <int LB, auto F>
__global__ __launch_bounds__(LB) void k_batched(const typename BodyArgs<F>::type* args) {
  const auto& x = args[blockIdx.y];
  if (x.fail && *x.fail) return;
  std::apply([](auto... a) { F(a...); }, x.t);
}
template <int LB, int WPE, auto F>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_batched_wpe(
    const typename BodyArgs<F>::type* args) {
  const auto& x = args[blockIdx.y];
  if (x.fail && *x.fail) return;
  std::apply([](auto... a) { F(a...); }, x.t);
}

template <auto F, class... X>
static typename BodyArgs<F>::type bargs(const unsigned* fail, X... x) {
  return typename BodyArgs<F>::type{typename BodyArgs<F>::tuple(x...), fail};
}

// Records one batched launch: grid (gx, queries), `block` threads.
template <int LB, auto F>
static void brec(BatchProgram& p, unsigned gx, unsigned block, const std::vector<typename BodyArgs<F>::type>& v) {
  if (v.empty() || gx == 0) return;
  const size_t off = p.put(v);
  p.add(reinterpret_cast<const void*>(&k_batched<LB, F>), dim3(gx, (unsigned)v.size()), dim3(block), off);
}
template <int LB, int WPE, auto F>
static void brec_wpe(BatchProgram& p, unsigned gx, unsigned block, const std::vector<typename BodyArgs<F>::type>& v) {
  if (v.empty() || gx == 0) return;
  const size_t off = p.put(v);
  p.add(reinterpret_cast<const void*>(&k_batched_wpe<LB, WPE, F>), dim3(gx, (unsigned)v.size()), dim3(block), off);
}

// Per query: the failure flag after a screen and its selection.  A query
// fails when its candidates overflowed the buffer (flag bit 1), when fewer
// than k pairs exist (bit 2), or when two narrowing steps left more records
// than the compaction holds (bit 4: many keys tied with the k-th); the
// single-query path handles all three by repeating steps.
__device__ __forceinline__ void k_batch_check_body(const unsigned long long* out_count, unsigned long long cap,
                                                   const unsigned long long* small_count, unsigned small_cap,
                                                   unsigned long long k, unsigned* fail) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long n = *out_count, m = *small_count;
  const unsigned why = (n > cap ? 1u : 0u) | (n < k ? 2u : 0u) | (m > (unsigned long long)small_cap ? 4u : 0u);
  if (why) *fail = why;
}

template <class T>
__device__ __forceinline__ int ec_blocks_d(const EcJob<T>& j) {
  if (j.nunits <= 0 || j.nuv < 2) return 0;
  const int tw = (j.nuv - 1 + kTile - 1) / kTile;
  return j.nunits * tw * tw;
}

// Fused evaluation + centres of many jobs: blockIdx.y = job.
template <class T>
struct EcBJob {
  EcJob<T> j;
  const unsigned* fail;
};
template <class T, int MAXD, bool RAT, bool KL>
__global__ __launch_bounds__(64) void k_eval_centres_b(const EcBJob<T>* jobs) {
  const EcBJob<T>& x = jobs[blockIdx.y];
  if (x.fail && *x.fail) return;
  if ((int)blockIdx.x >= ec_blocks_d(x.j)) return;
  ec_tile<T, MAXD, RAT, KL>(x.j.sd, x.j.reg, x.j.nunits, x.j.nuv, x.j.o, (int)blockIdx.x);
}

template <class T>
int eval_variant(const SurfDev<T>& sd) {
  EcJob<T> j;
  j.sd = sd;
  return ec_variant(j);
}
template int eval_variant<float>(const SurfDev<float>&);
template int eval_variant<double>(const SurfDev<double>&);

template <class T, int MAXD>
static void ec_brec_rat(BatchProgram& p, int var, unsigned gx, const std::vector<EcBJob<T>>& v) {
  const size_t off = p.put(v);
  const bool rat = (var / 3) % 2 == 1, kl = var >= 6;
  const void* f = rat ? (kl ? reinterpret_cast<const void*>(&k_eval_centres_b<T, MAXD, true, true>)
                            : reinterpret_cast<const void*>(&k_eval_centres_b<T, MAXD, true, false>))
                      : (kl ? reinterpret_cast<const void*>(&k_eval_centres_b<T, MAXD, false, true>)
                            : reinterpret_cast<const void*>(&k_eval_centres_b<T, MAXD, false, false>));
  p.add(f, dim3(gx, (unsigned)v.size()), dim3(64), off);
}

// Jobs grouped by kernel variant, one launch per variant present.
template <class T>
static void ec_brec(BatchProgram& p, const std::vector<EcBJob<T>>& jobs) {
  std::vector<int> vars;
  for (const auto& x : jobs) {
    const int v = ec_variant(x.j);
    if (std::find(vars.begin(), vars.end(), v) == vars.end()) vars.push_back(v);
  }
  for (int v : vars) {
    std::vector<EcBJob<T>> g;
    int gx = 0;
    for (const auto& x : jobs)
      if (ec_variant(x.j) == v) {
        g.push_back(x);
        gx = std::max(gx, ec_blocks(x.j));
      }
    if (gx <= 0) continue;
    switch (v % 3) {
      case 0: ec_brec_rat<T, 3>(p, v, (unsigned)gx, g); break;
      case 1: ec_brec_rat<T, 7>(p, v, (unsigned)gx, g); break;
      default: ec_brec_rat<T, kMaxDeg - 1>(p, v, (unsigned)gx, g); break;
    }
  }
}

// One query's view of a screening round.
template <class S>
struct TopkQ {
  const int *su1, *su2;
  const vec4_t<S>*C1, *C2, *lo1, *hi1, *lo2, *hi2;
  const double4 *R1, *R2;              // FP64 centres of the re-ranking (FP32SCREEN)
  const unsigned long long* exp_absmax;  // FP64: the expanded test's max |coordinate|
  const unsigned long long* absmax;      // FP32SCREEN: the widening's max |coordinate|
  uint64_t* list;
  unsigned* small;
  Cand* cand;
  Cand* topk;
  SelState* st;
  unsigned* hist;
  unsigned long long* range;
  unsigned long long* work;
  unsigned long long* best;
  Cand* sorted;
  Cand* last;
  unsigned* fail;
};

// The launches of one screen (dist_topk's sequence without its host round
// trips): seed threshold, tile-pair list, filter, re-rank, two narrowing
// steps, compaction, the failure check, rank sort.
template <class S>
static void topk_brec(BatchProgram& p, const DistBatchShape& sh, const std::vector<TopkQ<S>>& tq, int nslots,
                      int w1, int w2, bool rerank, bool exp) {
  const int B = (int)tq.size();
  const int fma = rerank ? 1 : 0;
  const int k = sh.k;
  const int tpu1 = tiles_per_unit_h(w1), tpu2 = tiles_per_unit_h(w2);
  const unsigned long long ntp = (unsigned long long)nslots * tpu1 * tpu2;
  const long long npos = (long long)nslots * tpu1;
  const int pairs_tp = std::min(w1 * w1, kTileCells) * std::min(w2 * w2, kTileCells);
  const unsigned long long want =
      ((unsigned long long)sh.tune.seed_pairs * (unsigned long long)k + pairs_tp - 1) / (unsigned long long)pairs_tp;
  const int nseed = nslots == 1 ? std::min(sh.tune.r1_seeds, tpu1) : (int)std::max(4ull, std::min(64ull, want));
  const unsigned long long cap = sh.cand_cap;
  const unsigned small_cap = sh.small_cap;
  auto thr = [](const TopkQ<S>& q) { return reinterpret_cast<S*>(q.small + 6); };
  auto out_count = [](const TopkQ<S>& q) { return reinterpret_cast<unsigned long long*>(q.small + 4); };
  auto small_count = [](const TopkQ<S>& q) { return reinterpret_cast<unsigned long long*>(q.small + 10); };
  bool counters_clear = false;
  if (npos <= 8192) {
    const int ordered = (nslots > 1 && tpu1 == 1 && tpu2 == 1) ? 1 : 0;
    if (!ordered) {
      using FT = std::integral_constant<decltype(&k_dist_seed_body<S>), &k_dist_seed_body<S>>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq) a.push_back(bargs<FT::value>(q.fail, nslots, q.su1, q.su2, w1, w2, q.lo1, q.hi1, q.lo2, q.hi2, k, q.best));
      brec<256, FT::value>(p, (unsigned)npos, 256, a);
      using GT = std::integral_constant<decltype(&k_seed_list_body<S>), &k_seed_list_body<S>>;
      std::vector<typename BodyArgs<GT::value>::type> b;
      for (const auto& q : tq) b.push_back(bargs<GT::value>(q.fail, q.best, (int)npos, tpu1, nseed, ordered, q.list, q.small));
      brec<1024, GT::value>(p, 1, 1024, b);
    }
    const int m = nseed < npos ? nseed : (int)npos;
    // the queries of a batch fill the chip together: about 4096 workgroups
    // over the batch, so fewer partial histograms are cleared and added to
    // the global ones than with the single-query split
    const long long mq = (long long)m * (long long)tq.size();
    int split = mq > 0 ? (int)std::min<long long>(kSeedSplitMax, 4096 / mq) : 1;
    split = split < 1 ? 1 : split;
    const int nb = m > 0 ? m * split : 1;
    auto gh = [](const TopkQ<S>& q) {
      return reinterpret_cast<unsigned*>(q.work + 2 * (size_t)kSeedMaxTiles * kSeedSplitMax);
    };
    auto hist_pass = [&](auto fh) {
      using FT = decltype(fh);
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq)
        a.push_back(bargs<FT::value>(q.fail, (const uint64_t*)q.list, m, ordered, q.su1, q.su2, w1, w2, q.C1, q.C2, split,
                             (const unsigned long long*)q.range, q.work, gh(q)));
      brec<256, FT::value>(p, (unsigned)nb, 256, a);
    };
    if (fma)
      hist_pass(std::integral_constant<decltype(&k_seed_hist_body<S, true, true>), &k_seed_hist_body<S, true, true>>());
    else
      hist_pass(
          std::integral_constant<decltype(&k_seed_hist_body<S, false, true>), &k_seed_hist_body<S, false, true>>());
    {
      using FT = std::integral_constant<decltype(&k_seed_range_body), &k_seed_range_body>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq) a.push_back(bargs<FT::value>(q.fail, (const unsigned long long*)q.work, nb, q.range, gh(q)));
      brec<1024, FT::value>(p, 1, 1024, a);
    }
    if (fma)
      hist_pass(
          std::integral_constant<decltype(&k_seed_hist_body<S, true, false>), &k_seed_hist_body<S, true, false>>());
    else
      hist_pass(
          std::integral_constant<decltype(&k_seed_hist_body<S, false, false>), &k_seed_hist_body<S, false, false>>());
    {
      using FT = std::integral_constant<decltype(&k_seed_thresh_body<S>), &k_seed_thresh_body<S>>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq)
        a.push_back(bargs<FT::value>(q.fail, (const unsigned*)gh(q), (const unsigned long long*)q.range, k, thr(q), q.small, 6));
      brec<1024, FT::value>(p, 1, kSeedHistBins, a);
    }
    if (rerank) {
      using FT = std::integral_constant<decltype(&k_inflate_body), &k_inflate_body>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq) a.push_back(bargs<FT::value>(q.fail, reinterpret_cast<float*>(thr(q)), q.absmax));
      brec<1024, FT::value>(p, 1, 64, a);
    }
    counters_clear = true;
  } else {
    using FT = std::integral_constant<decltype(&k_set_thresh_body<S>), &k_set_thresh_body<S>>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq) a.push_back(bargs<FT::value>(q.fail, thr(q), (const Cand*)q.cand, 0, 1));
    brec<1024, FT::value>(p, 1, 64, a);
  }
  auto fill = [&](auto ptr_of, unsigned long long n) {
    using FT = std::integral_constant<decltype(&k_fill_u32_body), &k_fill_u32_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq) a.push_back(bargs<FT::value>(q.fail, ptr_of(q), 0u, n));
    brec<1024, FT::value>(p, 1, 64, a);
  };
  if (!counters_clear) fill([](const TopkQ<S>& q) { return q.small; }, 1);
  {  // tile-pair list
    const long long rows = npos;
    auto list_pass = [&](auto fl, unsigned gx) {
      using FT = decltype(fl);
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq)
        a.push_back(bargs<FT::value>(q.fail, nslots, q.su1, q.su2, w1, w2, q.lo1, q.hi1, q.lo2, q.hi2, (const S*)thr(q),
                             sh.cull, q.list, q.small));
      brec<1024, FT::value>(p, gx, 256, a);
    };
    if (tpu2 >= 64) {
      if (fma)
        list_pass(std::integral_constant<decltype(&k_dist_list_rows_body<S, true>), &k_dist_list_rows_body<S, true>>(),
                  (unsigned)rows);
      else
        list_pass(
            std::integral_constant<decltype(&k_dist_list_rows_body<S, false>), &k_dist_list_rows_body<S, false>>(),
            (unsigned)rows);
    } else {
      const unsigned g = blocks_for((long long)ntp, 256);
      if (fma)
        list_pass(std::integral_constant<decltype(&k_dist_list_body<S, true>), &k_dist_list_body<S, true>>(), g);
      else
        list_pass(std::integral_constant<decltype(&k_dist_list_body<S, false>), &k_dist_list_body<S, false>>(), g);
    }
  }
  if (!counters_clear) fill([](const TopkQ<S>& q) { return q.small + 4; }, 2);
  // the filter
  if (w1 <= kTile && w2 <= kTile) {
    const unsigned gu = (unsigned)ntp;
    if (exp) {
      if constexpr (std::is_same<S, double>::value) {
        // as the single-query launch: workgroups walk at most 8 slots each
        const unsigned wk = (unsigned)std::min(16, std::max(1, sh.tune.filter_walk));  // <= kMaxWalk
        const unsigned gx = std::min(gu, std::max(1u, (gu + wk - 1u) / wk));
        const int nq = (w2 * w2 / 16) / 4;
        auto mf32 = [&](auto fm) {
          using FT = decltype(fm);
          std::vector<typename BodyArgs<FT::value>::type> a;
          for (const auto& q : tq)
            a.push_back(bargs<FT::value>(q.fail, (const uint64_t*)q.list, (const unsigned*)q.small, (int)gu, q.su1,
                                         q.su2, w1, w2, q.C1, q.C2, q.lo1, q.hi1, q.lo2, q.hi2, (const double*)thr(q),
                                         q.cand, out_count(q), cap, (unsigned long long*)nullptr));
          brec_wpe<256, 4, FT::value>(p, gx, 256, a);
        };
        auto mf = [&](auto fm, auto wpe) {
          using FT = decltype(fm);
          constexpr int W = decltype(wpe)::value;
          std::vector<typename BodyArgs<FT::value>::type> a;
          for (const auto& q : tq)
            a.push_back(bargs<FT::value>(q.fail, (const uint64_t*)q.list, (const unsigned*)q.small, (int)gu, q.su1, q.su2,
                                 w1, w2, q.C1, q.C2, (const double*)thr(q), q.cand, out_count(q), cap,
                                 (unsigned long long*)nullptr, q.exp_absmax));
          brec_wpe<256, W, FT::value>(p, gx, 256, a);
        };
#define JSI_MF(NQ)                                                                                         \
  if (sh.prec == JSI_PREC_FP32SCREEN)                                                                          \
    mf32(std::integral_constant<decltype(&k_dist_filter_mfma32_body<NQ>), &k_dist_filter_mfma32_body<NQ>>()); \
  else                                                                                                         \
    mf(std::integral_constant<decltype(&k_dist_filter_mfma_body<NQ>), &k_dist_filter_mfma_body<NQ>>(),        \
       std::integral_constant<int, (NQ < 4 ? 4 : 3)>())
        switch (nq) {
          case 0: JSI_MF(0); break;
          case 1: JSI_MF(1); break;
          case 2: JSI_MF(2); break;
          case 3: JSI_MF(3); break;
          default: JSI_MF(4); break;
        }
#undef JSI_MF
      }
    } else {
      auto fu = [&](auto ff) {
        using FT = decltype(ff);
        std::vector<typename BodyArgs<FT::value>::type> a;
        for (const auto& q : tq)
          a.push_back(bargs<FT::value>(q.fail, (const uint64_t*)q.list, (const unsigned*)q.small, q.su1, q.su2, w1, w2, q.C1,
                               q.C2, (const S*)thr(q), q.cand, out_count(q), cap, (unsigned long long*)nullptr,
                               (const unsigned long long*)nullptr));
        brec<256, FT::value>(p, gu, 256, a);
      };
      if (fma)
        fu(std::integral_constant<decltype(&k_dist_filter_units_body<S, true, false>),
                                  &k_dist_filter_units_body<S, true, false>>());
      else
        fu(std::integral_constant<decltype(&k_dist_filter_units_body<S, false, false>),
                                  &k_dist_filter_units_body<S, false, false>>());
    }
  } else {
    // round 1: persistent workgroups walking the query's listed tile pairs,
    // about twice the resident workgroups over the whole batch
    const unsigned long long g1 = 2ull * (unsigned long long)sh.grid * (unsigned long long)sh.tune.r1_scale / 100ull;
    const unsigned gx = (unsigned)std::min<unsigned long long>(
        ntp, std::max(16ull, (g1 + B - 1) / (unsigned long long)B));
    auto ft = [&](auto ff) {
      using FT = decltype(ff);
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq)
        a.push_back(bargs<FT::value>(q.fail, (const uint64_t*)q.list, (const unsigned*)q.small, q.su1, q.su2, w1, w2, q.C1,
                             q.C2, (const S*)thr(q), q.cand, out_count(q), cap, (unsigned long long*)nullptr, sh.cull,
                             exp ? q.exp_absmax : (const unsigned long long*)nullptr));
      brec<256, FT::value>(p, gx, 256, a);
    };
    if (exp) {
      if constexpr (std::is_same<S, double>::value)
        ft(std::integral_constant<decltype(&k_dist_filter_body<S, false, true>), &k_dist_filter_body<S, false, true>>());
    } else if (fma) {
      ft(std::integral_constant<decltype(&k_dist_filter_body<S, true, false>), &k_dist_filter_body<S, true, false>>());
    } else {
      ft(std::integral_constant<decltype(&k_dist_filter_body<S, false, false>),
                                &k_dist_filter_body<S, false, false>>());
    }
  }
  // selection over the query's own records (grid-stride over at most cap)
  const unsigned gs = (unsigned)std::max(1ull, std::min<unsigned long long>(64, (cap + 255) / 256));
  if (rerank) {
    using FT = std::integral_constant<decltype(&k_rerank_body), &k_rerank_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq)
      a.push_back(bargs<FT::value>(q.fail, q.cand, (const unsigned long long*)out_count(q), cap, q.su1, q.su2, w1 * w1, w2 * w2,
                           q.R1, q.R2));
    brec<1024, FT::value>(p, gs, 256, a);
  }
  const unsigned long long maxkey =
      rerank ? 0x7ff0000000000000ull : (std::is_same<S, float>::value ? 0x7f800000ull : 0x7ff0000000000000ull);
  {
    using FT = std::integral_constant<decltype(&k_sel_init_body), &k_sel_init_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq)
      a.push_back(bargs<FT::value>(q.fail, q.st, cap, (const unsigned long long*)out_count(q), (unsigned long long)k, maxkey,
                           small_cap, q.hist, small_count(q)));
    brec<256, FT::value>(p, 1, 256, a);
  }
  {
    using FT = std::integral_constant<decltype(&k_sel_minmax_body), &k_sel_minmax_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq) a.push_back(bargs<FT::value>(q.fail, (const Cand*)q.cand, cap, (const unsigned long long*)out_count(q), q.st));
    brec<256, FT::value>(p, gs, 256, a);
  }
  for (int it = 0; it < 2; ++it) {
    {
      using FT = std::integral_constant<decltype(&k_sel_hist_body), &k_sel_hist_body>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq)
        a.push_back(bargs<FT::value>(q.fail, (const Cand*)q.cand, cap, (const unsigned long long*)out_count(q),
                             (const SelState*)q.st, q.hist));
      brec<256, FT::value>(p, gs, 256, a);
    }
    {
      using FT = std::integral_constant<decltype(&k_sel_step_body), &k_sel_step_body>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : tq) a.push_back(bargs<FT::value>(q.fail, q.st, q.hist, small_cap));
      brec<1024, FT::value>(p, 1, 1024, a);
    }
  }
  {
    using FT = std::integral_constant<decltype(&k_sel_compact_body), &k_sel_compact_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq)
      a.push_back(bargs<FT::value>(q.fail, (const Cand*)q.cand, cap, (const unsigned long long*)out_count(q),
                           (const SelState*)q.st, q.topk, small_count(q), (unsigned long long)small_cap));
    brec<256, FT::value>(p, gs, 256, a);
  }
  {
    using FT = std::integral_constant<decltype(&k_batch_check_body), &k_batch_check_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq)
      a.push_back(bargs<FT::value>(q.fail, (const unsigned long long*)out_count(q), cap,
                           (const unsigned long long*)small_count(q), small_cap, (unsigned long long)k, q.fail));
    brec<1024, FT::value>(p, 1, 64, a);
  }
  if (tq.size() <= 4) {  // few queries: the rank sort spreads over the GPU
    using FT = std::integral_constant<decltype(&k_rank_sort_body), &k_rank_sort_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq)
      a.push_back(bargs<FT::value>(q.fail, (const Cand*)q.topk, (const unsigned long long*)small_count(q), small_cap,
                                   (unsigned)k, q.sorted, q.last));
    brec<1024, FT::value>(p, blocks_for(small_cap, 64), 64 * kRankWaves, a);
    return;
  }
  {  // sorted runs in the candidate buffer (free after the compaction)
    using FT = std::integral_constant<decltype(&k_run_sort_body), &k_run_sort_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq)
      a.push_back(bargs<FT::value>(q.fail, (const Cand*)q.topk, (const unsigned long long*)small_count(q), small_cap,
                                   q.cand));
    brec<kRun / 2, FT::value>(p, blocks_for(small_cap, kRun), kRun / 2, a);
  }
  {
    using FT = std::integral_constant<decltype(&k_run_rank_body), &k_run_rank_body>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : tq)
      a.push_back(bargs<FT::value>(q.fail, (const Cand*)q.cand, (const unsigned long long*)small_count(q), small_cap,
                                   (unsigned)k, q.sorted, q.last));
    brec<256, FT::value>(p, blocks_for(small_cap, 256), 256, a);
  }
}

template <class T>
int record_dist_batch(BatchProgram& p, const DistBatchShape& sh, const std::vector<DistBatchQ<T>>& qs) {
  using V = vec4_t<T>;
  const bool dbl = std::is_same<T, double>::value;
  const bool s32 = dbl && sh.prec == JSI_PREC_FP32SCREEN;
  const int k = sh.k;
  if (qs.empty()) return JSI_OK;
  if (k < 1 || k > kDedupMax || sh.rounds < 1 || sh.nuv1 < 2 || sh.nuv2 < 2 || sh.nref < 2) return JSI_ERR_INVALID;
  {
    using FT = std::integral_constant<decltype(&k_query_init_body<T>), &k_query_init_body<T>>;
    std::vector<typename BodyArgs<FT::value>::type> a;
    for (const auto& q : qs)
      a.push_back(bargs<FT::value>(q.fail, q.reg[0][0], q.g1, q.reg[1][0], q.g2, q.su[0][0], q.su[1][0],
                           dbl ? q.amax : (unsigned long long*)nullptr));
    brec<1024, FT::value>(p, 1, 64, a);
  }
  {  // round-1 evaluation of both sides
    std::vector<EcBJob<T>> jobs;
    for (int s = 0; s < 2; ++s)
      for (const auto& q : qs) {
        EcBJob<T> x;
        x.fail = q.fail;
        x.j.sd = s ? q.sd2 : q.sd1;
        x.j.reg = q.reg[s][0];
        x.j.nunits = 1;
        x.j.nuv = s ? sh.nuv2 : sh.nuv1;
        x.j.o.C = q.C[s][0];
        x.j.o.lo = q.tbl[s][0];
        x.j.o.hi = q.tbh[s][0];
        if (s32) {
          x.j.o.C32 = q.c32[s];
          x.j.o.lo32 = q.tb32l[s];
          x.j.o.hi32 = q.tb32h[s];
        }
        x.j.o.absmax = dbl ? q.amax : nullptr;
        jobs.push_back(x);
      }
    ec_brec<T>(p, jobs);
  }
  for (int round = 1; round <= sh.rounds; ++round) {
    const int pp = (round - 1) % 2, nx = round % 2;
    const int nuv1 = round == 1 ? sh.nuv1 : sh.nref, nuv2 = round == 1 ? sh.nuv2 : sh.nref;
    const int w1 = nuv1 - 1, w2 = nuv2 - 1;
    const int nslots = round == 1 ? 1 : k;
    if (s32 && round == 1) {
      std::vector<TopkQ<float>> tq;
      for (const auto& q : qs) {
        TopkQ<float> t{};
        t.su1 = q.su[0][pp];
        t.su2 = q.su[1][pp];
        t.C1 = q.c32[0];
        t.C2 = q.c32[1];
        t.lo1 = q.tb32l[0];
        t.hi1 = q.tb32h[0];
        t.lo2 = q.tb32l[1];
        t.hi2 = q.tb32h[1];
        t.R1 = reinterpret_cast<const double4*>(q.C[0][pp]);
        t.R2 = reinterpret_cast<const double4*>(q.C[1][pp]);
        t.absmax = q.amax;
        t.list = q.list; t.small = q.small; t.cand = q.cand; t.topk = q.topk; t.st = q.st; t.hist = q.hist;
        t.range = q.range; t.work = q.seedwork; t.best = q.best;
        t.sorted = q.sorted[round - 1];
        t.last = q.lastrec + (round - 1);
        t.fail = q.fail;
        tq.push_back(t);
      }
      topk_brec<float>(p, sh, tq, nslots, w1, w2, true, false);
    } else {
      std::vector<TopkQ<T>> tq;
      for (const auto& q : qs) {
        TopkQ<T> t{};
        t.su1 = q.su[0][pp];
        t.su2 = q.su[1][pp];
        t.C1 = q.C[0][pp];
        t.C2 = q.C[1][pp];
        t.lo1 = q.tbl[0][pp];
        t.hi1 = q.tbh[0][pp];
        t.lo2 = q.tbl[1][pp];
        t.hi2 = q.tbh[1][pp];
        t.exp_absmax = dbl ? q.amax : nullptr;
        t.list = q.list; t.small = q.small; t.cand = q.cand; t.topk = q.topk; t.st = q.st; t.hist = q.hist;
        t.range = q.range; t.work = q.seedwork; t.best = q.best;
        t.sorted = q.sorted[round - 1];
        t.last = q.lastrec + (round - 1);
        t.fail = q.fail;
        tq.push_back(t);
      }
      topk_brec<T>(p, sh, tq, nslots, w1, w2, false, dbl);
    }
    if (round == sh.rounds) break;
    // refinement units of the next round (launch_dist_dedup's sequence)
    {
      using FT = std::integral_constant<decltype(&k_dedup_mark_body), &k_dedup_mark_body>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : qs)
        a.push_back(bargs<FT::value>(q.fail, (const Cand*)(q.sorted[round - 1]), k, w1 * w1, w2 * w2, q.fp1, q.fp2));
      brec<256, FT::value>(p, blocks_for(k, 256), 256, a);
    }
    {
      using FT = std::integral_constant<decltype(&k_dedup_units_body), &k_dedup_units_body>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : qs)
        a.push_back(bargs<FT::value>(q.fail, (const Cand*)(q.sorted[round - 1]), k, w1 * w1, w2 * w2,
                             (const int*)q.fp1, (const int*)q.fp2, q.su[0][nx], q.su[1][nx], q.dwork,
                             q.nunits + 2 * (round - 1), dbl ? q.amax : (unsigned long long*)nullptr,
                             dbl ? kAmaxSlots : 0));
      brec<1024, FT::value>(p, 1, 1024, a);
    }
    {
      using FT = std::integral_constant<decltype(&k_dedup_apply_body<T>), &k_dedup_apply_body<T>>;
      std::vector<typename BodyArgs<FT::value>::type> a;
      for (const auto& q : qs)
        a.push_back(bargs<FT::value>(q.fail, (const Cand*)(q.sorted[round - 1]), k, w1 * w1, w2 * w2,
                             (const int*)q.su[0][pp], (const int*)q.su[1][pp], (const V*)q.reg[0][pp],
                             (const V*)q.reg[1][pp], nuv1, nuv2, (const unsigned char*)q.dwork,
                             (const int*)q.su[0][nx], (const int*)q.su[1][nx], q.reg[0][nx], q.reg[1][nx], q.fp1,
                             q.fp2));
      brec<256, FT::value>(p, blocks_for(k, 256), 256, a);
    }
    {
      std::vector<EcBJob<T>> jobs;
      for (int s = 0; s < 2; ++s)
        for (const auto& q : qs) {
          EcBJob<T> x;
          x.fail = q.fail;
          x.j.sd = s ? q.sd2 : q.sd1;
          x.j.reg = q.reg[s][nx];
          x.j.nunits = k;
          x.j.nuv = sh.nref;
          x.j.o.C = q.C[s][nx];
          x.j.o.lo = q.tbl[s][nx];
          x.j.o.hi = q.tbh[s][nx];
          x.j.o.absmax = dbl ? q.amax : nullptr;
          x.j.o.nunits_dev = q.nunits + 2 * (round - 1) + s;
          jobs.push_back(x);
        }
      ec_brec<T>(p, jobs);
    }
  }
  return JSI_OK;
}
template int record_dist_batch<float>(BatchProgram&, const DistBatchShape&, const std::vector<DistBatchQ<float>>&);
template int record_dist_batch<double>(BatchProgram&, const DistBatchShape&, const std::vector<DistBatchQ<double>>&);

}  // namespace jsi
