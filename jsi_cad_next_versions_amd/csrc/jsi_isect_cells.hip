// jsi_isect_cells.hip - the intersection refinement units of a round, for
// grids of nuv <= 8 points per side: evaluation with second derivatives,
// the unit's compute_K maxima, the padded cell AABBs and the unit's tile box
// in one pass.  File:line citations refer to JSI_CAD_NEXT_VERSION5/.
//
// After the first round, the refinement units are cells of the round before,
// sampled on nuv = 16 points per side, halved to 8 and then 4 once a round
// passes 512 pairs (intersection.cu:338-475).  k_points_tile gives each unit
// a 256-thread workgroup for a 16 x 16-point tile, so 16 or 64 of its threads
// hold a point and the fixed costs of a workgroup (knot staging, five
// dependent phases, barriers) set the speed; k_aabb and k_tilebox then read
// the points back from HBM.  Here a wavefront holds several units (a power
// of two lanes per unit, nuv^2 of them holding a point each) and carries
// them from their regions to their cells in registers and LDS:
//   1. spans and A2.3 rows (derivatives up to 2) of the units' columns and
//      rows, one per lane (the knots read through the L1 cache: staging
//      them in LDS would cost a load round trip, a barrier and the LDS that
//      limits the resident wavefronts);
//   2. the u-contractions of every (column, control row) of a unit, by the
//      unit's lanes, all u orders from one read of each control point;
//   3. each point adds its q+1 rows in s order for every (k, l) and applies
//      the quotient rule (EvaluateOneUV, evaluation.cu:259-356;
//      EG_EvaluateQuotientRule2, evaluation.cu:77-128);
//   4. the unit's compute_K maxima, a max over its lanes (common.cu:155-168);
//   5. each cell's AABB from its four corner points with the compute_K
//      padding (construct_aabbs_host, intersection.cu:106-148,187-206);
//   6. the unit's tile box (a unit of nuv <= 8 has one 16 x 16-cell tile).
// Every value goes through the operations k_points_tile, k_aabb and
// k_tilebox apply to it, in their order, so the cells are bit-identical.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "jsi_device.hpp"
#include "jsi_internal.hpp"
#include "jsi_nurbs.hpp"

namespace jsi {

namespace {

constexpr int kMaxDeg = 25;     // MAXDEG, evaluation.cu:13
constexpr int kIcMaxCols = 16;  // unit columns (and unit rows) per wavefront

// Lanes per unit: nuv^2 rounded up to a power of two, so a unit's lanes are
// an aligned group and xor shuffles below the group size stay inside it.
__host__ __device__ inline int ic_lanes(int nuv) {
  int l = 1;
  while (l < nuv * nuv) l <<= 1;
  return l;
}
// Units per wavefront: as many groups as fit, at most kIcMaxCols columns.
__host__ __device__ inline int ic_units_per_wave(int nuv) {
  const int a = 64 / ic_lanes(nuv), b = kIcMaxCols / nuv;
  return a < b ? a : b;
}

// control rows per chunk of u-contractions: a degree-3 unit inside one knot
// span reads 4 rows (one across a knot takes a second chunk)
constexpr int kIcRows = 4;

template <class T, int MAXD, bool RAT, bool BIC, bool PADK>
__global__ __launch_bounds__(64) void k_isect_cells(IcJob<T> j0, IcJob<T> j1, int nb0) {
  constexpr int NB = MAXD + 1;
  constexpr int NC = RAT ? 4 : 3;
  constexpr int NK = PADK ? 3 : 1;  // derivative orders kept per direction
  constexpr int kRows = kIcRows;
  // u-contractions of a column: kRows x NK entries of NC values, read as
  // 16-byte words; an odd number of such words per column puts the columns
  // that neighbouring lanes read on distinct banks
  constexpr int kRw = kRows * NK * NC;
  static_assert(kRw * sizeof(T) % 16 == 0, "column of u-contractions: whole 16-byte words");
  constexpr int kRs = (kRw * (int)sizeof(T) / 16) % 2 == 0 ? kRw + 16 / (int)sizeof(T) : kRw;
  constexpr int kRsz = kIcMaxCols * kRs;
  constexpr int kA = kRsz > 3 * 64 ? kRsz : 3 * 64;
  __shared__ T sN[2 * kIcMaxCols][NK][NB];  // A2.3 rows: columns [0, 16), rows [16, 32)
  __shared__ int sS[2 * kIcMaxCols];        // their spans
  // two lives: u-contractions [column][row][k][c] (phases 2-3), then the
  // points' coordinates [c][lane] (phase 5); a barrier separates them
  __shared__ T sA[kA];
  // entry (column, row, u order) of the u-contractions
  const auto sR = [&](int cx, int rr, int k) -> T* { return sA + cx * kRs + (rr * NK + k) * NC; };
  const bool second = (int)blockIdx.x >= nb0;  // uniform: one body, its arguments selected
  const SurfDev<T> sd = second ? j1.sd : j0.sd;
  const vec4_t<T>* const regions = second ? j1.reg : j0.reg;
  const int nunits = second ? j1.nunits : j0.nunits;
  const int nuv = second ? j1.nuv : j0.nuv;
  const int bid = second ? (int)blockIdx.x - nb0 : (int)blockIdx.x;
  const int lane = threadIdx.x;
  const int L = ic_lanes(nuv), upw = ic_units_per_wave(nuv);
  const int sub = lane / L, e = lane % L;  // unit of the wavefront, lane in the unit
  const int unit = bid * upw + sub;
  const int npts = nuv * nuv, w = nuv - 1;
  const bool live = sub < upw && unit < nunits;
  const bool active = live && e < npts;  // the lane holds point e of its unit
  const int iy = e / nuv, ix = e % nuv;
  const int p = sd.degu, q = sd.degv;
  const int du = p < 2 ? p : 2, dv = q < 2 ? q : 2;
  const int ku = PADK ? du : 0;  // u orders formed
  const int lv = PADK ? dv : 0;  // v orders formed
  const int ncol = upw * nuv;    // columns (and rows) of the wavefront's units
  // phase 1: spans and basis rows; the parameter follows the grid driver
  // (evaluation.cu:366) on the unit's own nuv x nuv grid
  if (!sd.zero && lane < 2 * ncol) {
    const int dir = lane >= ncol;
    const int t = dir ? lane - ncol : lane;
    const int un = bid * upw + t / nuv, idx = t % nuv;
    if (un < nunits) {
      const vec4_t<T> r = regions[un];
      const T start = dir ? r.y : r.x;
      const T step = ((dir ? r.w : r.z) - start) / T(nuv);
      const T tt = step * T(idx) + start;
      const int deg = dir ? q : p, der = PADK ? (dir ? dv : du) : 0;
      const T* knots = dir ? sd.V : sd.U;
      const int span = find_span(dir ? sd.nkv : sd.nku, deg, tt, knots);
      const int slot = (dir ? kIcMaxCols : 0) + t;
      if (BIC)  // degree 3 both ways: one recurrence, no degree switch
        basis_row_fixed<T, 3, NK, NB>(span, tt, knots, der, sN[slot]);
      else
        basis_row_any<T, MAXD, NK, NB>(span, deg, tt, knots, der, sN[slot]);
      sS[slot] = span;
    }
  }
  __syncthreads();
  // control rows a unit reads: [vlo - q, vhi] over its rows' spans (a
  // reduction over the unit's lane group), and the wavefront's largest count
  int vlo = INT_MAX, vhi = INT_MIN;
  if (live && !sd.zero && e < nuv) vlo = vhi = sS[kIcMaxCols + sub * nuv + e];
  for (int off = L >> 1; off > 0; off >>= 1) {
    vlo = min(vlo, __shfl_xor(vlo, off));
    vhi = max(vhi, __shfl_xor(vhi, off));
  }
  const int v0 = vlo - q;
  const int nrows = (live && !sd.zero) ? vhi - vlo + q + 1 : 0;
  int nrw = nrows;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nrw = max(nrw, __shfl_xor(nrw, off));
  T h[6][4];
#pragma unroll
  for (int m = 0; m < 6; ++m)
#pragma unroll
    for (int c = 0; c < 4; ++c) h[m][c] = T(0);
  const int cpt = sub * nuv + ix;  // the point's column slot
  const int rpt = kIcMaxCols + sub * nuv + iy;
  const int sv = active && !sd.zero ? sS[rpt] : 0;
  for (int r0 = 0; r0 < nrw; r0 += kRows) {
    const int nr = min(kRows, nrows - r0);  // <= 0 once the unit's rows are done
    // phase 2: u-contractions of the unit's rows [r0, r0 + nr) for each of
    // its columns, all u orders from one read of each control point
    for (int task = e; task < nuv * nr; task += L) {
      const int rr = task % nr, cx = task / nr;
      const int cs = sub * nuv + cx;
      const int base = sd.ncpu * (v0 + r0 + rr) + sS[cs] - p;
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
          const T n = sN[cs][k][j];
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
          for (int c = 0; c < NC; ++c) sR(cs, rr, k)[c] = acc[k][c];
    }
    __syncthreads();
    // phase 3: each point adds this chunk's rows, in s order, for every (k, l)
    if (active && !sd.zero) {
      if (nrows <= kRows) {  // the first chunk holds every row of the unit:
                             // the point's rows are sv - vlo + s2, all present
                             // (later chunks, run for other units, add none)
        const int rb = sv - vlo;
        if (r0 == 0) {
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
                const T nv = sN[rpt][l][s2];
#pragma unroll
                for (int c = 0; c < NC; ++c) h[m][c] += nv * sR(cpt, rb + s2, k)[c];
              }
            }
        }
      } else {
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
              const int rr = sv - q + s2 - v0 - r0;
              if (rr < 0 || rr >= nr) continue;
              const T nv = sN[rpt][l][s2];
#pragma unroll
              for (int c = 0; c < NC; ++c) h[m][c] += nv * sR(cpt, rr, k)[c];
            }
          }
      }
    }
    __syncthreads();  // sR is rewritten by the next chunk (or becomes the points)
  }
  if (RAT && active && !sd.zero) {
    // EG_EvaluateQuotientRule2 for dim 3, der 2, stride 4 (evaluation.cu:77-128)
    const T F = T(1) / h[0][3];
    if (!PADK) {
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
  // phase 4: compute_K maxima of the unit's points (max is exact and
  // order-free) and the padding of its cells (common.cu:170-171)
  T pad = T(0);
  if (PADK) {
    unsigned long long mk[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      const T v = cmax(fabs(h[3 + g][0]), cmax(fabs(h[3 + g][1]), fabs(h[3 + g][2])));
      mk[g] = active ? ord_bits(v) : 0ull;
      for (int off = L >> 1; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mk[g], off);
        mk[g] = o > mk[g] ? o : mk[g];
      }
    }
    if (live) {
      const T muu = from_bits<T>(mk[0]);
      const T muv = from_bits<T>(mk[1]);
      const T mvv = from_bits<T>(mk[2]);
      const int side = nuv > 1 ? nuv - 1 : 1;
      const T s2 = T(side * side);
      const T raw = (muu / s2 + T(2) * muv / s2 + mvv / s2) / T(8);
      const vec4_t<T> r = regions[unit];
      const T init_u = second ? j1.init_u : j0.init_u, init_v = second ? j1.init_v : j0.init_v;
      const T us = fabs(r.z - r.x) / init_u;
      const T vs = fabs(r.w - r.y) / init_v;
      const T ps = smax(us, vs);
      pad = raw * ps * ps;
      if (!isfinite(pad) || pad < T(0)) pad = T(0);
    }
  }
  // phase 5: the points into LDS, then each cell's AABB from its corners
  // (cell id = iv * (nuv - 1) + iu, corners (iv, iu), (iv, iu + 1),
  // (iv + 1, iu), (iv + 1, iu + 1)) and the unit's box of its padded cells
  T* const sP = sA;
  if (active) {
    sP[lane] = h[0][0];
    sP[64 + lane] = h[0][1];
    sP[128 + lane] = h[0][2];
  }
  __syncthreads();
  const T inf = infinity_v<T>();
  T mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
  T amax = T(0);
  // FP32SCREEN: the cells' float records and the unit's float box of them
  // (k_isect_boxes32 and k_tilebox<float>, in their operations)
  float4* const X32 = second ? j1.X : j0.X;
  const float finf = __uint_as_float(0x7f800000u);
  float fmn[3] = {finf, finf, finf}, fmx[3] = {-finf, -finf, -finf};
  float famax = 0.f;
  if (active && ix < w && iy < w) {
    const int p0 = lane, p1 = p0 + 1, p2 = p0 + nuv, p3 = p2 + 1;
    T lo[3], hi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T* b = sP + 64 * c;
      const T a0 = b[p0], a1 = b[p1], a2 = b[p2], a3 = b[p3];
      lo[c] = smin(a0, smin(a1, smin(a2, a3)));
      hi[c] = smax(a0, smax(a1, smax(a2, a3)));
    }
    vec4_t<T> ctr, rad;
    ctr.x = (lo[0] + hi[0]) / T(2);
    ctr.y = (lo[1] + hi[1]) / T(2);
    ctr.z = (lo[2] + hi[2]) / T(2);
    ctr.w = T(0);
    rad.x = (hi[0] - lo[0]) / T(2) + pad;
    rad.y = (hi[1] - lo[1]) / T(2) + pad;
    rad.z = (hi[2] - lo[2]) / T(2) + pad;
    rad.w = T(0);
    vec4_t<T>* const C = second ? j1.C : j0.C;
    vec4_t<T>* const R = second ? j1.R : j0.R;
    const size_t o = (size_t)unit * w * w + iy * w + ix;
    C[o] = ctr;
    R[o] = rad;
    const T v[3] = {ctr.x, ctr.y, ctr.z}, rr[3] = {rad.x, rad.y, rad.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = fmin(mn[a], v[a] - rr[a]);
      mx[a] = fmax(mx[a], v[a] + rr[a]);
      amax = fmax(amax, fabs(v[a]) + fabs(rr[a]));
    }
    if constexpr (std::is_same<T, double>::value) {
      if (X32) {  // uniform: a kernel argument
        float4 x, y;
        record32_axis(ctr.x, rad.x, false, 0.f, 0.f, &x.x, &y.x);
        record32_axis(ctr.y, rad.y, false, 0.f, 0.f, &x.y, &y.y);
        record32_axis(ctr.z, rad.z, false, 0.f, 0.f, &x.z, &y.z);
        x.w = y.w = 0.f;
        X32[o] = x;
        (second ? j1.Y : j0.Y)[o] = y;
        const float xv[3] = {x.x, x.y, x.z}, yv[3] = {y.x, y.y, y.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          fmn[a] = fminf(fmn[a], xv[a] - yv[a]);
          fmx[a] = fmaxf(fmx[a], xv[a] + yv[a]);
          famax = fmaxf(famax, fabsf(xv[a]) + fabsf(yv[a]));
        }
      }
    }
  }
  if (X32) {  // uniform: a kernel argument
    for (int off = L >> 1; off > 0; off >>= 1) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        fmn[a] = fminf(fmn[a], __shfl_xor(fmn[a], off));
        fmx[a] = fmaxf(fmx[a], __shfl_xor(fmx[a], off));
      }
      famax = fmaxf(famax, __shfl_xor(famax, off));
    }
    if (live && e == 0) {
      // the outward margin of k_tilebox<float>
      const float delta = famax * 9.5367431640625e-07f + 1e-30f;
      (second ? j1.flo : j0.flo)[unit] = make_float4(fmn[0] - delta, fmn[1] - delta, fmn[2] - delta, 0.f);
      (second ? j1.fhi : j0.fhi)[unit] = make_float4(fmx[0] + delta, fmx[1] + delta, fmx[2] + delta, 0.f);
    }
  }
  vec4_t<T>* const lo = second ? j1.lo : j0.lo;
  if (lo) {  // uniform: a kernel argument
    for (int off = L >> 1; off > 0; off >>= 1) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = fmin(mn[a], __shfl_xor(mn[a], off));
        mx[a] = fmax(mx[a], __shfl_xor(mx[a], off));
      }
      amax = fmax(amax, __shfl_xor(amax, off));
    }
    if (live && e == 0) {
      // outward margin of k_tilebox: 2^-20 of the largest |coordinate|
      const T delta = amax * T(9.5367431640625e-07) + T(1e-30);
      vec4_t<T> l, hh;
      l.x = mn[0] - delta;
      l.y = mn[1] - delta;
      l.z = mn[2] - delta;
      l.w = T(0);
      hh.x = mx[0] + delta;
      hh.y = mx[1] + delta;
      hh.z = mx[2] + delta;
      hh.w = T(0);
      vec4_t<T>* const hi = second ? j1.hi : j0.hi;
      lo[unit] = l;
      hi[unit] = hh;
    }
  }
}

// ---------------------------------------------------------------------------
// Overlap screen and refinement bookkeeping of a round whose units have at
// most 64 cells per side (nuv <= 9, one tile per unit), per slot.
//
// After a dense round the next one has as many slots as the dense round had
// pairs (10^7 after a 64 x 64 first grid whose padded cells mostly overlap),
// nearly all of them without a hit, so nothing here is sized per row (slot,
// i), and most slots end at the unit-box test (the tile-pair cull of
// k_isect_list with one tile per unit), which for such rounds runs first,
// one slot per lane, and lists the slots that pass in order.  A group of
// lanes then holds one slot, a lane per side-1 cell i (cells1
// rounded up to a power of two lanes); a lane's hits over the side-2 cells j
// form a bit mask, j ascending (the strict 3-axis test of overlap_aabb_host,
// intersection.cu:46-63).  k_isect_slot_count
// writes three counts per slot: its pairs, its non-empty rows (the side-1
// units it creates) and its distinct j (its side-2 units).  After their
// scans, k_isect_slot_emit recomputes the masks and writes the slot's pairs
// in (i, j) order at the slot's offset, so the list is in (slot, i, j) order
// with no atomics and no sort, and, when a round follows, the refinement
// units: units are numbered by first appearance in pair order, (slot, i) on
// side 1 and (slot, j) on side 2 (the reuse maps of intersection.cu:384-446),
// so a side-1 unit is its slot's offset plus the non-empty rows before it,
// and a side-2 unit its slot's offset plus the distinct j of the rows before
// its first row plus the new j below it in that row.
// ---------------------------------------------------------------------------
template <class T>
struct SlotScreen {
  long long nslots;
  int cells1, cells2;
  const int *su1, *su2;
  const vec4_t<T>*C1, *R1, *C2, *R2;
  const vec4_t<T>*lo1, *hi1, *lo2, *hi2;  // unit boxes (cull)
  int cull;
  SlotF32 f;  // FP32SCREEN (f.X1 non-null): float records and unit boxes
};
template <class T>
struct SlotUnits {
  const vec4_t<T>*reg1, *reg2;  // the slots' parent regions
  int pnuv1, pnuv2;             // their grids
  int *nsu1, *nsu2;             // per pair: its units
  vec4_t<T>*nreg1, *nreg2;      // per unit: its region
};

__host__ __device__ inline int slot_lanes(int cells1) {
  int l = 1;
  while (l < cells1) l <<= 1;
  return l;
}
// slot groups per wavefront; the emit pass keeps cells2 unit ids per group
// in LDS, at most 1024 per wavefront
__host__ __device__ inline int slot_groups(int cells1, int cells2) {
  const int a = 64 / slot_lanes(cells1), b = 1024 / cells2;
  return a < b ? a : (b > 0 ? b : 1);
}

// The unit-box test of one slot (the tile-pair cull of k_isect_list; the
// float boxes of the records under FP32SCREEN).
template <class B>
__device__ __forceinline__ bool boxes_meet(const B& l1, const B& h1, const B& l2, const B& h2) {
  return l1.x <= h2.x && l2.x <= h1.x && l1.y <= h2.y && l2.y <= h1.y && l1.z <= h2.z && l2.z <= h1.z;
}
template <class T>
__device__ __forceinline__ bool slot_boxes_meet(const SlotScreen<T>& s, long long slot) {
  if (!s.cull) return true;
  const int u1 = s.su1[slot], u2 = s.su2[slot];
  if (s.f.X1) return boxes_meet(s.f.lo1[u1], s.f.hi1[u1], s.f.lo2[u2], s.f.hi2[u2]);
  return boxes_meet(s.lo1[u1], s.hi1[u1], s.lo2[u2], s.hi2[u2]);
}

// The strict 3-axis overlap test of overlap_aabb_host (intersection.cu:46-63)
// in the records' own arithmetic (k_isect_pass's expression).
template <class V>
__device__ __forceinline__ bool cells_overlap(const V& c1, const V& r1, const V& c2, const V& r2) {
  // the three axis tests of overlap_aabb_host, evaluated without branches
  return (fabs(c1.x - c2.x) < (r1.x + r2.x)) & (fabs(c1.y - c2.y) < (r1.y + r2.y)) &
         (fabs(c1.z - c2.z) < (r1.z + r2.z));
}

// The hits of row (slot, i): bit j = cell j of the side-2 unit.  Under
// FP32SCREEN the float records' test picks the candidates first and the FP64
// test decides them (k_isect_retest's decision), so the hits are the FP64
// screen's.
template <class T>
__device__ __forceinline__ unsigned long long slot_row_mask(const SlotScreen<T>& s, long long slot, int i) {
  const int u1 = s.su1[slot], u2 = s.su2[slot];
  const size_t a = (size_t)u1 * s.cells1 + i, b0 = (size_t)u2 * s.cells2;
  unsigned long long cand = ~0ull >> (64 - s.cells2);
  if (s.f.X1) {
    const float4 x1 = s.f.X1[a], y1 = s.f.Y1[a];
    cand = 0;
    for (int j = 0; j < s.cells2; ++j)
      cand |= (unsigned long long)cells_overlap(x1, y1, s.f.X2[b0 + j], s.f.Y2[b0 + j]) << j;
    if (!cand) return 0;
  }
  const vec4_t<T> c1 = s.C1[a], r1 = s.R1[a];
  unsigned long long m = 0;
  if (!s.f.X1) {  // every cell a candidate: a counted loop, whose loads run ahead
#pragma unroll 4
    for (int j = 0; j < s.cells2; ++j)
      m |= (unsigned long long)cells_overlap(c1, r1, s.C2[b0 + j], s.R2[b0 + j]) << j;
    return m;
  }
  for (unsigned long long r = cand; r; r &= r - 1) {
    const int j = __ffsll((long long)r) - 1;
    m |= (unsigned long long)cells_overlap(c1, r1, s.C2[b0 + j], s.R2[b0 + j]) << j;
  }
  return m;
}

// Rounds with many slots are culled first, one slot per lane: a wavefront
// counts its 64 slots that pass (k_isect_slot_cull), and after a scan of the
// counts lists them in order (k_isect_slot_list).  The count and emit passes
// then run over the list (entry e holds slot list[e]; their counts and
// offsets are per entry), or, without a list, over every slot with the
// unit-box test in each group.
template <class T>
__global__ __launch_bounds__(256) void k_isect_slot_cull(SlotScreen<T> s, unsigned* wcount) {
  const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool pass = slot < s.nslots && slot_boxes_meet(s, slot);
  const unsigned long long m = __ballot(pass);
  if ((threadIdx.x & 63) == 0 && slot < s.nslots) wcount[slot >> 6] = (unsigned)__popcll(m);
}

template <class T>
__global__ __launch_bounds__(256) void k_isect_slot_list(SlotScreen<T> s, const unsigned long long* woff,
                                                         int* list) {
  const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool pass = slot < s.nslots && slot_boxes_meet(s, slot);
  const unsigned long long m = __ballot(pass);
  const int lane = threadIdx.x & 63;
  if (pass) list[woff[slot >> 6] + __popcll(m & ((1ull << lane) - 1))] = (int)slot;
}

// n entries: slots 0 .. n-1, or list[0 .. n-1]
template <class T>
__global__ __launch_bounds__(256) void k_isect_slot_count(SlotScreen<T> s, long long n, const int* list,
                                                          unsigned* npair, unsigned* nrow, unsigned* ncol,
                                                          unsigned long long* masks, int stage) {
  // stage (FP64 screen, launched with the LDS for it): each wavefront copies
  // the side-2 cells of its slots into LDS once, and the rows test them from
  // there instead of every lane loading each cell from device memory
  extern __shared__ unsigned char sdyn[];
  const int G = slot_lanes(s.cells1), spw = slot_groups(s.cells1, s.cells2);
  const long long nwaves = (n + spw - 1) / spw;
  const int lane = threadIdx.x & 63, grp = lane / G, i = lane % G;
  const int nst = spw * s.cells2;  // staged cells per wavefront
  vec4_t<T>* const stC = reinterpret_cast<vec4_t<T>*>(sdyn) + (size_t)(threadIdx.x >> 6) * 2 * nst;
  vec4_t<T>* const stR = stC + nst;
  for (long long base = (long long)blockIdx.x * 4; base < nwaves; base += (long long)gridDim.x * 4) {
    const long long wbase = (base + (threadIdx.x >> 6)) * spw;  // the wavefront's first entry
    const long long e = wbase + grp;
    const bool in = grp < spw && e < n;
    const long long slot = in ? (list ? (long long)list[e] : e) : 0;
    const bool valid = in && i < s.cells1 && (list || slot_boxes_meet(s, slot));
    unsigned long long m = 0;
    if (stage) {  // uniform: a kernel argument
      // FP32SCREEN stages the float records (the FP64 test of the candidates
      // reads its cells from device memory), the FP64 screen the cells
      float4* const sfX = reinterpret_cast<float4*>(stC);
      float4* const sfY = sfX + nst;
      for (int x = lane; x < nst; x += 64) {
        const int g = x / s.cells2, j = x - g * s.cells2;
        if (wbase + g < n) {
          const long long sg = list ? (long long)list[wbase + g] : wbase + g;
          const size_t b = (size_t)s.su2[sg] * s.cells2 + j;
          if (s.f.X1) {
            sfX[x] = s.f.X2[b];
            sfY[x] = s.f.Y2[b];
          } else {
            stC[x] = s.C2[b];
            stR[x] = s.R2[b];
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (valid) {
        const size_t a = (size_t)s.su1[slot] * s.cells1 + i;
        if (s.f.X1) {  // slot_row_mask's two tests, the float records from LDS
          const float4 x1 = s.f.X1[a], y1 = s.f.Y1[a];
          const float4* const gx = sfX + grp * s.cells2;
          const float4* const gy = sfY + grp * s.cells2;
          unsigned long long cand = 0;
#pragma unroll 4
          for (int j = 0; j < s.cells2; ++j) cand |= (unsigned long long)cells_overlap(x1, y1, gx[j], gy[j]) << j;
          if (cand) {
            const vec4_t<T> c1 = s.C1[a], r1 = s.R1[a];
            const size_t b0 = (size_t)s.su2[slot] * s.cells2;
            for (unsigned long long r = cand; r; r &= r - 1) {
              const int j = __ffsll((long long)r) - 1;
              m |= (unsigned long long)cells_overlap(c1, r1, s.C2[b0 + j], s.R2[b0 + j]) << j;
            }
          }
        } else {
          const vec4_t<T> c1 = s.C1[a], r1 = s.R1[a];
          const vec4_t<T>* const gc = stC + grp * s.cells2;
          const vec4_t<T>* const gr = stR + grp * s.cells2;
#pragma unroll 4
          for (int j = 0; j < s.cells2; ++j) m |= (unsigned long long)cells_overlap(c1, r1, gc[j], gr[j]) << j;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();  // the next entries' cells overwrite the staging
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      m = valid ? slot_row_mask(s, slot, i) : 0ull;
    }
    if (masks && in && i < s.cells1) masks[e * s.cells1 + i] = m;  // for the emit pass
    unsigned pc = (unsigned)__popcll(m), nz = m != 0;
    unsigned long long om = m;
    for (int off = G >> 1; off > 0; off >>= 1) {
      pc += __shfl_xor(pc, off);
      nz += __shfl_xor(nz, off);
      om |= __shfl_xor(om, off);
    }
    if (in && i == 0) {
      npair[e] = pc;
      nrow[e] = nz;
      ncol[e] = (unsigned)__popcll(om);
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_isect_slot_emit(SlotScreen<T> s, long long n, const int* list,
                                                         const unsigned long long* offp,
                                                         const unsigned long long* offr,
                                                         const unsigned long long* offc, IPair* out, int units,
                                                         SlotUnits<T> u, const unsigned long long* masks) {
  extern __shared__ int sTab[];  // per wavefront: spw groups x cells2 side-2 unit ids
  const int G = slot_lanes(s.cells1), spw = slot_groups(s.cells1, s.cells2);
  const long long nwaves = (n + spw - 1) / spw;
  const int lane = threadIdx.x & 63, grp = lane / G, i = lane % G;
  int* const tab = sTab + ((threadIdx.x >> 6) * spw + (grp < spw ? grp : 0)) * s.cells2;
  for (long long base = (long long)blockIdx.x * 4; base < nwaves; base += (long long)gridDim.x * 4) {
    const long long e = (base + (threadIdx.x >> 6)) * spw + grp;
    const bool in = grp < spw && e < n;
    const long long slot = in ? (list ? (long long)list[e] : e) : 0;
    // the rows' hit masks: written by the count pass, or recomputed
    const bool row = in && i < s.cells1;
    const bool valid = row && (masks || list || slot_boxes_meet(s, slot));
    const unsigned long long m = masks ? (row ? masks[e * s.cells1 + i] : 0ull)
                                       : (valid ? slot_row_mask(s, slot, i) : 0ull);
    // inclusive scans over the group's rows: pairs, non-empty rows, j seen
    unsigned pc = (unsigned)__popcll(m), nz = m != 0;
    unsigned long long om = m;
    for (int off = 1; off < G; off <<= 1) {
      const unsigned a = __shfl_up(pc, off, G), b = __shfl_up(nz, off, G);
      const unsigned long long c = __shfl_up(om, off, G);
      if (i >= off) {
        pc += a;
        nz += b;
        om |= c;
      }
    }
    unsigned long long prior = __shfl_up(om, 1, G);  // j of the rows before
    if (i == 0) prior = 0;
    const unsigned long long p0 = valid ? offp[e] + (pc - (unsigned)__popcll(m)) : 0ull;
    int u1 = 0;
    if (valid && m) {
      IPair pr;
      pr.slot = (int)slot;
      pr.i = i;
      unsigned long long p = p0;
      for (unsigned long long r = m; r; r &= r - 1) {
        pr.j = __ffsll((long long)r) - 1;
        out[p++] = pr;
      }
      if (units) {
        u1 = (int)(offr[e] + (nz - 1));
        u.nreg1[u1] = child_region<T>(u.reg1[s.su1[slot]], u.pnuv1, i);
        const unsigned long long nw = m & ~prior;
        const unsigned long long ub = offc[e] + (unsigned)__popcll(prior);
        const vec4_t<T> preg2 = u.reg2[s.su2[slot]];
        for (unsigned long long r = nw; r; r &= r - 1) {
          const int j = __ffsll((long long)r) - 1;
          const int idx = (int)(ub + (unsigned)__popcll(nw & ((1ull << j) - 1)));
          u.nreg2[idx] = child_region<T>(preg2, u.pnuv2, j);
          tab[j] = idx;
        }
      }
    }
    if (units) {  // uniform: a kernel argument
      __syncthreads();
      if (valid && m) {
        unsigned long long p = p0;
        for (unsigned long long r = m; r; r &= r - 1) {
          u.nsu1[p] = u1;
          u.nsu2[p] = tab[__ffsll((long long)r) - 1];
          ++p;
        }
      }
      __syncthreads();  // the table is rewritten by the next slots
    }
  }
}

template <class T>
int ic_blocks(const IcJob<T>& j) {
  if (j.nunits <= 0 || j.nuv < 2 || j.nuv > kIcMaxNuv) return 0;
  const int upw = ic_units_per_wave(j.nuv);
  return (j.nunits + upw - 1) / upw;
}

// variant of a job: 0-3 = degree class (3, 5, 7, kMaxDeg - 1), +4 rational,
// +8 bicubic (degree class 0 only)
template <class T>
int ic_variant(const IcJob<T>& j) {
  const int d = std::max(j.sd.degu, j.sd.degv);
  const int dc = d <= 3 ? 0 : d <= 5 ? 1 : d <= 7 ? 2 : 3;
  const bool bic = j.sd.degu == 3 && j.sd.degv == 3 && !j.sd.zero;
  return dc + (j.sd.rational ? 4 : 0) + (bic ? 8 : 0);
}

template <class T, int MAXD, bool RAT, bool BIC>
void ic_launch_k(int use_k, const IcJob<T>& j0, const IcJob<T>& j1, int nb0, int nb, hipStream_t s) {
  const dim3 g((unsigned)nb), b(64);
  if (use_k)
    hipLaunchKernelGGL((k_isect_cells<T, MAXD, RAT, BIC, true>), g, b, 0, s, j0, j1, nb0);
  else
    hipLaunchKernelGGL((k_isect_cells<T, MAXD, RAT, BIC, false>), g, b, 0, s, j0, j1, nb0);
}

template <class T, int MAXD>
void ic_launch_rat(int var, int use_k, const IcJob<T>& j0, const IcJob<T>& j1, int nb0, int nb, hipStream_t s) {
  const bool rat = (var / 4) % 2 == 1, bic = var >= 8;
  if constexpr (MAXD == 3) {
    if (bic) {
      if (rat)
        ic_launch_k<T, MAXD, true, true>(use_k, j0, j1, nb0, nb, s);
      else
        ic_launch_k<T, MAXD, false, true>(use_k, j0, j1, nb0, nb, s);
      return;
    }
  }
  if (rat)
    ic_launch_k<T, MAXD, true, false>(use_k, j0, j1, nb0, nb, s);
  else
    ic_launch_k<T, MAXD, false, false>(use_k, j0, j1, nb0, nb, s);
}

template <class T>
void ic_launch(int var, int use_k, const IcJob<T>& j0, const IcJob<T>& j1, int nb0, int nb, hipStream_t s) {
  if (nb <= 0) return;
  switch (var % 4) {
    case 0: ic_launch_rat<T, 3>(var, use_k, j0, j1, nb0, nb, s); break;
    case 1: ic_launch_rat<T, 5>(var, use_k, j0, j1, nb0, nb, s); break;
    case 2: ic_launch_rat<T, 7>(var, use_k, j0, j1, nb0, nb, s); break;
    default: ic_launch_rat<T, kMaxDeg - 1>(var, use_k, j0, j1, nb0, nb, s); break;
  }
}

}  // namespace

template <class T>
hipError_t launch_isect_cells(const IcJob<T>& a, const IcJob<T>* b, int use_k, hipStream_t s) {
  for (const IcJob<T>* j : {&a, b}) {
    if (!j || j->nunits <= 0) continue;
    if (j->nuv < 2 || j->nuv > kIcMaxNuv || !j->C || !j->R || !j->lo != !j->hi) return hipErrorInvalidValue;
  }
  const int na = ic_blocks(a), nb = b ? ic_blocks(*b) : 0;
  const int va = ic_variant(a);
  if (b && nb > 0 && ic_variant(*b) == va) {
    if ((long long)na + nb > INT_MAX) return hipErrorInvalidValue;
    ic_launch(va, use_k, a, *b, na, na + nb, s);
  } else {
    ic_launch(va, use_k, a, a, na, na, s);
    if (b) ic_launch(ic_variant(*b), use_k, *b, *b, nb, nb, s);
  }
  return hipGetLastError();
}

template <class T>
static SlotScreen<T> slot_screen(long long nslots, int w1, int w2, const int* su1, const int* su2, const vec4_t<T>* C1,
                                 const vec4_t<T>* R1, const vec4_t<T>* C2, const vec4_t<T>* R2, const vec4_t<T>* lo1,
                                 const vec4_t<T>* hi1, const vec4_t<T>* lo2, const vec4_t<T>* hi2, int cull,
                                 const SlotF32* f) {
  SlotScreen<T> s;
  s.f = f ? *f : SlotF32();
  s.nslots = nslots;
  s.cells1 = w1 * w1;
  s.cells2 = w2 * w2;
  s.su1 = su1;
  s.su2 = su2;
  s.C1 = C1;
  s.R1 = R1;
  s.C2 = C2;
  s.R2 = R2;
  s.lo1 = lo1;
  s.hi1 = hi1;
  s.lo2 = lo2;
  s.hi2 = hi2;
  s.cull = cull;
  return s;
}

static unsigned slot_blocks(long long n, int cells1, int cells2) {
  const int spw = slot_groups(cells1, cells2);
  const long long waves = (n + spw - 1) / spw;
  return (unsigned)std::min<long long>((waves + 3) / 4, 1ll << 20);
}

static bool slot_args_ok(long long nslots, int w1, int w2, int cull, const void* lo1, const void* hi1,
                         const void* lo2, const void* hi2, const SlotF32* f) {
  if (f && (!f->X1 || !f->Y1 || !f->X2 || !f->Y2)) return false;
  if (f) {
    lo1 = f->lo1;
    hi1 = f->hi1;
    lo2 = f->lo2;
    hi2 = f->hi2;
  }
  return w1 >= 1 && w2 >= 1 && w1 <= kIcMaxNuv && w2 <= kIcMaxNuv && nslots <= INT_MAX &&
         !(cull && (!lo1 || !hi1 || !lo2 || !hi2));
}

template <class T>
hipError_t launch_isect_slot_cull(long long nslots, int w1, int w2, const int* su1, const int* su2,
                                  const vec4_t<T>* lo1, const vec4_t<T>* hi1, const vec4_t<T>* lo2,
                                  const vec4_t<T>* hi2, unsigned* wcount, const unsigned long long* woff, int* list,
                                  const SlotF32* f, hipStream_t s) {
  if (nslots <= 0) return hipSuccess;
  if (!slot_args_ok(nslots, w1, w2, 1, lo1, hi1, lo2, hi2, f) || (!wcount == !list)) return hipErrorInvalidValue;
  const SlotScreen<T> sc = slot_screen<T>(nslots, w1, w2, su1, su2, nullptr, nullptr, nullptr, nullptr, lo1, hi1,
                                          lo2, hi2, 1, f);
  const dim3 g((unsigned)((nslots + 255) / 256)), b(256);
  if (wcount)
    hipLaunchKernelGGL(k_isect_slot_cull<T>, g, b, 0, s, sc, wcount);
  else
    hipLaunchKernelGGL(k_isect_slot_list<T>, g, b, 0, s, sc, woff, list);
  return hipGetLastError();
}

template <class T>
hipError_t launch_isect_slot_count(long long nslots, int w1, int w2, const int* su1, const int* su2,
                                   const vec4_t<T>* C1, const vec4_t<T>* R1, const vec4_t<T>* C2,
                                   const vec4_t<T>* R2, const vec4_t<T>* lo1, const vec4_t<T>* hi1,
                                   const vec4_t<T>* lo2, const vec4_t<T>* hi2, int cull, long long n,
                                   const int* list, unsigned* npair, unsigned* nrow, unsigned* ncol, const SlotF32* f,
                                   hipStream_t s, unsigned long long* masks) {
  if (n <= 0) return hipSuccess;
  if (!slot_args_ok(nslots, w1, w2, list ? 0 : cull, lo1, hi1, lo2, hi2, f) || n > nslots)
    return hipErrorInvalidValue;
  const SlotScreen<T> sc = slot_screen<T>(nslots, w1, w2, su1, su2, C1, R1, C2, R2, lo1, hi1, lo2, hi2, cull, f);
  // each wavefront stages its slots' side-2 cells (FP32SCREEN: their float
  // records) in LDS, up to 32 KB per workgroup
  const size_t st_bytes =
      (size_t)4 * slot_groups(sc.cells1, sc.cells2) * sc.cells2 * 2 * sizeof(vec4_t<T>);
  const int stage = st_bytes <= 32768 ? 1 : 0;
  hipLaunchKernelGGL(k_isect_slot_count<T>, dim3(slot_blocks(n, sc.cells1, sc.cells2)), dim3(256),
                     stage ? st_bytes : 0, s, sc, n, list, npair, nrow, ncol, masks, stage);
  return hipGetLastError();
}

template <class T>
hipError_t launch_isect_slot_emit(long long nslots, int w1, int w2, const int* su1, const int* su2,
                                  const vec4_t<T>* C1, const vec4_t<T>* R1, const vec4_t<T>* C2, const vec4_t<T>* R2,
                                  const vec4_t<T>* lo1, const vec4_t<T>* hi1, const vec4_t<T>* lo2,
                                  const vec4_t<T>* hi2, int cull, long long n, const int* list,
                                  const unsigned long long* offp, const unsigned long long* offr,
                                  const unsigned long long* offc, IPair* out, const vec4_t<T>* reg1,
                                  const vec4_t<T>* reg2, int pnuv1, int pnuv2, int* nsu1, int* nsu2,
                                  vec4_t<T>* nreg1, vec4_t<T>* nreg2, const SlotF32* f, hipStream_t s,
                                  const unsigned long long* masks) {
  if (n <= 0) return hipSuccess;
  if (!slot_args_ok(nslots, w1, w2, list ? 0 : cull, lo1, hi1, lo2, hi2, f) || n > nslots)
    return hipErrorInvalidValue;
  const SlotScreen<T> sc = slot_screen<T>(nslots, w1, w2, su1, su2, C1, R1, C2, R2, lo1, hi1, lo2, hi2, cull, f);
  SlotUnits<T> su;
  su.reg1 = reg1;
  su.reg2 = reg2;
  su.pnuv1 = pnuv1;
  su.pnuv2 = pnuv2;
  su.nsu1 = nsu1;
  su.nsu2 = nsu2;
  su.nreg1 = nreg1;
  su.nreg2 = nreg2;
  const int units = nsu1 != nullptr;
  if (units && (!nsu2 || !nreg1 || !nreg2 || !reg1 || !reg2)) return hipErrorInvalidValue;
  const size_t lds = units ? sizeof(int) * 4 * slot_groups(sc.cells1, sc.cells2) * sc.cells2 : 0;
  hipLaunchKernelGGL(k_isect_slot_emit<T>, dim3(slot_blocks(n, sc.cells1, sc.cells2)), dim3(256), lds, s, sc, n, list,
                     offp, offr, offc, out, units, su, masks);
  return hipGetLastError();
}


#define JSI_IS(T)                                                                                                  \
  template hipError_t launch_isect_slot_cull<T>(long long, int, int, const int*, const int*, const vec4_t<T>*,     \
                                                const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, unsigned*,    \
                                                const unsigned long long*, int*, const SlotF32*, hipStream_t);      \
  template hipError_t launch_isect_slot_count<T>(long long, int, int, const int*, const int*, const vec4_t<T>*,    \
                                                 const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*,             \
                                                 const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*,             \
                                                 const vec4_t<T>*, int, long long, const int*, unsigned*, unsigned*, \
                                                 unsigned*, const SlotF32*, hipStream_t, unsigned long long*);      \
  template hipError_t launch_isect_slot_emit<T>(                                                                   \
      long long, int, int, const int*, const int*, const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*,           \
      const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, const vec4_t<T>*, int, long long,    \
      const int*, const unsigned long long*, const unsigned long long*, const unsigned long long*, IPair*,         \
      const vec4_t<T>*, const vec4_t<T>*, int, int, int*, int*, vec4_t<T>*, vec4_t<T>*, const SlotF32*, hipStream_t, \
      const unsigned long long*);
JSI_IS(float)
JSI_IS(double)
#undef JSI_IS

template hipError_t launch_isect_cells<float>(const IcJob<float>&, const IcJob<float>*, int, hipStream_t);
template hipError_t launch_isect_cells<double>(const IcJob<double>&, const IcJob<double>*, int, hipStream_t);

}  // namespace jsi
