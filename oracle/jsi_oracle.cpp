// jsi_oracle.cpp - CPU oracle (test infrastructure, never part of the product).
//
// A restatement of the JSI_CAD_NEXT VERSION5 hot path, templated on the scalar
// type.  With T = float and -ffp-contract=off it reproduces the reference bit
// for bit (x86-64 g++ builds of the reference do not fuse multiply-adds).
// With T = double it is the FP64 oracle: same operation order, double
// arithmetic, surface data promoted from float.
//
// File:line citations refer to JSI_CAD_NEXT_VERSION5/.
#include "jsi_oracle.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <queue>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kMaxDeg = 25;  // MAXDEG, evaluation.cu:13

// ---------------------------------------------------------------------------
// Surface view (BuildEvalContext, evaluation.cu:239-257)
// ---------------------------------------------------------------------------
template <class T>
struct Net {
  int flags = 0, degu = 0, degv = 0, ncpu = 0, ncpv = 0, nku = 0, nkv = 0;
  bool rational = false;
  bool zero = false;  // periodic flags without the flat bit, or degree >= 25: all-zero output (:271-272)
  std::vector<T> U, V, CP, W;
};

template <class T>
Net<T> make_net(const int* ivec, const float* data) {
  Net<T> n;
  n.flags = ivec[0];
  n.degu = ivec[1];
  n.ncpu = ivec[2];
  n.nku = ivec[3];
  n.degv = ivec[4];
  n.ncpv = ivec[5];
  n.nkv = ivec[6];
  n.rational = (n.flags & 2) != 0;
  n.zero = ((n.flags & 1) == 0 && (n.flags & 12) != 0) || n.degu >= kMaxDeg || n.degv >= kMaxDeg;
  const float* p = data;
  n.U.assign(p, p + n.nku);
  p += n.nku;
  n.V.assign(p, p + n.nkv);
  p += n.nkv;
  const int ncp = n.ncpu * n.ncpv;
  n.CP.assign(p, p + 3 * ncp);
  p += 3 * ncp;
  if (n.rational) n.W.assign(p, p + ncp);
  return n;
}

// FindSpan, evaluation.cu:152-166.  The span is the unique m in [p, n-1] with
// U[m] <= u < U[m+1], clamped at both ends.
template <class T>
int find_span(int nk, int p, T u, const T* U) {
  if (u <= U[p]) return p;
  const int n = nk - p - 1;
  if (u >= U[n]) return n - 1;
  int lo = p, hi = n, mid = (lo + hi) / 2;
  while (u < U[mid] || u >= U[mid + 1]) {
    if (u < U[mid])
      hi = mid;
    else
      lo = mid;
    mid = (lo + hi) / 2;
  }
  return mid;
}

// Basis functions and derivatives, The NURBS Book algorithm A2.3 as used by
// DersBasisFuns, evaluation.cu:168-219.  ders is [der+1][kMaxDeg+1].
template <class T>
void basis_ders(int span, int p, T u, const T* U, int der, T ders[][kMaxDeg + 1]) {
  T ndu[kMaxDeg + 1][kMaxDeg + 1];
  T left[kMaxDeg + 1], right[kMaxDeg + 1];
  T a[2][kMaxDeg + 1];
  ndu[0][0] = T(1);
  for (int j = 1; j <= p; ++j) {
    left[j] = u - U[span + 1 - j];
    right[j] = U[span + j] - u;
    T saved = T(0);
    for (int r = 0; r < j; ++r) {
      ndu[j][r] = right[r + 1] + left[j - r];         // lower triangle: knot differences
      const T temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;        // upper triangle: basis values
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
  for (int j = 0; j <= p; ++j) ders[0][j] = ndu[j][p];
  for (int r = 0; r <= p; ++r) {
    int s1 = 0, s2 = 1;
    a[0][0] = T(1);
    for (int k = 1; k <= der; ++k) {
      T d = T(0);
      const int rk = r - k, pk = p - k;
      if (r >= k) {
        a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
        d = a[s2][0] * ndu[rk][pk];
      }
      const int j1 = (rk >= -1) ? 1 : -rk;
      const int j2 = (r - 1 <= pk) ? k - 1 : p - r;
      for (int j = j1; j <= j2; ++j) {
        a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
        d += a[s2][j] * ndu[rk + j][pk];
      }
      if (r <= pk) {
        a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
        d += a[s2][k] * ndu[r][pk];
      }
      ders[k][r] = d;
      std::swap(s1, s2);
    }
  }
  int r = p;
  for (int k = 1; k <= der; ++k) {
    for (int j = 0; j <= p; ++j) ders[k][j] *= T(r);
    r *= p - k;
  }
}

// Rational quotient rule for dim=3, der=2, stride 4 (EG_EvaluateQuotientRule2,
// evaluation.cu:77-128).  v holds homogeneous [S, Su, Sv, Suu, Suv, Svv] x
// (x, y, z, w).  One reciprocal of W scales all 24 entries, then first-order
// and second-order corrections.
template <class T>
void quotient_rule(T v[24]) {
  const T F = T(1) / v[3];
  for (int e = 0; e < 24; ++e) v[e] *= F;
  const T ws = -v[4 + 3];
  const T wt = -v[8 + 3];
  for (int c = 0; c < 3; ++c) {
    v[4 + c] += ws * v[c];
    v[8 + c] += wt * v[c];
  }
  const T wss = -v[12 + 3];
  const T wst = -v[16 + 3];
  const T wtt = -v[20 + 3];
  for (int c = 0; c < 3; ++c) {
    const T S = v[c], Su = v[4 + c], Sv = v[8 + c];
    v[12 + c] += wss * S + T(2) * ws * Su;
    v[16 + c] += wst * S + wt * Su + ws * Sv;
    v[20 + c] += wtt * S + T(2) * wt * Sv;
  }
}

// Output row of the (k = u-order, l = v-order) derivative before the 2<->3
// row swap (loop order of EvaluateOneUV, evaluation.cu:281-300): l outer, k inner.
inline int raw_row(int k, int l) { return l == 0 ? k : (l == 1 ? 3 + k : 5); }
// Final order [S, Su, Sv, Suu, Suv, Svv] after the swap (:301-310, :336-347).
inline int final_row(int k, int l) {
  const int m = raw_row(k, l);
  return m == 2 ? 3 : (m == 3 ? 2 : m);
}

// EvaluateOneUV, evaluation.cu:259-356.
template <class T>
void eval_point(const Net<T>& n, T u, T v, T out[18]) {
  for (int e = 0; e < 18; ++e) out[e] = T(0);
  if (n.zero) return;
  const int du = std::min(2, n.degu), dv = std::min(2, n.degv);
  T Nu[3][kMaxDeg + 1], Nv[3][kMaxDeg + 1];
  const int su = find_span(n.nku, n.degu, u, n.U.data());
  basis_ders(su, n.degu, u, n.U.data(), du, Nu);
  const int sv = find_span(n.nkv, n.degv, v, n.V.data());
  basis_ders(sv, n.degv, v, n.V.data(), dv, Nv);
  const int p = n.degu, q = n.degv;
  if (!n.rational) {
    for (int l = 0; l <= dv; ++l) {
      for (int k = 0; k <= 2 - l; ++k) {
        if (k > du) continue;
        T acc[3] = {T(0), T(0), T(0)};
        for (int s = 0; s <= q; ++s) {
          T row[3] = {T(0), T(0), T(0)};
          const int base = n.ncpu * (sv - q + s) + su - p;
          for (int j = 0; j <= p; ++j) {
            const T* P = &n.CP[3 * (base + j)];
            row[0] += Nu[k][j] * P[0];
            row[1] += Nu[k][j] * P[1];
            row[2] += Nu[k][j] * P[2];
          }
          acc[0] += Nv[l][s] * row[0];
          acc[1] += Nv[l][s] * row[1];
          acc[2] += Nv[l][s] * row[2];
        }
        const int m = final_row(k, l);
        out[3 * m + 0] = acc[0];
        out[3 * m + 1] = acc[1];
        out[3 * m + 2] = acc[2];
      }
    }
    return;
  }
  T h[24];
  for (int e = 0; e < 24; ++e) h[e] = T(0);
  for (int l = 0; l <= dv; ++l) {
    for (int k = 0; k <= 2 - l; ++k) {
      if (k > du) continue;
      T acc[4] = {T(0), T(0), T(0), T(0)};
      for (int s = 0; s <= q; ++s) {
        T row[4] = {T(0), T(0), T(0), T(0)};
        const int base = n.ncpu * (sv - q + s) + su - p;
        for (int j = 0; j <= p; ++j) {
          const int id = base + j;
          const T nw = Nu[k][j] * n.W[id];
          row[0] += nw * n.CP[3 * id + 0];
          row[1] += nw * n.CP[3 * id + 1];
          row[2] += nw * n.CP[3 * id + 2];
          row[3] += nw;
        }
        for (int c = 0; c < 4; ++c) acc[c] += Nv[l][s] * row[c];
      }
      const int m = final_row(k, l);
      for (int c = 0; c < 4; ++c) h[4 * m + c] = acc[c];
    }
  }
  quotient_rule(h);
  for (int m = 0; m < 6; ++m)
    for (int c = 0; c < 3; ++c) out[3 * m + c] = h[4 * m + c];
}

// Grid driver, evaluation.cu:358-369.
template <class T>
void eval_grid(const Net<T>& n, T ustep, T vstep, T ustart, T vstart, int nuv, T* out) {
#pragma omp parallel for schedule(static) collapse(2) if (nuv * nuv > 512)
  for (int idy = 0; idy < nuv; ++idy)
    for (int idx = 0; idx < nuv; ++idx)
      eval_point(n, ustep * T(idx) + ustart, vstep * T(idy) + vstart, out + size_t(idx + idy * nuv) * 18);
}

// std::min / std::max semantics (first argument wins unless the second is
// strictly smaller / larger): the nesting order fixes the sign of zero.
template <class T>
inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T>
inline T smax(T a, T b) { return (a < b) ? b : a; }

// Corner reduction of one cell (distance_analysis.cu:56-73, intersection.cu:115-132).
template <class T>
inline void cell_minmax(int nuv, const T* pts, int iv, int iu, int c, T& lo, T& hi) {
  const T a = pts[size_t(iv * nuv + iu) * 18 + c];
  const T b = pts[size_t(iv * nuv + iu + 1) * 18 + c];
  const T d = pts[size_t((iv + 1) * nuv + iu) * 18 + c];
  const T e = pts[size_t((iv + 1) * nuv + iu + 1) * 18 + c];
  lo = smin(a, smin(b, smin(d, e)));
  hi = smax(a, smax(b, smax(d, e)));
}

// Distance AABB centres, construct_aabbs_distance_host (distance_analysis.cu:45-86), k = 0.
template <class T>
void aabb_dist(int nuv, const T* pts, T* center) {
  const int w = nuv - 1;
  const T k = T(0);
  for (int iv = 0; iv < w; ++iv)
    for (int iu = 0; iu < w; ++iu)
      for (int c = 0; c < 3; ++c) {
        T lo, hi;
        cell_minmax(nuv, pts, iv, iu, c, lo, hi);
        center[size_t(iv * w + iu) * 3 + c] = (lo - k + hi + k) / T(2);
      }
}

// Intersection AABBs, construct_aabbs_host (intersection.cu:106-148).
template <class T>
void aabb_isect(int nuv, const T* pts, T pad, T* center, T* radius) {
  const int w = nuv - 1;
  for (int iv = 0; iv < w; ++iv)
    for (int iu = 0; iu < w; ++iu)
      for (int c = 0; c < 3; ++c) {
        T lo, hi;
        cell_minmax(nuv, pts, iv, iu, c, lo, hi);
        const size_t o = size_t(iv * w + iu) * 3 + c;
        center[o] = (lo + hi) / T(2);
        radius[o] = (hi - lo) / T(2) + pad;
      }
}

// compute_K, common.cu:149-172 (CadMax(a, b) = a > b ? a : b).
template <class T>
inline T cmax(T a, T b) { return a > b ? a : b; }

template <class T>
T compute_k(const T* pts, int nuv) {
  T muu = -std::numeric_limits<T>::max(), muv = muu, mvv = muu;
  const int npt = nuv * nuv;
  for (int i = 0; i < npt; ++i) {
    const T* r = pts + size_t(i) * 18;
    muu = cmax(muu, cmax(std::fabs(r[9]), cmax(std::fabs(r[10]), std::fabs(r[11]))));
    muv = cmax(muv, cmax(std::fabs(r[12]), cmax(std::fabs(r[13]), std::fabs(r[14]))));
    mvv = cmax(mvv, cmax(std::fabs(r[15]), cmax(std::fabs(r[16]), std::fabs(r[17]))));
  }
  const int side = nuv > 1 ? nuv - 1 : 1;
  const T s2 = T(side * side);
  return (muu / s2 + T(2) * muv / s2 + mvv / s2) / T(8);
}

// ---------------------------------------------------------------------------
// Query state
// ---------------------------------------------------------------------------
template <class T>
struct Region {
  T ustart, vstart, ustop, vstop;
  int nuv;
};

template <class T>
struct Unit {  // one evaluated region: AABBResult (distance_analysis.cuh:224-237)
  Region<T> reg;
  int w = 0;                // naabb_width = nuv - 1
  std::vector<T> center;    // 3 per cell
  std::vector<T> radius;    // 3 per cell (intersection only)
};

// Child region of cell bbox of a parent unit: make_dist_refine_task /
// make_intersect_refine_task (distance_analysis.cu:340-354, intersection.cu:311-325).
template <class T>
Region<T> child_region(const Unit<T>& parent, int bbox, int nuv) {
  const T ustep = (parent.reg.ustop - parent.reg.ustart) / T(parent.reg.nuv);
  const T vstep = (parent.reg.vstop - parent.reg.vstart) / T(parent.reg.nuv);
  const int x = bbox % parent.w, y = bbox / parent.w;
  Region<T> r;
  r.nuv = nuv;
  r.ustart = parent.reg.ustart + ustep * T(x);
  r.ustop = r.ustart + ustep;
  r.vstart = parent.reg.vstart + vstep * T(y);
  r.vstop = r.vstart + vstep;
  return r;
}

// run_evaluation_and_construction_task, distance flavour (distance_analysis.cu:88-116).
template <class T>
Unit<T> eval_unit_dist(const Net<T>& n, const Region<T>& r) {
  Unit<T> u;
  u.reg = r;
  u.w = r.nuv - 1;
  const T ustep = (r.ustop - r.ustart) / T(r.nuv);
  const T vstep = (r.vstop - r.vstart) / T(r.nuv);
  std::vector<T> pts(size_t(r.nuv) * r.nuv * 18);
  eval_grid(n, ustep, vstep, r.ustart, r.vstart, r.nuv, pts.data());
  u.center.resize(size_t(u.w) * u.w * 3);
  aabb_dist(r.nuv, pts.data(), u.center.data());
  return u;
}

// run_evaluation_and_construction_task, intersection flavour (intersection.cu:175-219).
template <class T>
Unit<T> eval_unit_isect(const Net<T>& n, const Region<T>& r, int side, bool use_k, const T init_u[3],
                        const T init_v[3]) {
  Unit<T> u;
  u.reg = r;
  u.w = r.nuv - 1;
  const T ustep = (r.ustop - r.ustart) / T(r.nuv);
  const T vstep = (r.vstop - r.vstart) / T(r.nuv);
  std::vector<T> pts(size_t(r.nuv) * r.nuv * 18);
  eval_grid(n, ustep, vstep, r.ustart, r.vstart, r.nuv, pts.data());
  T pad = T(0);
  if (use_k) {
    const T raw = compute_k(pts.data(), r.nuv);
    const int sd = (side == 2) ? 2 : 1;
    const T us = std::fabs(r.ustop - r.ustart) / init_u[sd];
    const T vs = std::fabs(r.vstop - r.vstart) / init_v[sd];
    const T ps = smax(us, vs);
    pad = raw * ps * ps;
    if (!std::isfinite(pad) || pad < T(0)) pad = T(0);
  }
  u.center.resize(size_t(u.w) * u.w * 3);
  u.radius.resize(size_t(u.w) * u.w * 3);
  aabb_isect(r.nuv, pts.data(), pad, u.center.data(), u.radius.data());
  return u;
}

// ---------------------------------------------------------------------------
// Distance screen: bounded top-k under the total order (d2, slot, i, j)
// (dist_pair_ref_less / dist_pair_top_push / dist_collect_top_pairs,
// distance_analysis.cu:202-338).  The k smallest elements of a totally
// ordered set are unique, so per-thread heaps merged at the end give exactly
// the list the reference's insertion-sorted vector holds.
// ---------------------------------------------------------------------------
template <class T>
struct DPair {
  T d2;
  int slot, i, j;
};
template <class T>
inline bool dless(const DPair<T>& a, const DPair<T>& b) {
  if (a.d2 != b.d2) return a.d2 < b.d2;
  if (a.slot != b.slot) return a.slot < b.slot;
  if (a.i != b.i) return a.i < b.i;
  return a.j < b.j;
}
template <class T>
struct DLessCmp {
  bool operator()(const DPair<T>& a, const DPair<T>& b) const { return dless(a, b); }
};

// center_distance_sq, distance_analysis.cu:202-207.
template <class T>
inline T center_d2(const T* a, const T* b) {
  const T dx = a[0] - b[0];
  const T dy = a[1] - b[1];
  const T dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}

template <class T>
std::vector<DPair<T>> topk_slots(const std::vector<const Unit<T>*>& s1, const std::vector<const Unit<T>*>& s2,
                                 int k) {
  const int limit = std::max(1, k);
  struct Row {
    int slot, i;
  };
  std::vector<Row> rows;
  for (size_t s = 0; s < s1.size(); ++s) {
    const int n1 = s1[s]->w * s1[s]->w;
    for (int i = 0; i < n1; ++i) rows.push_back({int(s), i});
  }
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  std::vector<std::vector<DPair<T>>> heaps(nthreads);
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    std::priority_queue<DPair<T>, std::vector<DPair<T>>, DLessCmp<T>> heap;
    T bound = std::numeric_limits<T>::infinity();
#pragma omp for schedule(static)
    for (long long r = 0; r < (long long)rows.size(); ++r) {
      const int slot = rows[r].slot, i = rows[r].i;
      const Unit<T>& a = *s1[slot];
      const Unit<T>& b = *s2[slot];
      const int n2 = b.w * b.w;
      const T* ci = &a.center[size_t(i) * 3];
      const T* cb = b.center.data();
      for (int j = 0; j < n2; ++j) {
        const T d2 = center_d2(ci, cb + size_t(j) * 3);
        if (d2 > bound) continue;  // cannot enter: the heap holds limit elements all <= bound
        const DPair<T> p{d2, slot, i, j};
        if ((int)heap.size() < limit) {
          heap.push(p);
          if ((int)heap.size() == limit) bound = heap.top().d2;
        } else if (dless(p, heap.top())) {
          heap.pop();
          heap.push(p);
          bound = heap.top().d2;
        }
      }
    }
    auto& out = heaps[tid];
    while (!heap.empty()) {
      out.push_back(heap.top());
      heap.pop();
    }
  }
  std::vector<DPair<T>> all;
  for (auto& h : heaps) all.insert(all.end(), h.begin(), h.end());
  std::sort(all.begin(), all.end(), dless<T>);
  if ((int)all.size() > limit) all.resize(limit);
  return all;
}

// ---------------------------------------------------------------------------
// Overlap screen (AabbOverlap / intersect_collect_pairs, intersection.cu:46-63,259-309):
// pairs in (slot, i, j) order.
// ---------------------------------------------------------------------------
template <class T>
inline bool overlap(const T* c1, const T* r1, const T* c2, const T* r2) {
  return std::fabs(c1[0] - c2[0]) < (r1[0] + r2[0]) && std::fabs(c1[1] - c2[1]) < (r1[1] + r2[1]) &&
         std::fabs(c1[2] - c2[2]) < (r1[2] + r2[2]);
}

struct IPair {
  int slot, i, j;
};

template <class T>
std::vector<IPair> overlap_slots(const std::vector<const Unit<T>*>& s1, const std::vector<const Unit<T>*>& s2) {
  struct Row {
    int slot, i;
  };
  std::vector<Row> rows;
  for (size_t s = 0; s < s1.size(); ++s) {
    const int n1 = s1[s]->w * s1[s]->w;
    for (int i = 0; i < n1; ++i) rows.push_back({int(s), i});
  }
  const long long nrows = (long long)rows.size();
  std::vector<long long> cnt(nrows + 1, 0);
  auto row_scan = [&](long long r, IPair* dst) -> long long {
    const int slot = rows[r].slot, i = rows[r].i;
    const Unit<T>& a = *s1[slot];
    const Unit<T>& b = *s2[slot];
    const int n2 = b.w * b.w;
    const T* c1 = &a.center[size_t(i) * 3];
    const T* r1 = &a.radius[size_t(i) * 3];
    long long c = 0;
    for (int j = 0; j < n2; ++j) {
      if (overlap(c1, r1, &b.center[size_t(j) * 3], &b.radius[size_t(j) * 3])) {
        if (dst) dst[c] = IPair{slot, i, j};
        ++c;
      }
    }
    return c;
  };
#pragma omp parallel for schedule(dynamic, 64)
  for (long long r = 0; r < nrows; ++r) cnt[r + 1] = row_scan(r, nullptr);
  for (long long r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
  std::vector<IPair> out(static_cast<size_t>(cnt[nrows]));
#pragma omp parallel for schedule(dynamic, 64)
  for (long long r = 0; r < nrows; ++r)
    if (cnt[r + 1] > cnt[r]) row_scan(r, out.data() + cnt[r]);
  return out;
}

// ---------------------------------------------------------------------------
// Trace hashing (FNV-1a 64 over the ordered pair list)
// ---------------------------------------------------------------------------
inline void fnv(uint64_t& h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
}
constexpr uint64_t kFnvSeed = 1469598103934665603ull;

// First-appearance dedup of (slot, cell) endpoints (find_or_add_unit /
// find_or_add_task, distance_analysis.cu:434-465, intersection.cu:421-447).
struct Dedup {
  std::unordered_map<uint64_t, int> index;
  std::vector<std::pair<int, int>> endpoints;  // (slot, bbox) in first-appearance order
  int add(int slot, int bbox) {
    const uint64_t key = (uint64_t(uint32_t(slot)) << 32) | uint32_t(bbox);
    auto it = index.emplace(key, int(endpoints.size()));
    if (it.second) endpoints.push_back({slot, bbox});
    return it.first->second;
  }
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
Region<T> to_region(const jsi_region& r) {
  return Region<T>{T(r.ustart), T(r.vstart), T(r.ustop), T(r.vstop), r.nuv};
}

// minimum_distance, distance_analysis.cu:356-506 (JSI_CAD_PRESERVE_PAIR_RELATION=1).
template <class T>
int min_distance(const jsi_params& P, const int* iv1, const float* d1, const jsi_region& r1, const int* iv2,
                 const float* d2, const jsi_region& r2, jsi_result* res, jsi_pair* out, int cap) {
  const double t0 = now_ms();
  const Net<T> n1 = make_net<T>(iv1, d1), n2 = make_net<T>(iv2, d2);
  std::vector<Unit<T>> units1, units2;  // all units ever evaluated (the reference's workspace)
  units1.reserve(1);
  std::vector<const Unit<T>*> slot1, slot2;
  double te = now_ms();
  units1.push_back(eval_unit_dist(n1, to_region<T>(r1)));
  units2.push_back(eval_unit_dist(n2, to_region<T>(r2)));
  res->ms_eval += now_ms() - te;
  slot1.push_back(&units1.back());
  slot2.push_back(&units2.back());
  std::deque<std::vector<Unit<T>>> keep;  // owns refinement units of every round (stable references)
  int nuv_ref = P.dist_refine_nuv;
  int round = 0;
  std::vector<DPair<T>> pairs;
  int cur_nuv = r1.nuv;
  while (true) {
    ++round;
    const double ts = now_ms();
    pairs = topk_slots(slot1, slot2, P.k);
    res->ms_screen += now_ms() - ts;
    jsi_round_info& ri = res->round[round - 1];
    ri.npairs = (int64_t)pairs.size();
    ri.slots = (int64_t)slot1.size();
    ri.nuv = cur_nuv;
    ri.kth_d2 = pairs.empty() ? 0.0 : double(pairs.back().d2);
    if (P.trace) {
      uint64_t h = kFnvSeed;
      for (const auto& p : pairs) {
        const double dd = double(p.d2);
        fnv(h, &dd, 8);
        fnv(h, &p.slot, 4);
        fnv(h, &p.i, 4);
        fnv(h, &p.j, 4);
      }
      ri.hash = h;
    }
    res->distance = pairs.empty() ? 0.0 : double(std::sqrt(pairs.back().d2));
    res->rounds = round;
    res->npairs = (int64_t)pairs.size();
    if (pairs.empty() || round == P.dist_max_rounds || round >= JSI_MAX_ROUNDS) break;
    Dedup u1, u2;
    std::vector<std::pair<int, int>> jobs;
    for (const auto& p : pairs) jobs.push_back({u1.add(p.slot, p.i), u2.add(p.slot, p.j)});
    ri.units1 = (int64_t)u1.endpoints.size();
    ri.units2 = (int64_t)u2.endpoints.size();
    te = now_ms();
    keep.emplace_back(u1.endpoints.size());
    auto& k1 = keep.back();
#pragma omp parallel for schedule(dynamic)
    for (long long e = 0; e < (long long)u1.endpoints.size(); ++e)
      k1[e] = eval_unit_dist(n1, child_region(*slot1[u1.endpoints[e].first], u1.endpoints[e].second, nuv_ref));
    keep.emplace_back(u2.endpoints.size());
    auto& k2 = keep.back();
#pragma omp parallel for schedule(dynamic)
    for (long long e = 0; e < (long long)u2.endpoints.size(); ++e)
      k2[e] = eval_unit_dist(n2, child_region(*slot2[u2.endpoints[e].first], u2.endpoints[e].second, nuv_ref));
    res->ms_eval += now_ms() - te;
    std::vector<const Unit<T>*> ns1, ns2;
    for (const auto& jb : jobs) {
      ns1.push_back(&k1[jb.first]);
      ns2.push_back(&k2[jb.second]);
    }
    slot1.swap(ns1);
    slot2.swap(ns2);
    cur_nuv = nuv_ref;
  }
  if (out) {
    const int n = std::min<int>(cap, (int)pairs.size());
    for (int e = 0; e < n; ++e) out[e] = jsi_pair{double(pairs[e].d2), pairs[e].slot, pairs[e].i, pairs[e].j, 0};
  }
  res->ms_total = now_ms() - t0;
  return JSI_OK;
}

// intersection, intersection.cu:338-475 (JSI_CAD_PRESERVE_PAIR_RELATION=1).
template <class T>
int intersect(const jsi_params& P, const int* iv1, const float* d1, const jsi_region& r1, const int* iv2,
              const float* d2, const jsi_region& r2, jsi_result* res, jsi_pair* out, int cap) {
  const double t0 = now_ms();
  const Net<T> n1 = make_net<T>(iv1, d1), n2 = make_net<T>(iv2, d2);
  const Region<T> g1 = to_region<T>(r1), g2 = to_region<T>(r2);
  const T tiny = T(1.0e-12f);
  T init_u[3] = {T(1), T(1), T(1)}, init_v[3] = {T(1), T(1), T(1)};
  init_u[1] = smax(std::fabs(g1.ustop - g1.ustart), tiny);
  init_v[1] = smax(std::fabs(g1.vstop - g1.vstart), tiny);
  init_u[2] = smax(std::fabs(g2.ustop - g2.ustart), tiny);
  init_v[2] = smax(std::fabs(g2.vstop - g2.vstart), tiny);
  const bool use_k = P.use_compute_k != 0;
  std::deque<std::vector<Unit<T>>> keep(2);
  double te = now_ms();
  keep[0].push_back(eval_unit_isect(n1, g1, 1, use_k, init_u, init_v));
  keep[1].push_back(eval_unit_isect(n2, g2, 2, use_k, init_u, init_v));
  res->ms_eval += now_ms() - te;
  std::vector<const Unit<T>*> slot1{&keep[0][0]}, slot2{&keep[1][0]};
  int g_nuv = P.isect_refine_nuv;
  int cur_nuv = r1.nuv;
  int round = 0;
  std::vector<IPair> pairs;
  while (true) {
    ++round;
    const double ts = now_ms();
    pairs = overlap_slots(slot1, slot2);
    res->ms_screen += now_ms() - ts;
    jsi_round_info& ri = res->round[round - 1];
    ri.npairs = (int64_t)pairs.size();
    ri.slots = (int64_t)slot1.size();
    ri.nuv = cur_nuv;
    if (P.trace) {
      uint64_t h = kFnvSeed;
      for (const auto& p : pairs) {
        fnv(h, &p.slot, 4);
        fnv(h, &p.i, 4);
        fnv(h, &p.j, 4);
      }
      ri.hash = h;
    }
    res->rounds = round;
    res->npairs = (int64_t)pairs.size();
    if (round == P.isect_max_rounds || round >= JSI_MAX_ROUNDS) break;
    if (pairs.empty()) break;
    if ((long long)pairs.size() > P.isect_halve_above && g_nuv > P.isect_min_nuv) g_nuv /= 2;
    Dedup u1, u2;
    std::vector<std::pair<int, int>> jobs;
    jobs.reserve(pairs.size());
    for (const auto& p : pairs) jobs.push_back({u1.add(p.slot, p.i), u2.add(p.slot, p.j)});
    ri.units1 = (int64_t)u1.endpoints.size();
    ri.units2 = (int64_t)u2.endpoints.size();
    te = now_ms();
    std::vector<Unit<T>> k1(u1.endpoints.size()), k2(u2.endpoints.size());
#pragma omp parallel for schedule(dynamic)
    for (long long e = 0; e < (long long)u1.endpoints.size(); ++e)
      k1[e] = eval_unit_isect(n1, child_region(*slot1[u1.endpoints[e].first], u1.endpoints[e].second, g_nuv), 1,
                              use_k, init_u, init_v);
#pragma omp parallel for schedule(dynamic)
    for (long long e = 0; e < (long long)u2.endpoints.size(); ++e)
      k2[e] = eval_unit_isect(n2, child_region(*slot2[u2.endpoints[e].first], u2.endpoints[e].second, g_nuv), 2,
                              use_k, init_u, init_v);
    res->ms_eval += now_ms() - te;
    keep.push_back(std::move(k1));
    auto& kk1 = keep.back();
    keep.push_back(std::move(k2));
    auto& kk2 = keep.back();
    std::vector<const Unit<T>*> ns1, ns2;
    ns1.reserve(jobs.size());
    ns2.reserve(jobs.size());
    for (const auto& jb : jobs) {
      ns1.push_back(&kk1[jb.first]);
      ns2.push_back(&kk2[jb.second]);
    }
    slot1.swap(ns1);
    slot2.swap(ns2);
    cur_nuv = g_nuv;
  }
  res->isect_refine_nuv = g_nuv;
  if (out) {
    const long long n = std::min<long long>(cap, (long long)pairs.size());
    for (long long e = 0; e < n; ++e) out[e] = jsi_pair{0.0, pairs[e].slot, pairs[e].i, pairs[e].j, 0};
  }
  res->ms_total = now_ms() - t0;
  return JSI_OK;
}

bool valid_surface(const int* ivec) {
  if (!ivec) return false;
  return ivec[1] >= 1 && ivec[4] >= 1 && ivec[2] > ivec[1] && ivec[5] > ivec[4] && ivec[3] == ivec[2] + ivec[1] + 1 &&
         ivec[6] == ivec[5] + ivec[4] + 1;
}

bool valid_query(const jsi_params* P, const int* iv1, const float* d1, const jsi_region* r1, const int* iv2,
                 const float* d2, const jsi_region* r2) {
  if (!P || !iv1 || !d1 || !r1 || !iv2 || !d2 || !r2) return false;
  if (!valid_surface(iv1) || !valid_surface(iv2)) return false;
  if (r1->nuv < 2 || r2->nuv < 2 || r1->nuv > JSI_MAX_NUV || r2->nuv > JSI_MAX_NUV) return false;
  return true;
}

}  // namespace

extern "C" {

int orc_eval_grid_f32(const int* ivec, const float* data, float ustep, float vstep, float ustart, float vstart,
                      int nuv, float* out) {
  if (!ivec || !data || !out || nuv < 1) return JSI_ERR_INVALID;
  const Net<float> n = make_net<float>(ivec, data);
  eval_grid(n, ustep, vstep, ustart, vstart, nuv, out);
  return JSI_OK;
}

int orc_eval_grid_f64(const int* ivec, const float* data, double ustep, double vstep, double ustart,
                      double vstart, int nuv, double* out) {
  if (!ivec || !data || !out || nuv < 1) return JSI_ERR_INVALID;
  const Net<double> n = make_net<double>(ivec, data);
  eval_grid(n, ustep, vstep, ustart, vstart, nuv, out);
  return JSI_OK;
}

int orc_basis_f32(int nknots, int degree, const float* knots, float u, int der, float* ders) {
  if (degree < 1 || degree >= kMaxDeg || der < 0 || der > 2) return JSI_ERR_INVALID;
  float d[3][kMaxDeg + 1];
  const int span = find_span(nknots, degree, u, knots);
  basis_ders(span, degree, u, knots, der, d);
  for (int k = 0; k <= der; ++k)
    for (int j = 0; j <= degree; ++j) ders[k * (degree + 1) + j] = d[k][j];
  return span;
}

int orc_basis_f64(int nknots, int degree, const double* knots, double u, int der, double* ders) {
  if (degree < 1 || degree >= kMaxDeg || der < 0 || der > 2) return JSI_ERR_INVALID;
  double d[3][kMaxDeg + 1];
  const int span = find_span(nknots, degree, u, knots);
  basis_ders(span, degree, u, knots, der, d);
  for (int k = 0; k <= der; ++k)
    for (int j = 0; j <= degree; ++j) ders[k * (degree + 1) + j] = d[k][j];
  return span;
}

int orc_aabb_dist_f32(int nuv, const float* pts18, float* center) {
  if (nuv < 2 || !pts18 || !center) return JSI_ERR_INVALID;
  aabb_dist(nuv, pts18, center);
  return JSI_OK;
}
int orc_aabb_dist_f64(int nuv, const double* pts18, double* center) {
  if (nuv < 2 || !pts18 || !center) return JSI_ERR_INVALID;
  aabb_dist(nuv, pts18, center);
  return JSI_OK;
}
int orc_aabb_isect_f32(int nuv, const float* pts18, float pad, float* center, float* radius) {
  if (nuv < 2 || !pts18 || !center || !radius) return JSI_ERR_INVALID;
  aabb_isect(nuv, pts18, pad, center, radius);
  return JSI_OK;
}
int orc_aabb_isect_f64(int nuv, const double* pts18, double pad, double* center, double* radius) {
  if (nuv < 2 || !pts18 || !center || !radius) return JSI_ERR_INVALID;
  aabb_isect(nuv, pts18, pad, center, radius);
  return JSI_OK;
}
float orc_compute_k_f32(const float* pts18, int nuv) { return compute_k(pts18, nuv); }
double orc_compute_k_f64(const double* pts18, int nuv) { return compute_k(pts18, nuv); }

int orc_min_distance(const jsi_params* P, const int* iv1, const float* d1, const jsi_region* r1, const int* iv2,
                     const float* d2, const jsi_region* r2, jsi_result* res, jsi_pair* out, int cap) {
  if (!valid_query(P, iv1, d1, r1, iv2, d2, r2) || !res || P->k < 1 || P->dist_max_rounds < 1 ||
      P->dist_refine_nuv < 2)
    return JSI_ERR_INVALID;
  std::memset(res, 0, sizeof(*res));
  if (P->precision == JSI_PREC_EXACT32) return min_distance<float>(*P, iv1, d1, *r1, iv2, d2, *r2, res, out, cap);
  return min_distance<double>(*P, iv1, d1, *r1, iv2, d2, *r2, res, out, cap);
}

int orc_intersect(const jsi_params* P, const int* iv1, const float* d1, const jsi_region* r1, const int* iv2,
                  const float* d2, const jsi_region* r2, jsi_result* res, jsi_pair* out, int cap) {
  if (!valid_query(P, iv1, d1, r1, iv2, d2, r2) || !res || P->isect_max_rounds < 1 || P->isect_refine_nuv < 2)
    return JSI_ERR_INVALID;
  std::memset(res, 0, sizeof(*res));
  if (P->precision == JSI_PREC_EXACT32) return intersect<float>(*P, iv1, d1, *r1, iv2, d2, *r2, res, out, cap);
  return intersect<double>(*P, iv1, d1, *r1, iv2, d2, *r2, res, out, cap);
}

int orc_topk_f32(int n1, const float* c1, int n2, const float* c2, int k, jsi_pair* out) {
  if (n1 < 0 || n2 < 0 || k < 1 || !out) return JSI_ERR_INVALID;
  // Wrap the flat arrays as single-slot units.  w is only used for n = w*w,
  // so pass square counts through a fake width when possible.
  Unit<float> a, b;
  a.center.assign(c1, c1 + size_t(n1) * 3);
  b.center.assign(c2, c2 + size_t(n2) * 3);
  a.w = int(std::lround(std::sqrt(double(n1))));
  b.w = int(std::lround(std::sqrt(double(n2))));
  if (a.w * a.w != n1 || b.w * b.w != n2) return JSI_ERR_INVALID;
  std::vector<const Unit<float>*> s1{&a}, s2{&b};
  auto v = topk_slots(s1, s2, k);
  for (size_t e = 0; e < v.size(); ++e) out[e] = jsi_pair{double(v[e].d2), v[e].slot, v[e].i, v[e].j, 0};
  return int(v.size());
}

// ---------------------------------------------------------------------------
// Operator ABI of the reference's hotspot library (V0 mask semantics):
// jsi_ascend_distance_topk_mask / jsi_ascend_intersection_mask,
// JSI_CAD_NEXT_VERSION5/src/ascend/hotspots_910b/src/jsi_ascend_hotspots.cpp:33-113.
// ---------------------------------------------------------------------------
int orc_distance_topk_mask(const float* c1, size_t n1, const float* c2, size_t n2, int topk, int* m1, int* m2,
                           float* kth) {
  if (!c1 || !c2 || !m1 || !m2 || !kth || n1 == 0 || n2 == 0 || topk <= 0) return -1;
  const size_t total = n1 * n2;
  std::vector<float> dist(total);
#pragma omp parallel for schedule(static)
  for (long long x = 0; x < (long long)total; ++x) {
    const size_t i = size_t(x) / n2, j = size_t(x) % n2;
    const float dx = std::fabs(c1[i * 3] - c2[j * 3]);
    const float dy = std::fabs(c1[i * 3 + 1] - c2[j * 3 + 1]);
    const float dz = std::fabs(c1[i * 3 + 2] - c2[j * 3 + 2]);
    dist[x] = std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  const size_t k = std::min<size_t>(size_t(topk), total) - 1;
  std::vector<float> tmp = dist;
  std::nth_element(tmp.begin(), tmp.begin() + (long long)k, tmp.end());
  *kth = tmp[k];
  for (size_t i = 0; i < n1; ++i) m1[i] = 0;
  for (size_t j = 0; j < n2; ++j) m2[j] = 0;
  for (size_t i = 0; i < n1; ++i)
    for (size_t j = 0; j < n2; ++j)
      if (dist[i * n2 + j] <= *kth) m1[i] = m2[j] = 1;
  return 0;
}

int orc_intersection_mask(const float* c1, const float* r1, int n1, const float* c2, const float* r2, int n2, int* m1,
                          int* m2) {
  if (!c1 || !r1 || !c2 || !r2 || !m1 || !m2 || n1 <= 0 || n2 <= 0) return -1;
  for (int i = 0; i < n1; ++i) m1[i] = 0;
  for (int j = 0; j < n2; ++j) m2[j] = 0;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n2; ++j)
      if (overlap(c1 + 3 * i, r1 + 3 * i, c2 + 3 * j, r2 + 3 * j)) m1[i] = m2[j] = 1;
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
