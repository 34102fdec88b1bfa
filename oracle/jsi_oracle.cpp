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
#include <array>
#include <deque>
#include <map>
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


// ---------------------------------------------------------------------------
// FP64 solves on the final cells (no counterpart in the reference: its
// minimum_distance returns the k-th centre distance, distance_analysis.cu:505,
// and intersection() only prints pair counts, intersection.cu:367).
//
// Closest points: from the parameter centres of a record's two cells, a
// damped Newton iteration on f(x) = |S1(u1,v1) - S2(u2,v2)|^2, x projected
// onto the query boxes; a coordinate on a bound whose gradient points
// outward is held (gradient entry 0, identity row and column).  With J = [S1u S1v -S2u -S2v] and d = S1 - S2:
// gradient g = J^T d, Hessian H = J^T J + the d-weighted second derivatives
// (+ for surface 1, - for surface 2), step z from (H + mu diag(J^T J + eps)) z
// = g by LDL^T, x' = clamp(x - z); accepted when f decreases (mu /= 10),
// otherwise mu *= 10.  Intersection points: Gauss-Newton with minimum-norm
// steps, x' = clamp(x - J^T (J J^T)^-1 d).  Both in double, a fixed
// operation order, no contraction.
// ---------------------------------------------------------------------------
struct Sol {
  double x[4] = {0, 0, 0, 0};
  double f = 0;          // |S1 - S2|^2
  double p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
  double t[3] = {0, 0, 0};  // intersection: (S1u x S1v) x (S2u x S2v)
  int iters = 0, flags = 0;
};

inline double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double bigger(double a, double b) { return b > a ? b : a; }
inline double clip(double t, double lo, double hi) { return t < lo ? lo : (t > hi ? hi : t); }

// LDL^T solve of a symmetric n x n system given by its lower triangle.
bool ldlt(int n, const double A[4][4], const double* rhs, double* sol) {
  double L[4][4] = {}, W[4][4] = {}, D[4] = {}, y[4] = {};
  for (int j = 0; j < n; ++j) {
    double dj = A[j][j];
    for (int k = 0; k < j; ++k) dj = dj - W[j][k] * L[j][k];
    if (!(dj > 0.0 && dj <= DBL_MAX)) return false;
    D[j] = dj;
    for (int i = j + 1; i < n; ++i) {
      double v = A[i][j];
      for (int k = 0; k < j; ++k) v = v - W[i][k] * L[j][k];
      W[i][j] = v;
      L[i][j] = v / dj;
    }
  }
  for (int i = 0; i < n; ++i) {
    double v = rhs[i];
    for (int k = 0; k < i; ++k) v = v - L[i][k] * y[k];
    y[i] = v;
  }
  for (int i = 0; i < n; ++i) y[i] = y[i] / D[i];
  for (int i = n - 1; i >= 0; --i) {
    double v = y[i];
    for (int k = n - 1; k > i; --k) v = v - L[k][i] * sol[k];
    sol[i] = v;
  }
  return true;
}

// Parameter centre of cell `cell` of a unit (cell iu spans [start + step iu,
// start + step (iu + 1)], step = (stop - start)/nuv, distance_analysis.cu:92-94).
template <class T>
void cell_mid(const Region<T>& r, int cell, double* o) {
  const int w = r.nuv - 1;
  const int iu = cell % w, iv = cell / w;
  const double us = (double(r.ustop) - double(r.ustart)) / double(r.nuv);
  const double vs = (double(r.vstop) - double(r.vstart)) / double(r.nuv);
  o[0] = double(r.ustart) + us * (double(iu) + 0.5);
  o[1] = double(r.vstart) + vs * (double(iv) + 0.5);
}

struct Boxes {
  double lo[4], hi[4];
};
Boxes boxes_of(const jsi_region& a, const jsi_region& b) {
  Boxes bx;
  const jsi_region* rs[2] = {&a, &b};
  for (int s = 0; s < 2; ++s) {
    const double us = rs[s]->ustart, ue = rs[s]->ustop, vs = rs[s]->vstart, ve = rs[s]->vstop;
    bx.lo[2 * s] = std::min(us, ue);
    bx.hi[2 * s] = std::max(us, ue);
    bx.lo[2 * s + 1] = std::min(vs, ve);
    bx.hi[2 * s + 1] = std::max(vs, ve);
  }
  return bx;
}

Sol newton_closest(const Net<double>& n1, const Net<double>& n2, const Boxes& bx, const double* start) {
  Sol r;
  double x[4], y[4];
  for (int q = 0; q < 4; ++q) x[q] = clip(start[q], bx.lo[q], bx.hi[q]);
  double A[18], B[18], An[18], Bn[18];
  eval_point(n1, x[0], x[1], A);
  eval_point(n2, x[2], x[3], B);
  double d[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
  double f = dot(d, d);
  double mu = 1.0e-3;
  int it = 0;
  while (it < 60) {
    ++it;
    const double *su1 = A + 3, *sv1 = A + 6, *su2 = B + 3, *sv2 = B + 6;
    double g[4] = {dot(su1, d), dot(sv1, d), -dot(su2, d), -dot(sv2, d)};
    bool held[4];
    for (int q = 0; q < 4; ++q) {
      held[q] = (x[q] <= bx.lo[q] && g[q] > 0.0) || (x[q] >= bx.hi[q] && g[q] < 0.0);
      if (held[q]) g[q] = 0.0;
    }
    if (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0 && g[3] == 0.0) {
      r.flags |= JSI_POLISH_STATIONARY;
      break;
    }
    const double jj[4] = {dot(su1, su1), dot(sv1, sv1), dot(su2, su2), dot(sv2, sv2)};
    double H[4][4] = {};
    H[0][0] = jj[0] + dot(d, A + 9);
    H[1][0] = dot(su1, sv1) + dot(d, A + 12);
    H[1][1] = jj[1] + dot(d, A + 15);
    H[2][0] = -dot(su1, su2);
    H[3][0] = -dot(su1, sv2);
    H[2][1] = -dot(sv1, su2);
    H[3][1] = -dot(sv1, sv2);
    H[2][2] = jj[2] - dot(d, B + 9);
    H[3][2] = dot(su2, sv2) - dot(d, B + 12);
    H[3][3] = jj[3] - dot(d, B + 15);
    const double eps = 1.0e-12 * bigger(bigger(jj[0], jj[1]), bigger(jj[2], jj[3]));
    for (int q = 0; q < 4; ++q) H[q][q] = H[q][q] + mu * (jj[q] + eps);
    for (int q = 0; q < 4; ++q) {
      if (!held[q]) continue;
      for (int r = 0; r < 4; ++r)
        if (r < q)
          H[q][r] = 0.0;
        else if (r > q)
          H[r][q] = 0.0;
      H[q][q] = 1.0;
    }
    double z[4];
    if (!ldlt(4, H, g, z)) {
      mu = mu * 10.0;
      if (mu > 1.0e12) {
        r.flags |= JSI_POLISH_STALLED;
        break;
      }
      continue;
    }
    double zbig = 0.0, xbig = 0.0;
    for (int q = 0; q < 4; ++q) {
      zbig = bigger(zbig, std::fabs(z[q]));
      xbig = bigger(xbig, std::fabs(x[q]));
    }
    if (zbig <= 1.0e-15 * (1.0 + xbig)) {  // step below the parameters' rounding
      r.flags |= JSI_POLISH_CONVERGED;
      break;
    }
    bool moved = false;
    for (int q = 0; q < 4; ++q) {
      y[q] = clip(x[q] - z[q], bx.lo[q], bx.hi[q]);
      moved = moved || y[q] != x[q];
    }
    if (!moved) {
      r.flags |= JSI_POLISH_FIXED;
      break;
    }
    eval_point(n1, y[0], y[1], An);
    eval_point(n2, y[2], y[3], Bn);
    const double dn[3] = {An[0] - Bn[0], An[1] - Bn[1], An[2] - Bn[2]};
    const double fn = dot(dn, dn);
    if (!(fn < f)) {
      if (zbig <= 1.0e-9 * (1.0 + xbig)) {  // no decrease from a step at f's resolution
        r.flags |= JSI_POLISH_CONVERGED;
        break;
      }
      mu = mu * 10.0;
      if (mu > 1.0e12) {
        r.flags |= JSI_POLISH_STALLED;
        break;
      }
      continue;
    }
    double step = 0.0, mag = 0.0;
    for (int q = 0; q < 4; ++q) {
      step = bigger(step, std::fabs(y[q] - x[q]));
      mag = bigger(mag, std::fabs(y[q]));
      x[q] = y[q];
    }
    std::memcpy(A, An, sizeof(A));
    std::memcpy(B, Bn, sizeof(B));
    for (int c = 0; c < 3; ++c) d[c] = dn[c];
    f = fn;
    mu = mu * 0.1;
    if (mu < 1.0e-15) mu = 1.0e-15;
    if (step <= 1.0e-15 * (1.0 + mag)) {
      r.flags |= JSI_POLISH_CONVERGED;
      break;
    }
  }
  for (int q = 0; q < 4; ++q) r.x[q] = x[q];
  r.f = f;
  for (int c = 0; c < 3; ++c) {
    r.p1[c] = A[c];
    r.p2[c] = B[c];
  }
  r.iters = it;
  return r;
}

Sol newton_meet(const Net<double>& n1, const Net<double>& n2, const Boxes& bx, const double* start) {
  Sol r;
  double x[4];
  for (int q = 0; q < 4; ++q) x[q] = clip(start[q], bx.lo[q], bx.hi[q]);
  double A[18], B[18];
  eval_point(n1, x[0], x[1], A);
  eval_point(n2, x[2], x[3], B);
  double d[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
  double f = dot(d, d);
  const double scale = 1.0 + bigger(bigger(std::fabs(A[0]), std::fabs(A[1])), std::fabs(A[2]));
  const double tl = 1.0e-14 * scale;
  const double tol2 = tl * tl;
  int it = 0;
  while (true) {
    if (f <= tol2) {
      r.flags |= JSI_POLISH_CONVERGED;
      break;
    }
    if (it == 40) break;
    ++it;
    const double *su1 = A + 3, *sv1 = A + 6, *su2 = B + 3, *sv2 = B + 6;
    double G[4][4] = {};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j <= i; ++j) G[i][j] = su1[i] * su1[j] + sv1[i] * sv1[j] + su2[i] * su2[j] + sv2[i] * sv2[j];
    double lam[3];
    if (!ldlt(3, G, d, lam)) {
      r.flags |= JSI_POLISH_STALLED;
      break;
    }
    const double step[4] = {-dot(su1, lam), -dot(sv1, lam), dot(su2, lam), dot(sv2, lam)};
    double y[4];
    bool moved = false;
    for (int q = 0; q < 4; ++q) {
      y[q] = clip(x[q] + step[q], bx.lo[q], bx.hi[q]);
      moved = moved || y[q] != x[q];
    }
    if (!moved) {
      r.flags |= JSI_POLISH_FIXED;
      break;
    }
    std::memcpy(x, y, sizeof(x));
    eval_point(n1, x[0], x[1], A);
    eval_point(n2, x[2], x[3], B);
    for (int c = 0; c < 3; ++c) d[c] = A[c] - B[c];
    f = dot(d, d);
  }
  for (int q = 0; q < 4; ++q) r.x[q] = x[q];
  r.f = f;
  for (int c = 0; c < 3; ++c) {
    r.p1[c] = A[c];
    r.p2[c] = B[c];
  }
  const double n1v[3] = {A[4] * A[8] - A[5] * A[7], A[5] * A[6] - A[3] * A[8], A[3] * A[7] - A[4] * A[6]};
  const double n2v[3] = {B[4] * B[8] - B[5] * B[7], B[5] * B[6] - B[3] * B[8], B[3] * B[7] - B[4] * B[6]};
  r.t[0] = n1v[1] * n2v[2] - n1v[2] * n2v[1];
  r.t[1] = n1v[2] * n2v[0] - n1v[0] * n2v[2];
  r.t[2] = n1v[0] * n2v[1] - n1v[1] * n2v[0];
  r.iters = it;
  return r;
}

struct Cell3 {
  long long a, b, c;
  bool operator<(const Cell3& o) const { return a != o.a ? a < o.a : (b != o.b ? b < o.b : c < o.c); }
};
Cell3 cell3(const double* p, double s) {
  return Cell3{(long long)std::floor(p[0] / s), (long long)std::floor(p[1] / s), (long long)std::floor(p[2] / s)};
}
double sqdist(const double* p, const double* q) {
  const double e[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
  return dot(e, e);
}

// Tracing of the intersection curves from the projected points.
//  * seeds: projections with residual^2 <= tol^2, in pair order, thinned to
//    spacing h (dropped when a kept seed is closer than h; h = spacing, else
//    1e-3 of the accepted points' bounding-box diagonal, 1 if that is 0);
//    seeds with a zero tangent t = (S1u x S1v) x (S2u x S2v) are left out;
//    T = t / |t|;
//  * march from a seed along +-T with step h: predictor = per-surface
//    least-squares parameter step for a move of h T; corrector = Gauss-Newton
//    (minimum norm) on S1 - S2 = 0 and T.(S1 - p_prev) = h from the predicted
//    point clamped into the query boxes; when a corrector iterate leaves the
//    boxes, it is clamped, the coordinate furthest outside (relative to its
//    box width) is fixed, S1 - S2 = 0 is solved over the other three and the
//    march ends at that edge point; after an interior step the tangent is
//    re-oriented along the previous one; a new point closer than h/2 to a
//    seed (lowest index; the start seed from the second step on) ends the
//    march at that seed;
//  * curves: every seed that no march reaches heads an open curve (backward
//    march reversed, then the chain of seeds and forward points); leftover
//    seeds form closed curves when their chain returns to its start.
struct Seed {
  double x[4], p[3], T[3], r2;
};
struct Step {
  double x[4], p[3], r2;
};
struct Trace {
  int succ = -1;
  int flags = 0;
  std::vector<Step> pts;
};

enum { kConv, kFail, kLeft, kNoConv };

int corrector(const Net<double>& n1, const Net<double>& n2, const Boxes& bx, double* x, double* A, double* B, int freem,
              bool plane, const double* T, const double* p0, double h, double tol2) {
  for (int it = 0;; ++it) {
    double F[4] = {A[0] - B[0], A[1] - B[1], A[2] - B[2], 0.0};
    double r2 = dot(F, F);
    if (plane) {
      const double q[3] = {A[0] - p0[0], A[1] - p0[1], A[2] - p0[2]};
      F[3] = dot(T, q) - h;
      r2 = r2 + F[3] * F[3];
    }
    if (r2 <= tol2) return kConv;
    if (it == 20) return kNoConv;
    double J[4][4];
    for (int r = 0; r < 3; ++r) {
      J[r][0] = A[3 + r];
      J[r][1] = A[6 + r];
      J[r][2] = -B[3 + r];
      J[r][3] = -B[6 + r];
    }
    J[3][0] = dot(T, A + 3);
    J[3][1] = dot(T, A + 6);
    J[3][2] = 0.0;
    J[3][3] = 0.0;
    for (int q = 0; q < 4; ++q)
      if (!(freem & (1 << q)))
        for (int r = 0; r < 4; ++r) J[r][q] = 0.0;
    const int m = plane ? 4 : 3;
    double G[4][4] = {};
    for (int i = 0; i < m; ++i)
      for (int j = 0; j <= i; ++j) G[i][j] = J[i][0] * J[j][0] + J[i][1] * J[j][1] + J[i][2] * J[j][2] + J[i][3] * J[j][3];
    double lam[4];
    if (!ldlt(m, G, F, lam)) return kFail;
    bool outside = false;
    for (int q = 0; q < 4; ++q) {
      double corr = J[0][q] * lam[0] + J[1][q] * lam[1] + J[2][q] * lam[2];
      if (plane) corr = corr + J[3][q] * lam[3];
      const double v = x[q] - corr;
      const bool out = v < bx.lo[q] || v > bx.hi[q];
      outside = outside || (plane && out);
      x[q] = (out && !plane) ? clip(v, bx.lo[q], bx.hi[q]) : v;
    }
    if (outside) return kLeft;
    eval_point(n1, x[0], x[1], A);
    eval_point(n2, x[2], x[3], B);
  }
}

// Clamps into the boxes; returns the bit of the coordinate that was furthest
// outside relative to its box width (first on ties), 0 if none was outside.
int pin_to_box(double* x, const Boxes& bx) {
  int pick = -1;
  double far = 0.0;
  for (int q = 0; q < 4; ++q) {
    const double width = bx.hi[q] - bx.lo[q];
    const double over = x[q] < bx.lo[q] ? (bx.lo[q] - x[q]) / width : (x[q] > bx.hi[q] ? (x[q] - bx.hi[q]) / width : 0.0);
    if (over > far) {
      far = over;
      pick = q;
    }
    x[q] = clip(x[q], bx.lo[q], bx.hi[q]);
  }
  return pick < 0 ? 0 : 1 << pick;
}

Trace march(const Net<double>& n1, const Net<double>& n2, const Boxes& bx, const std::vector<Seed>& seeds, int from,
            double dir, double h) {
  Trace tr;
  const Seed& sd = seeds[from];
  double x[4], T[3], p[3], A[18], B[18];
  std::memcpy(x, sd.x, sizeof(x));
  for (int c = 0; c < 3; ++c) {
    T[c] = dir * sd.T[c];
    p[c] = sd.p[c];
  }
  eval_point(n1, x[0], x[1], A);
  eval_point(n2, x[2], x[3], B);
  const double scale = 1.0 + bigger(bigger(std::fabs(p[0]), std::fabs(p[1])), std::fabs(p[2]));
  const double tl = 1.0e-14 * scale;
  const double tol2 = tl * tl;
  const double half = 0.5 * h;
  const double half2 = half * half;
  auto record = [&]() {
    Step st;
    std::memcpy(st.x, x, sizeof(x));
    for (int c = 0; c < 3; ++c) st.p[c] = A[c];
    const double d[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    st.r2 = dot(d, d);
    tr.pts.push_back(st);
  };
  int steps = 0;
  for (; steps < 8192; ++steps) {
    const double mv[3] = {h * T[0], h * T[1], h * T[2]};
    double G1[4][4] = {}, G2[4][4] = {};
    G1[0][0] = dot(A + 3, A + 3);
    G1[1][0] = dot(A + 3, A + 6);
    G1[1][1] = dot(A + 6, A + 6);
    G2[0][0] = dot(B + 3, B + 3);
    G2[1][0] = dot(B + 3, B + 6);
    G2[1][1] = dot(B + 6, B + 6);
    const double b1[2] = {dot(A + 3, mv), dot(A + 6, mv)};
    const double b2[2] = {dot(B + 3, mv), dot(B + 6, mv)};
    double d1[2], d2[2];
    if (!ldlt(2, G1, b1, d1) || !ldlt(2, G2, b2, d2)) {
      tr.flags |= 4;
      break;
    }
    x[0] = x[0] + d1[0];
    x[1] = x[1] + d1[1];
    x[2] = x[2] + d2[0];
    x[3] = x[3] + d2[1];
    pin_to_box(x, bx);
    eval_point(n1, x[0], x[1], A);
    eval_point(n2, x[2], x[3], B);
    const int st = corrector(n1, n2, bx, x, A, B, 15, true, T, p, h, tol2);
    int pinned = 0;
    if (st == kLeft) {
      pinned = pin_to_box(x, bx);
      eval_point(n1, x[0], x[1], A);
      eval_point(n2, x[2], x[3], B);
      if (corrector(n1, n2, bx, x, A, B, 15 & ~pinned, false, T, p, h, tol2) != kConv) {
        tr.flags |= 4;
        break;
      }
    } else if (st != kConv) {
      tr.flags |= 4;
      break;
    }
    if (!pinned) {
      const double c1[3] = {A[4] * A[8] - A[5] * A[7], A[5] * A[6] - A[3] * A[8], A[3] * A[7] - A[4] * A[6]};
      const double c2[3] = {B[4] * B[8] - B[5] * B[7], B[5] * B[6] - B[3] * B[8], B[3] * B[7] - B[4] * B[6]};
      const double t[3] = {c1[1] * c2[2] - c1[2] * c2[1], c1[2] * c2[0] - c1[0] * c2[2],
                           c1[0] * c2[1] - c1[1] * c2[0]};
      const double tt = dot(t, t);
      if (!(tt > 0.0)) {
        tr.flags |= 4;
        break;
      }
      const double inv = 1.0 / std::sqrt(tt);
      double u[3] = {t[0] * inv, t[1] * inv, t[2] * inv};
      if (dot(u, T) < 0.0)
        for (double& v : u) v = -v;
      for (int c = 0; c < 3; ++c) T[c] = u[c];
    }
    for (int c = 0; c < 3; ++c) p[c] = A[c];
    int hit = -1;
    for (int q = 0; q < (int)seeds.size() && hit < 0; ++q)
      if (!(q == from && steps == 0) && sqdist(p, seeds[q].p) < half2) hit = q;
    if (hit >= 0) {
      tr.succ = hit;
      tr.flags |= 1;
      break;
    }
    record();
    if (pinned) {
      tr.flags |= 2;
      break;
    }
  }
  if (steps == 8192) tr.flags |= 8;
  return tr;
}

void trace_curves(const Net<double>& n1, const Net<double>& n2, const Boxes& bx, const std::vector<Sol>& sols,
                  double tol, double spacing, std::vector<jsi_curve_point>& out, int* ncurves) {
  out.clear();
  *ncurves = 0;
  std::vector<const Sol*> acc;
  for (const Sol& s : sols)
    if (s.f <= tol * tol) acc.push_back(&s);
  if (acc.empty()) return;
  double h = spacing;
  if (!(h > 0.0)) {
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) lo[c] = hi[c] = acc[0]->p1[c];
    for (const Sol* s : acc)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], s->p1[c]);
        hi[c] = std::max(hi[c], s->p1[c]);
      }
    const double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    h = 1.0e-3 * std::sqrt(dot(ext, ext));
    if (!(h > 0.0)) h = 1.0;
  }
  std::vector<Seed> seeds;
  {
    std::map<Cell3, std::vector<int>> grid;
    for (const Sol* s : acc) {
      const Cell3 k = cell3(s->p1, h);
      bool close = false;
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int c = -1; c <= 1; ++c) {
            auto it = grid.find(Cell3{k.a + a, k.b + b, k.c + c});
            if (it == grid.end()) continue;
            for (int q : it->second) close = close || sqdist(s->p1, seeds[q].p) < h * h;
          }
      const double tt = dot(s->t, s->t);
      if (close || !(tt > 0.0)) continue;
      Seed sd;
      std::memcpy(sd.x, s->x, sizeof(sd.x));
      std::memcpy(sd.p, s->p1, sizeof(sd.p));
      const double inv = 1.0 / std::sqrt(tt);
      for (int c = 0; c < 3; ++c) sd.T[c] = s->t[c] * inv;
      sd.r2 = s->f;
      grid[k].push_back((int)seeds.size());
      seeds.push_back(sd);
    }
  }
  const int ns = (int)seeds.size();
  std::vector<Trace> fwd(ns);
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < ns; ++s) fwd[s] = march(n1, n2, bx, seeds, s, 1.0, h);
  std::vector<int> into(ns, 0);
  for (const Trace& t : fwd)
    if (t.succ >= 0) ++into[t.succ];
  std::vector<int> heads;
  for (int s = 0; s < ns; ++s)
    if (into[s] == 0) heads.push_back(s);
  std::vector<Trace> bwd(heads.size());
#pragma omp parallel for schedule(dynamic)
  for (long long e = 0; e < (long long)heads.size(); ++e) bwd[e] = march(n1, n2, bx, seeds, heads[e], -1.0, h);
  std::vector<char> done(ns, 0);
  int curve = 0;
  auto put = [&](const double* x, const double* p, double r2, int flags) {
    out.push_back(jsi_curve_point{p[0], p[1], p[2], x[0], x[1], x[2], x[3], std::sqrt(r2), curve, flags});
  };
  auto follow = [&](int s, int flags) {
    for (; s >= 0 && !done[s]; s = fwd[s].succ) {
      done[s] = 1;
      put(seeds[s].x, seeds[s].p, seeds[s].r2, flags);
      for (const Step& st : fwd[s].pts) put(st.x, st.p, st.r2, flags);
    }
  };
  for (size_t e = 0; e < heads.size(); ++e) {
    if (done[heads[e]]) continue;
    for (auto it = bwd[e].pts.rbegin(); it != bwd[e].pts.rend(); ++it) put(it->x, it->p, it->r2, 0);
    follow(heads[e], 0);
    ++curve;
  }
  for (int s = 0; s < ns; ++s) {
    if (done[s]) continue;
    // closed when following the successors from s comes back to s before
    // meeting a seed already on a curve
    std::vector<char> mark(ns, 0);
    int t = s;
    while (t >= 0 && !done[t] && !mark[t]) {
      mark[t] = 1;
      t = fwd[t].succ;
    }
    follow(s, t == s ? JSI_CURVE_CLOSED : 0);
    ++curve;
  }
  *ncurves = curve;
}

// Rounds of a distance query: dist_max_rounds, or with param_tol > 0 the
// first round whose cells span <= param_tol in u and v on both surfaces.
int rounds_for(const jsi_params& P, const jsi_region& r1, const jsi_region& r2) {
  if (!(P.param_tol > 0.0)) return P.dist_max_rounds;
  double span = 0.0;
  for (const jsi_region* r : {&r1, &r2}) {
    span = std::max(span, std::fabs(double(r->ustop) - double(r->ustart)) / double(r->nuv));
    span = std::max(span, std::fabs(double(r->vstop) - double(r->vstart)) / double(r->nuv));
  }
  int n = 1;
  for (; span > P.param_tol && n < JSI_MAX_ROUNDS; ++n) span = span / double(P.dist_refine_nuv);
  return n;
}

struct CurveOut {
  double spacing = 0.0, tol = 1.0e-9;
  std::vector<jsi_curve_point> pts;
  int ncurves = 0;
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
  const int max_rounds = rounds_for(P, r1, r2);
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
    if (pairs.empty() || round == max_rounds || round >= JSI_MAX_ROUNDS) break;
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
  if (P.polish > 0 && !pairs.empty()) {
    const Net<double> m1 = make_net<double>(iv1, d1), m2 = make_net<double>(iv2, d2);
    const Boxes bx = boxes_of(r1, r2);
    const int m = std::min<int>(P.polish, (int)pairs.size());
    std::vector<Sol> sols(m);
#pragma omp parallel for schedule(dynamic)
    for (int e = 0; e < m; ++e) {
      double x0[4];
      cell_mid(slot1[pairs[e].slot]->reg, pairs[e].i, x0);
      cell_mid(slot2[pairs[e].slot]->reg, pairs[e].j, x0 + 2);
      sols[e] = newton_closest(m1, m2, bx, x0);
    }
    int win = 0;
    for (int e = 1; e < m; ++e)
      if (sols[e].f < sols[win].f) win = e;
    jsi_closest& c = res->closest;
    c.distance = std::sqrt(sols[win].f);
    c.u1 = sols[win].x[0];
    c.v1 = sols[win].x[1];
    c.u2 = sols[win].x[2];
    c.v2 = sols[win].x[3];
    for (int q = 0; q < 3; ++q) {
      c.p1[q] = sols[win].p1[q];
      c.p2[q] = sols[win].p2[q];
    }
    c.candidate = win;
    c.iterations = sols[win].iters;
    c.polished = m;
    c.flags = sols[win].flags;
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
              const float* d2, const jsi_region& r2, jsi_result* res, jsi_pair* out, int cap,
              CurveOut* co = nullptr) {
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
  if (co) {
    const Net<double> m1 = make_net<double>(iv1, d1), m2 = make_net<double>(iv2, d2);
    const Boxes bx = boxes_of(r1, r2);
    std::vector<Sol> sols(pairs.size());
#pragma omp parallel for schedule(dynamic, 256)
    for (long long e = 0; e < (long long)pairs.size(); ++e) {
      double x0[4];
      cell_mid(slot1[pairs[e].slot]->reg, pairs[e].i, x0);
      cell_mid(slot2[pairs[e].slot]->reg, pairs[e].j, x0 + 2);
      sols[e] = newton_meet(m1, m2, bx, x0);
    }
    trace_curves(m1, m2, bx, sols, co->tol, co->spacing, co->pts, &co->ncurves);
  }
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
      P->dist_refine_nuv < 2 || P->polish < 0 || !(P->param_tol >= 0.0) ||
      (P->param_tol > 0.0 && P->precision == JSI_PREC_EXACT32))
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

int orc_intersect_curves(const jsi_params* P, const int* iv1, const float* d1, const jsi_region* r1,
                         const int* iv2, const float* d2, const jsi_region* r2, double spacing, double tol,
                         jsi_result* res, jsi_curve_point* pts, int cap, int* npts, int* ncurves) {
  if (!valid_query(P, iv1, d1, r1, iv2, d2, r2) || !res || !npts || !ncurves || P->isect_max_rounds < 1 ||
      P->isect_refine_nuv < 2 || tol < 0.0)
    return JSI_ERR_INVALID;
  std::memset(res, 0, sizeof(*res));
  CurveOut co;
  co.spacing = spacing;
  co.tol = tol > 0.0 ? tol : 1.0e-9;
  int rc;
  if (P->precision == JSI_PREC_EXACT32)
    rc = intersect<float>(*P, iv1, d1, *r1, iv2, d2, *r2, res, nullptr, 0, &co);
  else
    rc = intersect<double>(*P, iv1, d1, *r1, iv2, d2, *r2, res, nullptr, 0, &co);
  if (rc != JSI_OK) return rc;
  *npts = (int)co.pts.size();
  *ncurves = co.ncurves;
  if (pts) std::memcpy(pts, co.pts.data(), sizeof(jsi_curve_point) * std::min<size_t>(co.pts.size(), size_t(cap)));
  return JSI_OK;
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
