// jsi_refine.hip - FP64 solves on the final cells of a query, beyond what the
// reference computes (its minimum_distance returns the k-th centre distance,
// distance_analysis.cu:505, and its intersection() only prints pair counts,
// intersection.cu:367):
//
//  * k_polish: closest points.  Each of the first m records of the final
//    top-k list starts at the parameter centres of its two cells and runs a
//    damped projected Newton iteration on f = |S1(u1,v1) - S2(u2,v2)|^2 in the
//    query regions (coordinates on a bound whose gradient points outward are
//    held fixed).
//  * k_curve_proj: intersection points.  Each final overlap pair starts at its
//    two cell centres and runs Gauss-Newton with minimum-norm steps on
//    S1(u1,v1) - S2(u2,v2) = 0.
//  * k_march: tracing of the intersection curves from the projected points
//    (seeds) by predictor-corrector steps of a fixed length; each seed marches
//    (one wavefront per march) until the next seed along the curve, a domain
//    edge or a failure, and the
//    host chains seeds by these successor links into ordered curves
//    (curve_seeds, assemble_curves).
//
// Surface points use EvaluateOneUV's operation order in double
// (evaluation.cu:259-356), as the grid kernels do.  Every step below has a
// fixed operation order, and the library is built without contraction, so a
// CPU restatement of the same steps reproduces the results exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/jsi_mi355x.h"
#include "jsi_internal.hpp"
#include "jsi_nurbs.hpp"

namespace jsi {

namespace {

constexpr int kMaxDegP = 24;     // highest degree the library evaluates
constexpr int kPolishIters = 60;  // Newton iterations (accepted and rejected steps)
constexpr int kCurveIters = 40;   // Gauss-Newton iterations
constexpr double kMuStart = 1.0e-3, kMuMin = 1.0e-15, kMuMax = 1.0e12;

__host__ __device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__host__ __device__ inline double dmax(double a, double b) { return b > a ? b : a; }
__host__ __device__ inline double clampd(double t, double lo, double hi) { return t < lo ? lo : (t > hi ? hi : t); }
__host__ __device__ inline void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// S, Su, Sv, Suu, Suv, Svv at (u, v): EvaluateOneUV (evaluation.cu:259-356)
// and the quotient rule (:77-128) in double, any degree <= MAXP.
template <int MAXP>
__device__ void eval_generic(const SurfDev<double>& sd, double u, double v, double out[18]) {
#pragma unroll
  for (int e = 0; e < 18; ++e) out[e] = 0.0;
  if (sd.zero) return;
  const int p = sd.degu, q = sd.degv;
  const int du = p < 2 ? p : 2, dv = q < 2 ? q : 2;
  double Nu[3][MAXP + 1], Nv[3][MAXP + 1];
  const int su = find_span<double>(sd.nku, p, u, sd.U);
  basis_ders<double, MAXP>(su, p, u, sd.U, du, Nu);
  const int sv = find_span<double>(sd.nkv, q, v, sd.V);
  basis_ders<double, MAXP>(sv, q, v, sd.V, dv, Nv);
  double h[6][4];
  for (int m = 0; m < 6; ++m)
    for (int c = 0; c < 4; ++c) h[m][c] = 0.0;
  // loop order of EvaluateOneUV: l outer, k inner; rows 2 and 3 swapped on output
  for (int l = 0; l <= dv; ++l) {
    for (int k = 0; k <= 2 - l; ++k) {
      if (k > du) continue;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (int s = 0; s <= q; ++s) {
        double row[4] = {0.0, 0.0, 0.0, 0.0};
        const int base = sd.ncpu * (sv - q + s) + su - p;
        for (int j = 0; j <= p; ++j) {
          const int id = base + j;
          if (sd.rational) {
            const double nw = Nu[k][j] * sd.W[id];
            row[0] += nw * sd.CP[3 * id + 0];
            row[1] += nw * sd.CP[3 * id + 1];
            row[2] += nw * sd.CP[3 * id + 2];
            row[3] += nw;
          } else {
            row[0] += Nu[k][j] * sd.CP[3 * id + 0];
            row[1] += Nu[k][j] * sd.CP[3 * id + 1];
            row[2] += Nu[k][j] * sd.CP[3 * id + 2];
          }
        }
        for (int c = 0; c < 4; ++c) acc[c] += Nv[l][s] * row[c];
      }
      const int raw = (l == 0) ? k : (l == 1 ? 3 + k : 5);
      const int m = raw == 2 ? 3 : (raw == 3 ? 2 : raw);
      for (int c = 0; c < 4; ++c) h[m][c] = acc[c];
    }
  }
  if (sd.rational) {
    const double F = 1.0 / h[0][3];
    for (int m = 0; m < 6; ++m)
      for (int c = 0; c < 4; ++c) h[m][c] *= F;
    const double ws = -h[1][3], wt = -h[2][3];
    for (int c = 0; c < 3; ++c) {
      h[1][c] += ws * h[0][c];
      h[2][c] += wt * h[0][c];
    }
    const double wss = -h[3][3], wst = -h[4][3], wtt = -h[5][3];
    for (int c = 0; c < 3; ++c) {
      const double S = h[0][c], Su = h[1][c], Sv = h[2][c];
      h[3][c] += wss * S + 2.0 * ws * Su;
      h[4][c] += wst * S + wt * Su + ws * Sv;
      h[5][c] += wtt * S + 2.0 * wt * Sv;
    }
  }
  for (int m = 0; m < 6; ++m)
    for (int c = 0; c < 3; ++c) out[3 * m + c] = h[m][c];
}

// A2.3 for a compile-time degree P (derivatives up to min(ORD, P)): the
// operations of basis_ders with every loop unrolled, so the tables stay in
// registers; results are bit-identical (the rows of order <= ORD do not
// depend on the higher ones).
template <int P, int ORD>
__device__ __forceinline__ void basis_fixed(int span, double u, const double* U, double (&ders)[3][P + 1]) {
  constexpr int D = P < ORD ? P : ORD;
  double ndu[P + 1][P + 1], left[P + 1], right[P + 1], a[2][P + 1];
  ndu[0][0] = 1.0;
#pragma unroll
  for (int j = 1; j <= P; ++j) {
    left[j] = u - U[span + 1 - j];
    right[j] = U[span + j] - u;
    double saved = 0.0;
#pragma unroll
    for (int r = 0; r < j; ++r) {
      ndu[j][r] = right[r + 1] + left[j - r];
      const double temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
#pragma unroll
  for (int j = 0; j <= P; ++j) ders[0][j] = ndu[j][P];
#pragma unroll
  for (int r = 0; r <= P; ++r) {
    a[0][0] = 1.0;
#pragma unroll
    for (int k = 1; k <= D; ++k) {
      const int s1 = (k - 1) & 1, s2 = k & 1;  // the row swap of A2.3 after each k
      double d = 0.0;
      const int rk = r - k, pk = P - k;
      if (r >= k) {
        a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
        d = a[s2][0] * ndu[rk][pk];
      }
      const int j1 = (rk >= -1) ? 1 : -rk;
      const int j2 = (r - 1 <= pk) ? k - 1 : P - r;
#pragma unroll
      for (int j = 1; j <= P; ++j) {
        if (j < j1 || j > j2) continue;
        a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
        d += a[s2][j] * ndu[rk + j][pk];
      }
      if (r <= pk) {
        a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
        d += a[s2][k] * ndu[r][pk];
      }
      ders[k][r] = d;
    }
  }
  int f = P;
#pragma unroll
  for (int k = 1; k <= D; ++k) {
#pragma unroll
    for (int j = 0; j <= P; ++j) ders[k][j] *= (double)f;
    f *= P - k;
  }
}

// eval_generic for compile-time degrees (PU, PV).  Each control row is read
// once: its u-sums for the derivative orders k (the same sums EvaluateOneUV
// forms for every l) feed the accumulators of all (k, l), in the same order.
template <int PU, int PV, int ORD>
__device__ __forceinline__ void eval_fixed(const SurfDev<double>& sd, double u, double v, double out[18]) {
  constexpr int DU = PU < ORD ? PU : ORD, DV = PV < ORD ? PV : ORD;
  constexpr int NR = ORD == 1 ? 3 : 6;  // output rows: S, Su, Sv (, Suu, Suv, Svv)
  double Nu[3][PU + 1], Nv[3][PV + 1];
  const int su = find_span<double>(sd.nku, PU, u, sd.U);
  basis_fixed<PU, ORD>(su, u, sd.U, Nu);
  const int sv = find_span<double>(sd.nkv, PV, v, sd.V);
  basis_fixed<PV, ORD>(sv, v, sd.V, Nv);
  double acc[NR][4];
#pragma unroll
  for (int m = 0; m < NR; ++m)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[m][c] = 0.0;
  const bool rat = sd.rational != 0;
#pragma unroll
  for (int s = 0; s <= PV; ++s) {
    const int base = sd.ncpu * (sv - PV + s) + su - PU;
    double P[PU + 1][4];
#pragma unroll
    for (int j = 0; j <= PU; ++j) {
      P[j][0] = sd.CP[3 * (base + j) + 0];
      P[j][1] = sd.CP[3 * (base + j) + 1];
      P[j][2] = sd.CP[3 * (base + j) + 2];
      P[j][3] = rat ? sd.W[base + j] : 0.0;
    }
    double row[DU + 1][4];
#pragma unroll
    for (int k = 0; k <= DU; ++k) {
#pragma unroll
      for (int c = 0; c < 4; ++c) row[k][c] = 0.0;
#pragma unroll
      for (int j = 0; j <= PU; ++j) {
        if (rat) {
          const double nw = Nu[k][j] * P[j][3];
          row[k][0] += nw * P[j][0];
          row[k][1] += nw * P[j][1];
          row[k][2] += nw * P[j][2];
          row[k][3] += nw;
        } else {
          row[k][0] += Nu[k][j] * P[j][0];
          row[k][1] += Nu[k][j] * P[j][1];
          row[k][2] += Nu[k][j] * P[j][2];
        }
      }
    }
#pragma unroll
    for (int l = 0; l <= DV; ++l)
#pragma unroll
      for (int k = 0; k <= ORD - l; ++k) {
        if (k > DU) continue;
        const int raw = (l == 0) ? k : (l == 1 ? 3 + k : 5);
        const int m = raw == 2 ? 3 : (raw == 3 ? 2 : raw);
        if (m >= NR) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[m][c] += Nv[l][s] * row[k][c];
      }
  }
  if (rat) {
    const double F = 1.0 / acc[0][3];
#pragma unroll
    for (int m = 0; m < NR; ++m)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[m][c] *= F;
    const double ws = -acc[1][3], wt = -acc[2][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[1][c] += ws * acc[0][c];
      acc[2][c] += wt * acc[0][c];
    }
    if constexpr (NR == 6) {
      const double wss = -acc[3][3], wst = -acc[4][3], wtt = -acc[5][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double S = acc[0][c], Su = acc[1][c], Sv = acc[2][c];
        acc[3][c] += wss * S + 2.0 * ws * Su;
        acc[4][c] += wst * S + wt * Su + ws * Sv;
        acc[5][c] += wtt * S + 2.0 * wt * Sv;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < NR; ++m)
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * m + c] = acc[m][c];
}

// Degrees 2 and 3 in each direction take the register-resident evaluator
// (FAST: the host checked that both surfaces are of these degrees); others the
// generic one.
__host__ __device__ inline bool fast_degrees(const SurfDev<double>& sd) {
  return !sd.zero && sd.degu >= 2 && sd.degu <= 3 && sd.degv >= 2 && sd.degv <= 3;
}
// ORD = 1: only S, Su, Sv (out[0..8]) are written on the register path (the
// generic path writes all 18).
template <int MAXP, bool FAST, int ORD = 2>
__device__ __forceinline__ void eval_point64(const SurfDev<double>& sd, double u, double v, double out[18]) {
  if (FAST) {
    if (sd.degu == 3) {
      if (sd.degv == 3)
        eval_fixed<3, 3, ORD>(sd, u, v, out);
      else
        eval_fixed<3, 2, ORD>(sd, u, v, out);
    } else {
      if (sd.degv == 3)
        eval_fixed<2, 3, ORD>(sd, u, v, out);
      else
        eval_fixed<2, 2, ORD>(sd, u, v, out);
    }
  } else {
    eval_generic<MAXP>(sd, u, v, out);
  }
}

// Solves M z = r for symmetric positive definite M (lower triangle used) by
// LDL^T without square roots; false when a pivot is not positive and finite.
template <int N>
__device__ bool ldl_solve(const double (&M)[4][4], const double* r, double* z) {
  double L[4][4], LD[4][4], D[4], y[4];
  for (int j = 0; j < N; ++j) {
    double s = M[j][j];
    for (int k = 0; k < j; ++k) s = s - LD[j][k] * L[j][k];
    if (!(s > 0.0 && s <= DBL_MAX)) return false;
    D[j] = s;
    for (int i = j + 1; i < N; ++i) {
      double t = M[i][j];
      for (int k = 0; k < j; ++k) t = t - LD[i][k] * L[j][k];
      LD[i][j] = t;
      L[i][j] = t / s;
    }
  }
  for (int i = 0; i < N; ++i) {
    double t = r[i];
    for (int k = 0; k < i; ++k) t = t - L[i][k] * y[k];
    y[i] = t;
  }
  for (int i = 0; i < N; ++i) y[i] = y[i] / D[i];
  for (int i = N - 1; i >= 0; --i) {
    double t = y[i];
    for (int k = N - 1; k > i; --k) t = t - L[k][i] * z[k];
    z[i] = t;
  }
  return true;
}

// Parameter centre of cell `cell` of a unit region sampled at nuv (the
// sample grid of run_evaluation_and_construction_task,
// distance_analysis.cu:92-94: step = (stop - start)/nuv, cell iu spans
// [start + step iu, start + step (iu + 1)]).
template <class TR>
__device__ inline void cell_centre(const vec4_t<TR>& R, int nuv, int cell, double* o) {
  const int w = nuv - 1;
  const int iu = cell % w, iv = cell / w;
  const double us = ((double)R.z - (double)R.x) / (double)nuv;
  const double vs = ((double)R.w - (double)R.y) / (double)nuv;
  o[0] = (double)R.x + us * ((double)iu + 0.5);
  o[1] = (double)R.y + vs * ((double)iv + 0.5);
}

constexpr int kPolishKnots = 256;  // knots per vector staged in LDS

// The four knot vectors in LDS when each fits (the span searches of every
// evaluation are chains of dependent loads), s1 / s2 then pointing there.
// Every thread of the workgroup calls it.
__device__ __forceinline__ void stage_knots(double (*sk)[kPolishKnots], SurfDev<double>& s1, SurfDev<double>& s2) {
  if (s1.nku <= kPolishKnots && s1.nkv <= kPolishKnots && s2.nku <= kPolishKnots && s2.nkv <= kPolishKnots) {
    for (int i = threadIdx.x; i < s1.nku; i += blockDim.x) sk[0][i] = s1.U[i];
    for (int i = threadIdx.x; i < s1.nkv; i += blockDim.x) sk[1][i] = s1.V[i];
    for (int i = threadIdx.x; i < s2.nku; i += blockDim.x) sk[2][i] = s2.U[i];
    for (int i = threadIdx.x; i < s2.nkv; i += blockDim.x) sk[3][i] = s2.V[i];
    __syncthreads();
    s1.U = sk[0];
    s1.V = sk[1];
    s2.U = sk[2];
    s2.V = sk[3];
  }
}

template <class TR, int MAXP, bool FAST>
__device__ __forceinline__ void k_polish_body(SurfDev<double> s1g, SurfDev<double> s2g, const Cand* rec, int m,
                                               const int* su1, const int* su2, const vec4_t<TR>* reg1,
                                               const vec4_t<TR>* reg2, int nuv1, int nuv2, double4 box1,
                                               double4 box2, PolishRec* out) {
  // the four knot vectors in LDS when they fit: the span searches of every
  // evaluation are chains of dependent loads
  __shared__ double sk[4][kPolishKnots];
  SurfDev<double> s1 = s1g, s2 = s2g;
  stage_knots(sk, s1, s2);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const Cand c = rec[e];
  const int slot = (int)(c.lo >> 40), ci = (int)((c.lo >> 20) & 0xfffff), cj = (int)(c.lo & 0xfffff);
  const double lo[4] = {box1.x, box1.y, box2.x, box2.y};
  const double hi[4] = {box1.z, box1.w, box2.z, box2.w};
  double x[4], xn[4];
  cell_centre<TR>(reg1[su1[slot]], nuv1, ci, x);
  cell_centre<TR>(reg2[su2[slot]], nuv2, cj, x + 2);
  for (int q = 0; q < 4; ++q) x[q] = clampd(x[q], lo[q], hi[q]);
  double E1[18], E2[18], F1[18], F2[18];
  eval_point64<MAXP, FAST>(s1, x[0], x[1], E1);
  eval_point64<MAXP, FAST>(s2, x[2], x[3], E2);
  double d[3] = {E1[0] - E2[0], E1[1] - E2[1], E1[2] - E2[2]};
  double f = dot3(d, d);
  double mu = kMuStart;
  int it = 0, flags = 0;
  while (it < kPolishIters) {
    ++it;
    const double* a = E1 + 3;
    const double* b = E1 + 6;
    const double* cc = E2 + 3;
    const double* ee = E2 + 6;
    double g[4];
    g[0] = dot3(a, d);
    g[1] = dot3(b, d);
    g[2] = -dot3(cc, d);
    g[3] = -dot3(ee, d);
    // coordinates on a bound whose descent direction points outward stay fixed
    int act = 0;
    for (int q = 0; q < 4; ++q)
      if ((x[q] <= lo[q] && g[q] > 0.0) || (x[q] >= hi[q] && g[q] < 0.0)) {
        act |= 1 << q;
        g[q] = 0.0;
      }
    if (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0 && g[3] == 0.0) {
      flags |= JSI_POLISH_STATIONARY;
      break;
    }
    const double J0 = dot3(a, a), J1 = dot3(b, b), J2 = dot3(cc, cc), J3 = dot3(ee, ee);
    double M[4][4];
    M[0][0] = J0 + dot3(d, E1 + 9);
    M[1][0] = dot3(a, b) + dot3(d, E1 + 12);
    M[1][1] = J1 + dot3(d, E1 + 15);
    M[2][0] = -dot3(a, cc);
    M[3][0] = -dot3(a, ee);
    M[2][1] = -dot3(b, cc);
    M[3][1] = -dot3(b, ee);
    M[2][2] = J2 - dot3(d, E2 + 9);
    M[3][2] = dot3(cc, ee) - dot3(d, E2 + 12);
    M[3][3] = J3 - dot3(d, E2 + 15);
    const double eps = 1.0e-12 * dmax(dmax(J0, J1), dmax(J2, J3));
    M[0][0] = M[0][0] + mu * (J0 + eps);
    M[1][1] = M[1][1] + mu * (J1 + eps);
    M[2][2] = M[2][2] + mu * (J2 + eps);
    M[3][3] = M[3][3] + mu * (J3 + eps);
    for (int q = 0; q < 4; ++q)
      if ((act >> q) & 1) {
        for (int r = 0; r < q; ++r) M[q][r] = 0.0;
        for (int r = q + 1; r < 4; ++r) M[r][q] = 0.0;
        M[q][q] = 1.0;
      }
    double z[4];
    if (!ldl_solve<4>(M, g, z)) {
      mu = mu * 10.0;
      if (mu > kMuMax) {
        flags |= JSI_POLISH_STALLED;
        break;
      }
      continue;
    }
    // a step below the rounding of the parameters ends the iteration
    double zm = 0.0, xm0 = 0.0;
    for (int q = 0; q < 4; ++q) {
      zm = dmax(zm, fabs(z[q]));
      xm0 = dmax(xm0, fabs(x[q]));
    }
    if (zm <= 1.0e-15 * (1.0 + xm0)) {
      flags |= JSI_POLISH_CONVERGED;
      break;
    }
    bool same = true;
    for (int q = 0; q < 4; ++q) {
      xn[q] = clampd(x[q] - z[q], lo[q], hi[q]);
      same = same && xn[q] == x[q];
    }
    if (same) {
      flags |= JSI_POLISH_FIXED;
      break;
    }
    eval_point64<MAXP, FAST>(s1, xn[0], xn[1], F1);
    eval_point64<MAXP, FAST>(s2, xn[2], xn[3], F2);
    const double dn[3] = {F1[0] - F2[0], F1[1] - F2[1], F1[2] - F2[2]};
    const double fn = dot3(dn, dn);
    if (fn < f) {
      double step = 0.0, xm = 0.0;
      for (int q = 0; q < 4; ++q) {
        step = dmax(step, fabs(xn[q] - x[q]));
        xm = dmax(xm, fabs(xn[q]));
        x[q] = xn[q];
      }
      for (int t = 0; t < 18; ++t) {
        E1[t] = F1[t];
        E2[t] = F2[t];
      }
      for (int t = 0; t < 3; ++t) d[t] = dn[t];
      f = fn;
      mu = mu * 0.1;
      if (mu < kMuMin) mu = kMuMin;
      if (step <= 1.0e-15 * (1.0 + xm)) {
        flags |= JSI_POLISH_CONVERGED;
        break;
      }
    } else {
      // f no longer decreases: a rejected step this small means the minimum
      // is resolved to the precision f allows (~sqrt(eps) in the parameters)
      if (zm <= 1.0e-9 * (1.0 + xm0)) {
        flags |= JSI_POLISH_CONVERGED;
        break;
      }
      mu = mu * 10.0;
      if (mu > kMuMax) {
        flags |= JSI_POLISH_STALLED;
        break;
      }
    }
  }
  PolishRec o;
  for (int q = 0; q < 4; ++q) o.x[q] = x[q];
  o.f = f;
  for (int t = 0; t < 3; ++t) {
    o.p1[t] = E1[t];
    o.p2[t] = E2[t];
  }
  o.iters = it;
  o.flags = flags;
  out[e] = o;
}
template <class TR, int MAXP, bool FAST>
__global__ __launch_bounds__(64) void k_polish(SurfDev<double> s1g, SurfDev<double> s2g, const Cand* rec, int m,
                                               const int* su1, const int* su2, const vec4_t<TR>* reg1,
                                               const vec4_t<TR>* reg2, int nuv1, int nuv2, double4 box1,
                                               double4 box2, PolishRec* out) {
  k_polish_body<TR, MAXP, FAST>(s1g, s2g, rec, m, su1, su2, reg1, reg2, nuv1, nuv2, box1, box2, out);
}

template <class TR, int MAXP, bool FAST>
__global__ __launch_bounds__(64) void k_curve_proj(SurfDev<double> s1, SurfDev<double> s2, const IPair* pairs,
                                                   long long n, const int* su1, const int* su2,
                                                   const vec4_t<TR>* reg1, const vec4_t<TR>* reg2, int nuv1,
                                                   int nuv2, double4 box1, double4 box2, CurveRec* out) {
  __shared__ double sk[4][kPolishKnots];
  stage_knots(sk, s1, s2);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const IPair pr = pairs[e];
  const double lo[4] = {box1.x, box1.y, box2.x, box2.y};
  const double hi[4] = {box1.z, box1.w, box2.z, box2.w};
  double x[4];
  cell_centre<TR>(reg1[su1[pr.slot]], nuv1, pr.i, x);
  cell_centre<TR>(reg2[su2[pr.slot]], nuv2, pr.j, x + 2);
  for (int q = 0; q < 4; ++q) x[q] = clampd(x[q], lo[q], hi[q]);
  double E1[18], E2[18];
  eval_point64<MAXP, FAST, 1>(s1, x[0], x[1], E1);
  eval_point64<MAXP, FAST, 1>(s2, x[2], x[3], E2);
  double d[3] = {E1[0] - E2[0], E1[1] - E2[1], E1[2] - E2[2]};
  double r2 = dot3(d, d);
  const double sc = 1.0 + dmax(dmax(fabs(E1[0]), fabs(E1[1])), fabs(E1[2]));
  const double tl = 1.0e-14 * sc;
  const double tol2 = tl * tl;
  int it = 0, flags = 0;
  for (;;) {
    if (r2 <= tol2) {
      flags |= JSI_POLISH_CONVERGED;
      break;
    }
    if (it == kCurveIters) break;
    ++it;
    const double* a = E1 + 3;
    const double* b = E1 + 6;
    const double* cc = E2 + 3;
    const double* ee = E2 + 6;
    double A[4][4];
    for (int r = 0; r < 3; ++r)
      for (int s = 0; s <= r; ++s) A[r][s] = a[r] * a[s] + b[r] * b[s] + cc[r] * cc[s] + ee[r] * ee[s];
    double y[3];
    if (!ldl_solve<3>(A, d, y)) {
      flags |= JSI_POLISH_STALLED;
      break;
    }
    const double dx[4] = {-dot3(a, y), -dot3(b, y), dot3(cc, y), dot3(ee, y)};
    double xn[4];
    bool same = true;
    for (int q = 0; q < 4; ++q) {
      xn[q] = clampd(x[q] + dx[q], lo[q], hi[q]);
      same = same && xn[q] == x[q];
    }
    if (same) {
      flags |= JSI_POLISH_FIXED;
      break;
    }
    for (int q = 0; q < 4; ++q) x[q] = xn[q];
    eval_point64<MAXP, FAST, 1>(s1, x[0], x[1], E1);
    eval_point64<MAXP, FAST, 1>(s2, x[2], x[3], E2);
    for (int t = 0; t < 3; ++t) d[t] = E1[t] - E2[t];
    r2 = dot3(d, d);
  }
  CurveRec o;
  for (int q = 0; q < 4; ++q) o.x[q] = x[q];
  for (int t = 0; t < 3; ++t) o.p[t] = E1[t];
  double n1[3], n2[3];
  cross3(E1 + 3, E1 + 6, n1);
  cross3(E2 + 3, E2 + 6, n2);
  cross3(n1, n2, o.t);
  o.r2 = r2;
  o.iters = it;
  o.flags = flags;
  out[e] = o;
}

inline int max_degree(const SurfDev<double>& a, const SurfDev<double>& b) {
  return std::max(std::max(a.degu, a.degv), std::max(b.degu, b.degv));
}

}  // namespace

// Kernel variant of a surface pair: 0 = both of degrees 2-3 (register
// evaluator), else the generic evaluator sized for the highest degree.
static int variant(const SurfDev<double>& s1, const SurfDev<double>& s2) {
  if (fast_degrees(s1) && fast_degrees(s2)) return 0;
  const int md = max_degree(s1, s2);
  return md <= 3 ? 1 : (md <= 7 ? 2 : 3);
}

#define JSI_REFINE_DISPATCH(KERNEL, ...)                                   \
  switch (variant(s1, s2)) {                                               \
    case 0: hipLaunchKernelGGL((KERNEL<3, true>), __VA_ARGS__); break;     \
    case 1: hipLaunchKernelGGL((KERNEL<3, false>), __VA_ARGS__); break;    \
    case 2: hipLaunchKernelGGL((KERNEL<7, false>), __VA_ARGS__); break;    \
    default: hipLaunchKernelGGL((KERNEL<kMaxDegP, false>), __VA_ARGS__); break; \
  }

template <class TR>
hipError_t launch_polish(const SurfDev<double>& s1, const SurfDev<double>& s2, const Cand* rec, int m, const int* su1,
                         const int* su2, const vec4_t<TR>* reg1, const vec4_t<TR>* reg2, int nuv1, int nuv2,
                         double4 box1, double4 box2, PolishRec* out, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  const dim3 g((unsigned)((m + 63) / 64));
  switch (variant(s1, s2)) {
    case 0:
      hipLaunchKernelGGL((k_polish<TR, 3, true>), g, dim3(64), 0, s, s1, s2, rec, m, su1, su2, reg1, reg2, nuv1, nuv2,
                         box1, box2, out);
      break;
    case 1:
      hipLaunchKernelGGL((k_polish<TR, 3, false>), g, dim3(64), 0, s, s1, s2, rec, m, su1, su2, reg1, reg2, nuv1,
                         nuv2, box1, box2, out);
      break;
    case 2:
      hipLaunchKernelGGL((k_polish<TR, 7, false>), g, dim3(64), 0, s, s1, s2, rec, m, su1, su2, reg1, reg2, nuv1,
                         nuv2, box1, box2, out);
      break;
    default:
      hipLaunchKernelGGL((k_polish<TR, kMaxDegP, false>), g, dim3(64), 0, s, s1, s2, rec, m, su1, su2, reg1, reg2,
                         nuv1, nuv2, box1, box2, out);
  }
  return hipGetLastError();
}

template <class TR>
hipError_t launch_curve_proj(const SurfDev<double>& s1, const SurfDev<double>& s2, const IPair* pairs, long long n,
                             const int* su1, const int* su2, const vec4_t<TR>* reg1, const vec4_t<TR>* reg2, int nuv1,
                             int nuv2, double4 box1, double4 box2, CurveRec* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g((unsigned)((n + 63) / 64));
  switch (variant(s1, s2)) {
    case 0:
      hipLaunchKernelGGL((k_curve_proj<TR, 3, true>), g, dim3(64), 0, s, s1, s2, pairs, n, su1, su2, reg1, reg2, nuv1,
                         nuv2, box1, box2, out);
      break;
    case 1:
      hipLaunchKernelGGL((k_curve_proj<TR, 3, false>), g, dim3(64), 0, s, s1, s2, pairs, n, su1, su2, reg1, reg2,
                         nuv1, nuv2, box1, box2, out);
      break;
    case 2:
      hipLaunchKernelGGL((k_curve_proj<TR, 7, false>), g, dim3(64), 0, s, s1, s2, pairs, n, su1, su2, reg1, reg2,
                         nuv1, nuv2, box1, box2, out);
      break;
    default:
      hipLaunchKernelGGL((k_curve_proj<TR, kMaxDegP, false>), g, dim3(64), 0, s, s1, s2, pairs, n, su1, su2, reg1,
                         reg2, nuv1, nuv2, box1, box2, out);
  }
  return hipGetLastError();
}

template hipError_t launch_polish<float>(const SurfDev<double>&, const SurfDev<double>&, const Cand*, int, const int*,
                                         const int*, const float4*, const float4*, int, int, double4, double4,
                                         PolishRec*, hipStream_t);
template hipError_t launch_polish<double>(const SurfDev<double>&, const SurfDev<double>&, const Cand*, int, const int*,
                                          const int*, const double4*, const double4*, int, int, double4, double4,
                                          PolishRec*, hipStream_t);
template hipError_t launch_curve_proj<float>(const SurfDev<double>&, const SurfDev<double>&, const IPair*, long long,
                                             const int*, const int*, const float4*, const float4*, int, int, double4,
                                             double4, CurveRec*, hipStream_t);
template hipError_t launch_curve_proj<double>(const SurfDev<double>&, const SurfDev<double>&, const IPair*, long long,
                                              const int*, const int*, const double4*, const double4*, int, int,
                                              double4, double4, CurveRec*, hipStream_t);

// ---------------------------------------------------------------------------
// Curve tracing.
//
// Seeds: the projections with r2 <= tol^2, in pair order, thinned to a
// spacing h (a seed is dropped when a kept seed lies closer than h); h =
// `spacing`, or 1e-3 of the accepted points' bounding-box diagonal (1 if that
// is 0).  Each seed carries its unit tangent T = t / |t|, t = (S1u x S1v) x
// (S2u x S2v).
//
// March (k_march), from a seed in direction dir T (dir = +1 forward, -1
// backward), until kMarchCap steps:
//  * predictor: the parameter step of each surface that moves its point by
//    h T in the least-squares sense, [a b]^T [a b] y = [a b]^T (h T) with
//    a, b its first derivatives;
//  * the predicted coordinates are clamped into the query boxes and the
//    corrector solves S1 - S2 = 0 and T.(S1 - p) = h (p the previous point)
//    by Gauss-Newton with minimum-norm steps (x -= J^T (J J^T)^-1 F);
//  * if an iterate leaves the boxes, its coordinates are clamped, the one that
//    was furthest outside (relative to its box width) is fixed, and the
//    corrector solves S1 - S2 = 0 over the other three (iterates clamped):
//    an edge point, after which the march ends;
//  * converged: |F|^2 <= (1e-14 (1 + max |seed coordinate|))^2, at most
//    kCorrIters iterations;
//  * after an interior step the tangent is recomputed and oriented along the
//    previous one; the march ends at the first seed closer than h/2 to the new
//    point (its successor; the start seed only after the first step),
//    otherwise the point is recorded.
// Host (assemble_curves): a seed no other seed reaches is the head of an
// open curve: its backward march (reversed), then the head, its forward
// points, its successor and so on until a march without successor.  Seeds
// left over lie on closed curves: from the lowest-index one, its chain is
// followed until it returns.
// ---------------------------------------------------------------------------
namespace {

constexpr int kCorrIters = 20;
constexpr int kMarchCap = 8192;
enum { kCorrConv = 0, kCorrFail = 1, kCorrLeft = 2, kCorrNoConv = 3 };

template <int MAXP, bool FAST>
__device__ __forceinline__ int correct(const SurfDev<double>& s1, const SurfDev<double>& s2, double* x, double* E1, double* E2,
                       const double* lo, const double* hi, int freem, bool plane, const double* T, const double* p0,
                       double h, double tol2) {
  for (int it = 0;; ++it) {
    double F[4];
    for (int c = 0; c < 3; ++c) F[c] = E1[c] - E2[c];
    double r2 = dot3(F, F);
    const int n = plane ? 4 : 3;
    if (plane) {
      const double q[3] = {E1[0] - p0[0], E1[1] - p0[1], E1[2] - p0[2]};
      F[3] = dot3(T, q) - h;
      r2 = r2 + F[3] * F[3];
    }
    if (r2 <= tol2) return kCorrConv;
    if (it == kCorrIters) return kCorrNoConv;
    const double* a = E1 + 3;
    const double* b = E1 + 6;
    const double* c = E2 + 3;
    const double* e = E2 + 6;
    double J[4][4];
    for (int r = 0; r < 3; ++r) {
      J[r][0] = a[r];
      J[r][1] = b[r];
      J[r][2] = -c[r];
      J[r][3] = -e[r];
    }
    J[3][0] = dot3(T, a);
    J[3][1] = dot3(T, b);
    J[3][2] = 0.0;
    J[3][3] = 0.0;
    for (int r = 0; r < 4; ++r)
      for (int q = 0; q < 4; ++q)
        if (!((freem >> q) & 1)) J[r][q] = 0.0;
    double A[4][4];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) A[i][j] = J[i][0] * J[j][0] + J[i][1] * J[j][1] + J[i][2] * J[j][2] + J[i][3] * J[j][3];
    double y[4];
    const bool ok = plane ? ldl_solve<4>(A, F, y) : ldl_solve<3>(A, F, y);
    if (!ok) return kCorrFail;
    bool left = false;
    for (int q = 0; q < 4; ++q) {
      double sum = J[0][q] * y[0] + J[1][q] * y[1] + J[2][q] * y[2];
      if (plane) sum = sum + J[3][q] * y[3];
      double xn = x[q] - sum;
      if (xn < lo[q] || xn > hi[q]) {
        if (plane)
          left = true;
        else
          xn = clampd(xn, lo[q], hi[q]);
      }
      x[q] = xn;
    }
    if (left) return kCorrLeft;
    eval_point64<MAXP, FAST, 1>(s1, x[0], x[1], E1);
    eval_point64<MAXP, FAST, 1>(s2, x[2], x[3], E2);
  }
}

// Clamps the coordinates outside the boxes; the one outside by the largest
// fraction of its box width (lowest index on ties) is returned as the fixed
// coordinate (bit mask, 0 when all were inside).
__device__ inline int clamp_fix(double* x, const double* lo, const double* hi) {
  int worst = -1;
  double wv = 0.0;
  for (int q = 0; q < 4; ++q) {
    double v = 0.0;
    if (x[q] < lo[q])
      v = (lo[q] - x[q]) / (hi[q] - lo[q]);
    else if (x[q] > hi[q])
      v = (x[q] - hi[q]) / (hi[q] - lo[q]);
    if (v > wv) {
      wv = v;
      worst = q;
    }
    x[q] = clampd(x[q], lo[q], hi[q]);
  }
  return worst < 0 ? 0 : 1 << worst;
}

// One wave per march: every lane steps the same march (the values are the
// same in all lanes), and the search for a seed within h/2 of the new point
// runs over 64 seeds at a time (the lowest index with a hit wins, as in a
// sequential scan; seeds are at least h apart, so at most one can hit).
// Counting pass (pts == nullptr): out[k] receives the successor, the number of
// recorded points and the end flags; writing pass: the same march again,
// recording its points at pts + offs[k].  Lane 0 writes.
template <int MAXP, bool FAST>
__global__ __launch_bounds__(64) void k_march(SurfDev<double> s1, SurfDev<double> s2, const MarchSeed* seeds,
                                              int nseeds, const int* starts, int nstart, double dir, double h,
                                              double4 box1, double4 box2, const long long* offs, MarchPt* pts,
                                              MarchOut* out) {
  const int k = blockIdx.x;
  const int lane = threadIdx.x;
  if (k >= nstart) return;  // whole workgroup
  __shared__ double sk[4][kPolishKnots];
  stage_knots(sk, s1, s2);
  const int s = starts[k];
  const double lo[4] = {box1.x, box1.y, box2.x, box2.y};
  const double hi[4] = {box1.z, box1.w, box2.z, box2.w};
  const MarchSeed sd = seeds[s];
  double x[4], T[3], p[3];
  for (int q = 0; q < 4; ++q) x[q] = sd.x[q];
  for (int c = 0; c < 3; ++c) {
    T[c] = dir * sd.T[c];
    p[c] = sd.p[c];
  }
  double E1[18], E2[18];
  eval_point64<MAXP, FAST, 1>(s1, x[0], x[1], E1);
  eval_point64<MAXP, FAST, 1>(s2, x[2], x[3], E2);
  const double sc = 1.0 + dmax(dmax(fabs(p[0]), fabs(p[1])), fabs(p[2]));
  const double tl = 1.0e-14 * sc;
  const double tol2 = tl * tl;
  const double rr = 0.5 * h;
  const double rr2 = rr * rr;
  MarchPt* dst = pts ? pts + offs[k] : nullptr;
  // the writing pass repeats the counting pass's march; its count bounds the writes
  const int room = pts ? out[k].n : 0;
  int n = 0, succ = -1, flags = 0, step = 0;
  for (; step < kMarchCap; ++step) {
    const double* a = E1 + 3;
    const double* b = E1 + 6;
    const double* c = E2 + 3;
    const double* e = E2 + 6;
    const double dl[3] = {h * T[0], h * T[1], h * T[2]};
    double G1[4][4], G2[4][4];
    G1[0][0] = dot3(a, a);
    G1[1][0] = dot3(a, b);
    G1[1][1] = dot3(b, b);
    G2[0][0] = dot3(c, c);
    G2[1][0] = dot3(c, e);
    G2[1][1] = dot3(e, e);
    const double r1[2] = {dot3(a, dl), dot3(b, dl)};
    const double r2v[2] = {dot3(c, dl), dot3(e, dl)};
    double y1[2], y2[2];
    if (!ldl_solve<2>(G1, r1, y1) || !ldl_solve<2>(G2, r2v, y2)) {
      flags |= JSI_MARCH_FAILED;
      break;
    }
    x[0] = x[0] + y1[0];
    x[1] = x[1] + y1[1];
    x[2] = x[2] + y2[0];
    x[3] = x[3] + y2[1];
    (void)clamp_fix(x, lo, hi);
    eval_point64<MAXP, FAST, 1>(s1, x[0], x[1], E1);
    eval_point64<MAXP, FAST, 1>(s2, x[2], x[3], E2);
    int fixed = 0;
    const int st = correct<MAXP, FAST>(s1, s2, x, E1, E2, lo, hi, 15, true, T, p, h, tol2);
    if (st == kCorrLeft) {
      fixed = clamp_fix(x, lo, hi);
      eval_point64<MAXP, FAST, 1>(s1, x[0], x[1], E1);
      eval_point64<MAXP, FAST, 1>(s2, x[2], x[3], E2);
      if (correct<MAXP, FAST>(s1, s2, x, E1, E2, lo, hi, 15 & ~fixed, false, T, p, h, tol2) != kCorrConv) {
        flags |= JSI_MARCH_FAILED;
        break;
      }
    } else if (st != kCorrConv) {
      flags |= JSI_MARCH_FAILED;
      break;
    }
    if (fixed == 0) {
      double n1[3], n2[3], t[3];
      cross3(E1 + 3, E1 + 6, n1);
      cross3(E2 + 3, E2 + 6, n2);
      cross3(n1, n2, t);
      const double tt = dot3(t, t);
      if (!(tt > 0.0)) {
        flags |= JSI_MARCH_FAILED;
        break;
      }
      const double inv = 1.0 / sqrt(tt);
      double tn[3] = {t[0] * inv, t[1] * inv, t[2] * inv};
      if (dot3(tn, T) < 0.0)
        for (int q = 0; q < 3; ++q) tn[q] = -tn[q];
      for (int q = 0; q < 3; ++q) T[q] = tn[q];
    }
    for (int q = 0; q < 3; ++q) p[q] = E1[q];
    // four runs of 64 seeds per iteration, so their loads are in flight
    // together; the lowest index with a hit wins
    int hit = -1;
    for (int q0 = 0; q0 < nseeds && hit < 0; q0 += 256) {
      unsigned long long m[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int q = q0 + 64 * c + lane;
        bool near = false;
        // the start seed counts only once the march has left it
        if (q < nseeds && !(q == s && step == 0)) {
          const double dq[3] = {p[0] - seeds[q].p[0], p[1] - seeds[q].p[1], p[2] - seeds[q].p[2]};
          near = dot3(dq, dq) < rr2;
        }
        m[c] = __ballot(near);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (hit < 0 && m[c] != 0ull) hit = q0 + 64 * c + __ffsll((long long)m[c]) - 1;
    }
    if (hit >= 0) {
      succ = hit;
      flags |= JSI_MARCH_SEED;
      break;
    }
    if (dst && n < room && lane == 0) {
      MarchPt& o = dst[n];
      for (int q = 0; q < 4; ++q) o.x[q] = x[q];
      for (int q = 0; q < 3; ++q) o.p[q] = E1[q];
      const double d[3] = {E1[0] - E2[0], E1[1] - E2[1], E1[2] - E2[2]};
      o.r2 = dot3(d, d);
    }
    ++n;
    if (fixed != 0) {
      flags |= JSI_MARCH_EDGE;
      break;
    }
  }
  if (step == kMarchCap) flags |= JSI_MARCH_CAP;
  MarchOut o;
  o.succ = succ;
  o.n = n;
  o.flags = flags;
  o.pad = 0;
  if (lane == 0) out[k] = o;
}

}  // namespace

hipError_t launch_march(const SurfDev<double>& s1, const SurfDev<double>& s2, const MarchSeed* seeds, int nseeds,
                        const int* starts, int nstart, double dir, double h, double4 box1, double4 box2,
                        const long long* offs, MarchPt* pts, MarchOut* out, hipStream_t s) {
  if (nstart <= 0) return hipSuccess;
  const dim3 g((unsigned)nstart);
  JSI_REFINE_DISPATCH(k_march, g, dim3(64), 0, s, s1, s2, seeds, nseeds, starts, nstart, dir, h, box1, box2, offs,
                      pts, out)
  return hipGetLastError();
}

namespace {

struct CellKey {
  int64_t a, b, c;
  bool operator==(const CellKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct CellHash {
  size_t operator()(const CellKey& k) const {
    uint64_t v = (uint64_t)k.a * 0x9E3779B97F4A7C15ull;
    v ^= (uint64_t)k.b + 0x7F4A7C159E3779B9ull + (v << 6) + (v >> 2);
    v ^= (uint64_t)k.c + 0x94D049BB133111EBull + (v << 6) + (v >> 2);
    return (size_t)v;
  }
};

inline CellKey cell_of(const double* p, double cs) {
  return CellKey{(int64_t)std::floor(p[0] / cs), (int64_t)std::floor(p[1] / cs), (int64_t)std::floor(p[2] / cs)};
}

inline double dist2(const double* p, const double* q) {
  const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
  return dot3(d, d);
}

// Cells of the kept seeds: open addressing with linear probing over a
// power-of-two table (at most half full); each cell holds the index of its
// last seed, earlier ones are chained through `next`.
struct SeedCells {
  std::vector<CellKey> key;
  std::vector<int> head;  // -1: empty slot
  std::vector<int> next;  // per seed
  size_t used = 0;
  SeedCells() : key(1024), head(1024, -1) {}
  size_t slot(const CellKey& k) const {
    const size_t mask = head.size() - 1;
    size_t i = CellHash()(k) & mask;
    while (head[i] >= 0 && !(key[i] == k)) i = (i + 1) & mask;
    return i;
  }
  int first(const CellKey& k) const { return head[slot(k)]; }
  void add(const CellKey& k, int seed) {
    if (2 * (used + 1) > head.size()) grow();
    const size_t i = slot(k);
    if (head[i] < 0) {
      key[i] = k;
      ++used;
    }
    next.push_back(head[i]);
    head[i] = seed;
  }
  void grow() {
    std::vector<CellKey> ok;
    std::vector<int> oh;
    ok.swap(key);
    oh.swap(head);
    key.assign(ok.size() * 2, CellKey{0, 0, 0});
    head.assign(oh.size() * 2, -1);
    for (size_t i = 0; i < oh.size(); ++i)
      if (oh[i] >= 0) {
        const size_t j = slot(ok[i]);
        key[j] = ok[i];
        head[j] = oh[i];
      }
  }
};

}  // namespace

double curve_seeds(const CurveRec* recs, long long n, double tol, double spacing, std::vector<MarchSeed>& seeds) {
  seeds.clear();
  const double tol2 = tol * tol;
  std::vector<long long> acc;
  acc.reserve((size_t)n);
  for (long long e = 0; e < n; ++e)
    if (recs[e].r2 <= tol2) acc.push_back(e);
  double h = spacing;
  if (acc.empty()) return h > 0.0 ? h : 1.0;
  if (!(h > 0.0)) {
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) lo[c] = hi[c] = recs[acc[0]].p[c];
    for (long long e : acc)
      for (int c = 0; c < 3; ++c) {
        const double v = recs[e].p[c];
        if (v < lo[c]) lo[c] = v;
        if (v > hi[c]) hi[c] = v;
      }
    const double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    h = 1.0e-3 * std::sqrt(dot3(ext, ext));
    if (!(h > 0.0)) h = 1.0;
  }
  const double h2 = h * h;
  SeedCells grid;
  // The seed that dropped the previous record is tried first: records in pair
  // order mostly lie close to each other.  Whether a record is dropped depends
  // only on whether some kept seed lies closer than h, not on which one.
  int last = -1;
  for (long long e : acc) {
    const CurveRec& r = recs[e];
    if (last >= 0 && dist2(r.p, seeds[last].p) < h2) continue;
    const CellKey k = cell_of(r.p, h);
    int near = -1;
    for (int da = -1; da <= 1 && near < 0; ++da)
      for (int db = -1; db <= 1 && near < 0; ++db)
        for (int dc = -1; dc <= 1 && near < 0; ++dc)
          for (int q = grid.first(CellKey{k.a + da, k.b + db, k.c + dc}); q >= 0; q = grid.next[q])
            if (dist2(r.p, seeds[q].p) < h2) {
              near = q;
              break;
            }
    if (near >= 0) {
      last = near;
      continue;
    }
    const double tt = dot3(r.t, r.t);
    if (!(tt > 0.0)) continue;  // tangential contact: no direction to march
    MarchSeed sd;
    for (int q = 0; q < 4; ++q) sd.x[q] = r.x[q];
    for (int c = 0; c < 3; ++c) sd.p[c] = r.p[c];
    const double inv = 1.0 / std::sqrt(tt);
    for (int c = 0; c < 3; ++c) sd.T[c] = r.t[c] * inv;
    sd.r2 = r.r2;
    grid.add(k, (int)seeds.size());
    seeds.push_back(sd);
  }
  return h;
}

void assemble_curves(const std::vector<MarchSeed>& seeds, const std::vector<MarchOut>& fwd,
                     const std::vector<long long>& fwd_off, const std::vector<MarchPt>& fwd_pts,
                     const std::vector<int>& heads, const std::vector<MarchOut>& bwd,
                     const std::vector<long long>& bwd_off, const std::vector<MarchPt>& bwd_pts,
                     std::vector<jsi_curve_point>& out, int* ncurves) {
  out.clear();
  const int ns = (int)seeds.size();
  std::vector<char> used(ns, 0);
  int cid = 0;
  auto emit = [&](const double* x, const double* p, double r2, int curve) {
    jsi_curve_point cp;
    cp.x = p[0];
    cp.y = p[1];
    cp.z = p[2];
    cp.u1 = x[0];
    cp.v1 = x[1];
    cp.u2 = x[2];
    cp.v2 = x[3];
    cp.residual = std::sqrt(r2);
    cp.curve = curve;
    cp.flags = 0;
    out.push_back(cp);
  };
  // a chain from seed s: the seed, its forward points, then its successor ...
  auto chain = [&](int s, int curve) -> bool {  // true when it returns to a used seed
    while (s >= 0) {
      if (used[s]) return true;
      used[s] = 1;
      emit(seeds[s].x, seeds[s].p, seeds[s].r2, curve);
      for (int q = 0; q < fwd[s].n; ++q) {
        const MarchPt& m = fwd_pts[fwd_off[s] + q];
        emit(m.x, m.p, m.r2, curve);
      }
      s = fwd[s].succ;
    }
    return false;
  };
  for (size_t e = 0; e < heads.size(); ++e) {
    const int hd = heads[e];
    if (used[hd]) continue;
    for (int q = bwd[e].n - 1; q >= 0; --q) {
      const MarchPt& m = bwd_pts[bwd_off[e] + q];
      emit(m.x, m.p, m.r2, cid);
    }
    chain(hd, cid);
    ++cid;
  }
  for (int s = 0; s < ns; ++s) {
    if (used[s]) continue;
    const size_t first = out.size();
    // closed when the chain comes back to its own first seed
    int t = s;
    bool closed = false;
    std::vector<int> visit;
    while (t >= 0 && !used[t]) {
      visit.push_back(t);
      used[t] = 1;
      t = fwd[t].succ;
    }
    closed = (t == s);
    for (int v : visit) used[v] = 0;
    chain(s, cid);
    if (closed)
      for (size_t q = first; q < out.size(); ++q) out[q].flags = JSI_CURVE_CLOSED;
    ++cid;
  }
  *ncurves = cid;
}

// ---------------------------------------------------------------------------
// Polish of a batch of distance queries (jsi_internal.hpp): one launch per
// kernel variant present, blockIdx.y = query, each query's records polished
// as alone.
// ---------------------------------------------------------------------------
int polish_variant(const SurfDev<double>& s1, const SurfDev<double>& s2) { return variant(s1, s2); }

template <class TR>
struct PolishBArgs {
  SurfDev<double> s1, s2;
  const Cand* rec;
  int m;
  const int *su1, *su2;
  const vec4_t<TR>*reg1, *reg2;
  int nuv1, nuv2;
  double4 box1, box2;
  PolishRec* out;
  const unsigned* fail;
};

template <class TR, int MAXP, bool FAST>
__global__ __launch_bounds__(64) void k_polish_b(const PolishBArgs<TR>* args) {
  const PolishBArgs<TR>& x = args[blockIdx.y];
  if (x.fail && *x.fail) return;
  if ((int)blockIdx.x * 64 >= x.m) return;
  k_polish_body<TR, MAXP, FAST>(x.s1, x.s2, x.rec, x.m, x.su1, x.su2, x.reg1, x.reg2, x.nuv1, x.nuv2, x.box1, x.box2,
                                x.out);
}

template <class T>
int record_polish_batch(BatchProgram& p, const DistBatchShape& sh, const std::vector<DistBatchQ<T>>& qs) {
  if (sh.polish <= 0 || qs.empty()) return JSI_OK;
  const int m = std::min(sh.k, sh.polish);
  const int last = (sh.rounds - 1) % 2;
  const int nuv1 = sh.rounds == 1 ? sh.nuv1 : sh.nref, nuv2 = sh.rounds == 1 ? sh.nuv2 : sh.nref;
  for (int v = 0; v < 4; ++v) {
    std::vector<PolishBArgs<T>> a;
    for (const auto& q : qs) {
      if (variant(q.p1, q.p2) != v) continue;
      PolishBArgs<T> x;
      x.s1 = q.p1;
      x.s2 = q.p2;
      x.rec = q.sorted[sh.rounds - 1];
      x.m = m;
      x.su1 = q.su[0][last];
      x.su2 = q.su[1][last];
      x.reg1 = q.reg[0][last];
      x.reg2 = q.reg[1][last];
      x.nuv1 = nuv1;
      x.nuv2 = nuv2;
      x.box1 = q.box1;
      x.box2 = q.box2;
      x.out = q.pol;
      x.fail = q.fail;
      a.push_back(x);
    }
    if (a.empty()) continue;
    const size_t off = p.put(a);
    const void* f = v == 0   ? reinterpret_cast<const void*>(&k_polish_b<T, 3, true>)
                    : v == 1 ? reinterpret_cast<const void*>(&k_polish_b<T, 3, false>)
                    : v == 2 ? reinterpret_cast<const void*>(&k_polish_b<T, 7, false>)
                             : reinterpret_cast<const void*>(&k_polish_b<T, kMaxDegP, false>);
    p.add(f, dim3((unsigned)((m + 63) / 64), (unsigned)a.size()), dim3(64), off);
  }
  return JSI_OK;
}
template int record_polish_batch<float>(BatchProgram&, const DistBatchShape&, const std::vector<DistBatchQ<float>>&);
template int record_polish_batch<double>(BatchProgram&, const DistBatchShape&, const std::vector<DistBatchQ<double>>&);

}  // namespace jsi
