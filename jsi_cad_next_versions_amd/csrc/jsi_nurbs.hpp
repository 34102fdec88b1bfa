// jsi_nurbs.hpp - knot span search and B-spline basis functions with
// derivatives on the device: FindSpan / DersBasisFuns
// (JSI_CAD_NEXT_VERSION5/src/evaluation/evaluation.cu:152-219; The NURBS Book
// A2.1 and A2.3), in the reference's operation order.  Shared by the grid
// kernels and the FP64 point solvers.
#pragma once

#include <hip/hip_runtime.h>

namespace jsi {

template <class T>
__device__ int find_span(int nk, int p, T u, const T* U) {
  if (u <= U[p]) return p;
  const int n = nk - p - 1;
  if (u >= U[n]) return n - 1;
  int lo = p, hi = n, mid = (lo + hi) / 2;
  // A non-decreasing knot vector ends the search within log2(n)+1 steps; the
  // cap only keeps a malformed vector from spinning forever.
  for (int it = 0; it < 64 && (u < U[mid] || u >= U[mid + 1]); ++it) {
    if (u < U[mid])
      hi = mid;
    else
      lo = mid;
    mid = (lo + hi) / 2;
  }
  return mid;
}

// basis_ders with the degree P fixed at compile time: every loop unrolls and
// the arrays stay in registers; the operations and their order are those of
// basis_ders (der <= 2).
template <class T, int P>
__device__ __forceinline__ void basis_ders_p(int span, T u, const T* U, int der, T (&ders)[3][P + 1]) {
  T ndu[P + 1][P + 1];
  T left[P + 1], right[P + 1];
  T a[2][P + 1];
  ndu[0][0] = T(1);
#pragma unroll
  for (int j = 1; j <= P; ++j) {
    left[j] = u - U[span + 1 - j];
    right[j] = U[span + j] - u;
    T saved = T(0);
#pragma unroll
    for (int r = 0; r < j; ++r) {
      ndu[j][r] = right[r + 1] + left[j - r];
      const T temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
#pragma unroll
  for (int j = 0; j <= P; ++j) ders[0][j] = ndu[j][P];
#pragma unroll
  for (int r = 0; r <= P; ++r) {
    a[0][0] = T(1);
#pragma unroll
    for (int k = 1; k <= 2; ++k) {
      if (k > der) break;
      const int s1 = (k - 1) & 1, s2 = k & 1;  // the rows swap after every order
      T d = T(0);
      const int rk = r - k, pk = P - k;
      if (r >= k) {
        a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
        d = a[s2][0] * ndu[rk][pk];
      }
      const int j1 = (rk >= -1) ? 1 : -rk;
      const int j2 = (r - 1 <= pk) ? k - 1 : P - r;
#pragma unroll
      for (int j = j1; j <= j2; ++j) {
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
  int r = P;
#pragma unroll
  for (int k = 1; k <= 2; ++k) {
    if (k > der) break;
#pragma unroll
    for (int j = 0; j <= P; ++j) ders[k][j] *= T(r);
    r *= P - k;
  }
}

template <class T, int MAXP>
__device__ void basis_ders(int span, int p, T u, const T* U, int der, T (&ders)[3][MAXP + 1]) {
  T ndu[MAXP + 1][MAXP + 1];
  T left[MAXP + 1], right[MAXP + 1];
  T a[2][MAXP + 1];
  ndu[0][0] = T(1);
  for (int j = 1; j <= p; ++j) {
    left[j] = u - U[span + 1 - j];
    right[j] = U[span + j] - u;
    T saved = T(0);
    for (int r = 0; r < j; ++r) {
      ndu[j][r] = right[r + 1] + left[j - r];
      const T temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
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
      const int tmp = s1;
      s1 = s2;
      s2 = tmp;
    }
  }
  int r = p;
  for (int k = 1; k <= der; ++k) {
    for (int j = 0; j <= p; ++j) ders[k][j] *= T(r);
    r *= p - k;
  }
}

// A2.3 rows of a runtime degree p <= MAXP written to row[k][j] (k < NK; orders
// above der are zero): degrees up to min(MAXP, 7) take the fixed-degree
// recurrence, whose arrays stay in registers (the general one indexes its
// arrays at run time and lives in scratch memory); the results are the same.
template <class T, int P, int NK, int NB>
__device__ __forceinline__ void basis_row_fixed(int span, T u, const T* U, int der, T (*row)[NB]) {
  T d[3][P + 1];
  basis_ders_p<T, P>(span, u, U, der, d);
#pragma unroll
  for (int k = 0; k < NK; ++k)
#pragma unroll
    for (int j = 0; j <= P; ++j) row[k][j] = k <= der ? d[k][j] : T(0);
}

template <class T, int MAXP, int NK, int NB>
__device__ __forceinline__ void basis_row_any(int span, int p, T u, const T* U, int der, T (*row)[NB]) {
  static_assert(NB >= MAXP + 1, "row too short");
  switch (p) {
    case 0: basis_row_fixed<T, 0, NK, NB>(span, u, U, der, row); return;
    case 1: if constexpr (MAXP >= 1) { basis_row_fixed<T, 1, NK, NB>(span, u, U, der, row); return; } break;
    case 2: if constexpr (MAXP >= 2) { basis_row_fixed<T, 2, NK, NB>(span, u, U, der, row); return; } break;
    case 3: if constexpr (MAXP >= 3) { basis_row_fixed<T, 3, NK, NB>(span, u, U, der, row); return; } break;
    case 4: if constexpr (MAXP >= 4) { basis_row_fixed<T, 4, NK, NB>(span, u, U, der, row); return; } break;
    case 5: if constexpr (MAXP >= 5) { basis_row_fixed<T, 5, NK, NB>(span, u, U, der, row); return; } break;
    case 6: if constexpr (MAXP >= 6) { basis_row_fixed<T, 6, NK, NB>(span, u, U, der, row); return; } break;
    case 7: if constexpr (MAXP >= 7) { basis_row_fixed<T, 7, NK, NB>(span, u, U, der, row); return; } break;
    default: break;
  }
  if constexpr (MAXP > 7) {
    T d[3][MAXP + 1];
    basis_ders<T, MAXP>(span, p, u, U, der, d);
    for (int k = 0; k < NK; ++k)
      for (int j = 0; j <= p; ++j) row[k][j] = k <= der ? d[k][j] : T(0);
  }
}

}  // namespace jsi
