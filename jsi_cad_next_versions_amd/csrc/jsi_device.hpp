// jsi_device.hpp - small device helpers shared by the kernel files: the
// reference's min / max conventions, order-preserving bit views of positive
// values, infinities and the output order of EvaluateOneUV's derivatives.
#pragma once

#include <hip/hip_runtime.h>

#include "jsi_internal.hpp"

namespace jsi {

// std::min / std::max: the first argument wins unless the second is strictly
// smaller / larger; the nesting order fixes the sign of zero.
template <class T>
__device__ __forceinline__ T smin(T a, T b) {
  return (b < a) ? b : a;
}
template <class T>
__device__ __forceinline__ T smax(T a, T b) {
  return (a < b) ? b : a;
}
// CadMax (common.cu:132-135)
template <class T>
__device__ __forceinline__ T cmax(T a, T b) {
  return a > b ? a : b;
}

// Bit patterns of non-negative values order like the values, so maxima of
// them can be taken with integer atomics.
__device__ __forceinline__ unsigned long long ord_bits(float x) { return (unsigned long long)__float_as_uint(x); }
__device__ __forceinline__ unsigned long long ord_bits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}
template <class T>
__device__ __forceinline__ T from_bits(unsigned long long b);
template <>
__device__ __forceinline__ float from_bits<float>(unsigned long long b) {
  return __uint_as_float((unsigned)b);
}
template <>
__device__ __forceinline__ double from_bits<double>(unsigned long long b) {
  return __longlong_as_double((long long)b);
}

template <class T>
__device__ __forceinline__ T infinity_v();
template <>
__device__ __forceinline__ float infinity_v<float>() {
  return __uint_as_float(0x7f800000u);
}
template <>
__device__ __forceinline__ double infinity_v<double>() {
  return __longlong_as_double(0x7ff0000000000000ll);
}

// Position of derivative (k = u order, l = v order) in the output
// [S, Su, Sv, Suu, Suv, Svv] (EvaluateOneUV loop order + row swap,
// evaluation.cu:281-310).
__device__ __forceinline__ int out_row(int k, int l) {
  const int m = (l == 0) ? k : (l == 1 ? 3 + k : 5);
  return m == 2 ? 3 : (m == 3 ? 2 : m);
}

// Region of cell `bbox` (id = y * (pnuv - 1) + x) of a unit sampled on a
// pnuv x pnuv grid: the refinement task of that cell (make_dist_refine_task,
// distance_analysis.cu:340-354; the same split in intersection.cu:388-420).
template <class T>
__device__ __forceinline__ vec4_t<T> child_region(const vec4_t<T>& pr, int pnuv, int bbox) {
  const int w = pnuv - 1;
  const T ustep = (pr.z - pr.x) / T(pnuv);
  const T vstep = (pr.w - pr.y) / T(pnuv);
  const int x = bbox % w, y = bbox / w;
  vec4_t<T> c;
  c.x = pr.x + ustep * T(x);
  c.z = c.x + ustep;
  c.y = pr.y + vstep * T(y);
  c.w = c.y + vstep;
  return c;
}

constexpr int kEcKnots = 128;  // knots per direction staged in LDS (when both vectors fit)

}  // namespace jsi
