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

// One axis of a cell's float overlap record (FP32SCREEN intersection): the
// hull [lo, hi] of the cell's FP64 acceptance interval [c - r (1 + 4 u64),
// c + r (1 + 4 u64)], widened by 4 u64 (|c| + r) for the FP64 arithmetic, and
// with has32 also of the float interval [c32 - r32 (1 + 3u), c32 + r32 (1 +
// 3u)] of VERSION5's own float boxes; stored as xo = fl32(mid) and yo =
// ru32((half + |xo - mid|)(1 + 3u) + 4 u64 (|mid| + half) + 1e-30), so two
// records whose hulls meet pass the float test.
__device__ __forceinline__ void record32_axis(double c64, double r64, bool has32, float c32, float r32, float* xo,
                                              float* yo) {
  constexpr double u = 5.9604644775390625e-08, u64 = 1.1102230246251565e-16;
  const double w = r64 * (1.0 + 4.0 * u64) + 4.0 * u64 * (fabs(c64) + r64);
  double lo = c64 - w, hi = c64 + w;
  if (has32) {
    const double w32 = (double)r32 * (1.0 + 3.0 * u) + 4.0 * u64 * (fabs((double)c32) + (double)r32);
    lo = fmin(lo, (double)c32 - w32);
    hi = fmax(hi, (double)c32 + w32);
  }
  const double mid = 0.5 * (lo + hi), half = 0.5 * (hi - lo);
  const float cf = __double2float_rn(mid);
  *xo = cf;
  *yo = __double2float_ru((half + fabs((double)cf - mid)) * (1.0 + 3.0 * u) + 4.0 * u64 * (fabs(mid) + half) +
                          1e-30);
}

constexpr int kEcKnots = 128;  // knots per direction staged in LDS (when both vectors fit)

}  // namespace jsi
