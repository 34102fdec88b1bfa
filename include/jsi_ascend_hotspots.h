/*
 * jsi_ascend_hotspots.h - the JSI_CAD_NEXT hotspot operator ABI, served by
 * libjsi_mi355x.so on an MI355X.  Same symbols, arguments and return codes as
 * JSI_CAD_NEXT_VERSION5/src/ascend/hotspots_910b/include/jsi_ascend_hotspots.h:10-13,
 * which the reference's loader resolves with dlsym
 * (src/ascend/ascend_hotspots.cpp:16-21): pointing its dlopen name
 * "libjsi_ascend_hotspots.so" at this library (a symlink) swaps in the GPU.
 * All pointers are host pointers; 0 = ok, -1 = bad argument or failure.
 */
#ifndef JSI_MI355X_ASCEND_HOTSPOTS_H
#define JSI_MI355X_ASCEND_HOTSPOTS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* k-th smallest centre distance over all cnt1 x cnt2 pairs (centres are AoS
 * xyz floats), and masks of the cells that take part in a pair at or below it
 * (jsi_ascend_hotspots.cpp:33-81). */
int jsi_ascend_distance_topk_mask(const float* center1, size_t cnt1, const float* center2, size_t cnt2, int topk,
                                  int* mask1, int* mask2, float* kth_distance);

/* masks of the cells that overlap any cell of the other side
 * (|c1-c2| < r1+r2 on all axes, jsi_ascend_hotspots.cpp:83-113). */
int jsi_ascend_intersection_mask(const float* center1, const float* radius1, int cnt1, const float* center2,
                                 const float* radius2, int cnt2, int* mask1, int* mask2);

#ifdef __cplusplus
}
#endif

#endif
