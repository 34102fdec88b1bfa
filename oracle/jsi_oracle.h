/*
 * jsi_oracle.h - CPU oracle for the JSI_CAD_NEXT VERSION5 minimum-distance and
 * intersection hot path.  TEST INFRASTRUCTURE ONLY: it is the checker that the
 * HIP library is compared against, and the CPU baseline leg of bench.py.  The
 * product (libjsi_mi355x.so) never links, loads or calls it.
 *
 * Each function restates the reference algorithm in the same floating-point
 * operation order (no contraction) and is templated on the scalar: float
 * reproduces VERSION5 bit for bit, double is the FP64 oracle.
 */
#ifndef JSI_ORACLE_H
#define JSI_ORACLE_H

#include "../include/jsi_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* evaluation.cu:358-374 (grid of nuv*nuv points, 18 values each) */
int orc_eval_grid_f32(const int* ivec, const float* data, float ustep, float vstep, float ustart, float vstart,
                      int nuv, float* out);
int orc_eval_grid_f64(const int* ivec, const float* data, double ustep, double vstep, double ustart,
                      double vstart, int nuv, double* out);

/* basis functions and derivatives at one parameter: evaluation.cu:152-219.
 * ders receives 3*(p+1) values (row k = derivative order); returns the span. */
int orc_basis_f32(int nknots, int degree, const float* knots, float u, int der, float* ders);
int orc_basis_f64(int nknots, int degree, const double* knots, double u, int der, double* ders);

/* distance AABB centres (distance_analysis.cu:45-86): center[3*(nuv-1)^2] */
int orc_aabb_dist_f32(int nuv, const float* pts18, float* center);
int orc_aabb_dist_f64(int nuv, const double* pts18, double* center);
/* intersection AABBs (intersection.cu:106-148) */
int orc_aabb_isect_f32(int nuv, const float* pts18, float pad, float* center, float* radius);
int orc_aabb_isect_f64(int nuv, const double* pts18, double pad, double* center, double* radius);
/* compute_K (common.cu:149-172) */
float orc_compute_k_f32(const float* pts18, int nuv);
double orc_compute_k_f64(const double* pts18, int nuv);

/* Whole queries.  params->precision selects float (EXACT32) or double (FP64,
 * FP32SCREEN).  pairs_out receives the last round's pair list.  params->polish
 * and params->param_tol as in jsi_min_distance (FP64 closest points restated). */
int orc_min_distance(const jsi_params* params, const int* ivec1, const float* data1, const jsi_region* r1,
                     const int* ivec2, const float* data2, const jsi_region* r2, jsi_result* result,
                     jsi_pair* pairs_out, int pairs_cap);
int orc_intersect(const jsi_params* params, const int* ivec1, const float* data1, const jsi_region* r1,
                  const int* ivec2, const float* data2, const jsi_region* r2, jsi_result* result,
                  jsi_pair* pairs_out, int pairs_cap);

/* Intersection curves: orc_intersect's rounds, then the FP64 projection of
 * the last round's pairs onto the curve and their ordering (the steps of
 * jsi_intersect_curves, restated). */
int orc_intersect_curves(const jsi_params* params, const int* ivec1, const float* data1, const jsi_region* r1,
                         const int* ivec2, const float* data2, const jsi_region* r2, double spacing, double tol,
                         jsi_result* result, jsi_curve_point* pts, int cap, int* npts, int* ncurves);

/* Screen only: top-k of one slot of distance centres (distance_analysis.cu:202-338)
 * given as float AoS arrays; for brute-force checks. */
int orc_topk_f32(int n1, const float* c1, int n2, const float* c2, int k, jsi_pair* out);

/* operator ABI of the reference's hotspot library (jsi_ascend_hotspots.cpp:33-113) */
int orc_distance_topk_mask(const float* c1, size_t n1, const float* c2, size_t n2, int topk, int* m1, int* m2,
                           float* kth);
int orc_intersection_mask(const float* c1, const float* r1, int n1, const float* c2, const float* r2, int n2, int* m1,
                          int* m2);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
