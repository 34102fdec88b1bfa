// jsi_cad.hpp - the JSI_CAD_NEXT VERSION5 query API, served by
// libjsi_mi355x.so on an MI355X.  Replaces JSI_CAD_NEXT_VERSION5/include/jsi_cad.hpp:1-48
// with the same include guard, names, argument meaning and C++ signatures, so
// the reference's drivers (examples/*_example.cu, which include this header,
// nurbs_data.hpp and common/common.cuh) compile against this include/
// directory alone, or beside the reference's own headers (whichever copy is
// included first wins, the guards being the same).  The memory helpers
// (cadMalloc, cadMemcpy, CadMemcpyKind ...) live in common/common.cuh, as in
// the reference.  Differences:
//   * the workspace is device memory (HBM): d_ivec/d_data of a task and all
//     pointers from allocate_from_workspace are device pointers, filled with
//     cadMemcpy(..., CadMemcpyKind::HostToDevice) exactly as the drivers do;
//   * results are computed by the HIP kernels in EXACT32 mode, bit-exact with
//     the reference's float path.
#ifndef _JSI_CAD_HPP
#define _JSI_CAD_HPP

#include <string>
#include <cstdio>
#include <cstdlib>

using cudaError_t = int;
inline constexpr int cudaSuccess = 0;

// jsi_cad.hpp:11-23
struct EvalAndConstructTask {
  int* d_ivec;
  float* d_data;
  int d_ndata;

  float ustart;
  float vstart;
  float ustop;
  float vstop;
  int nuv;

  int eval_task_side;
};

void init_cad();

void deinit_cad();

void* allocate_from_workspace(size_t size);

// Returns sqrt of the k-th (k = 8) smallest AABB-centre squared distance of
// the last refinement round (distance_analysis.cu:356-506).
float minimum_distance(const EvalAndConstructTask& t1, const EvalAndConstructTask& t2);

// Prints "Round r: overlap pairs = N" per round (intersection.cu:338-475).
void intersection(EvalAndConstructTask t1, EvalAndConstructTask t2);

static int default_global_dist_nuv = 512;
static int default_global_intersect_nuv = 16;

inline void checkError(cudaError_t error, std::string msg) {
  if (error != cudaSuccess) {
    printf("%s: %d\n", msg.c_str(), error);
    exit(1);
  }
}

#define ONE_UV_SIZE 18
// results receives 18*nuv*nuv floats; all pointers may be host or device.
void evalWithCuda(int* ivec, float* data, int ndata, float ustep, float vstep, float ustart, float vstart, int nuv,
                  float* results);

#endif
