// nurbs_data.hpp - the NurbsFace surface record of JSI_CAD_NEXT VERSION5
// (JSI_CAD_NEXT_VERSION5/include/nurbs_data.hpp:1-28): the reference's sample
// data files (examples/data/*_data.cu) define these objects, its drivers read
// them.  Layout read by the library (evaluation.cu:239-257):
//   ivec[7] = {flags, degU, nCPu, nKu, degV, nCPv, nKv}
//   data    = U[nKu] | V[nKv] | CP xyz[3*nCPu*nCPv] (u fastest) | w[nCPu*nCPv] if flags & 2
#pragma once

struct NurbsFace {
  int* ivec;
  int ndata;
  float* data;

  float get_ubegin() { return data[0]; }
  float get_uend() { return data[ivec[3] - 1]; }
  float get_vbegin() { return data[ivec[3]]; }
  float get_vend() { return data[ivec[3] + ivec[6] - 1]; }
};

// Sample surfaces the reference's drivers name (defined by the data files).
extern struct NurbsFace dist_case1_obj1;
extern struct NurbsFace dist_case1_obj2;
extern struct NurbsFace dist_case2_obj1;
extern struct NurbsFace dist_case2_obj2;
extern struct NurbsFace dist_case_hard_parallel_obj1;
extern struct NurbsFace dist_case_hard_parallel_obj2;
extern struct NurbsFace dist_hard_wave_obj1;
extern struct NurbsFace dist_hard_wave_obj2;

extern struct NurbsFace inter_case1_obj1;
extern struct NurbsFace inter_case1_obj2;
extern struct NurbsFace inter_case2_obj1;
extern struct NurbsFace inter_case2_obj2;
extern struct NurbsFace inter_case3_obj1;
extern struct NurbsFace inter_case3_obj2;
extern struct NurbsFace inter_case4_obj1;
extern struct NurbsFace inter_case4_obj2;
