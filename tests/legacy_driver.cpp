// legacy_driver.cpp - a driver of the VERSION5 C++ API written the way the
// reference's examples are (examples/distance_analysis_example.cu:39-163,
// intersection_example.cu:30-96, evaluation_example.cu:7-27): it includes
// jsi_cad.hpp, nurbs_data.hpp and common/common.cuh, places surfaces in the
// workspace with allocate_from_workspace + cadMemcpy, and calls
// minimum_distance / intersection / evalWithCuda / compute_K / PrefixSumAdd
// through their C++ (mangled) symbols in libjsi_mi355x.so.
//
// Commands (one per argument group, run in order in one process, so the
// intersection refinement nuv carries across calls as in the reference):
//   dist  <surface file 1> <surface file 2> <nuv>
//   isect <surface file 1> <surface file 2> <nuv> <use_k 0|1>
//   eval  <surface file> <nuv> <output file>      (18*nuv*nuv floats, then compute_K)
//   prefix <n>
// A surface file holds int32 ivec[7], int32 ndata, float32 data[ndata].
// Results go to stdout as "RESULT ..." lines between the library's own output.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "jsi_cad.hpp"
#include "nurbs_data.hpp"
#include "common/common.cuh"

namespace {

struct HostFace {
  int ivec[7];
  std::vector<float> data;
  NurbsFace face() {
    NurbsFace f;
    f.ivec = ivec;
    f.ndata = (int)data.size();
    f.data = data.data();
    return f;
  }
};

HostFace read_face(const char* path) {
  HostFace h;
  FILE* f = std::fopen(path, "rb");
  int nd = 0;
  if (!f || std::fread(h.ivec, sizeof(int), 7, f) != 7 || std::fread(&nd, sizeof(int), 1, f) != 1 || nd <= 0) {
    std::printf("cannot read %s\n", path);
    std::exit(2);
  }
  h.data.resize((size_t)nd);
  if (std::fread(h.data.data(), sizeof(float), (size_t)nd, f) != (size_t)nd) {
    std::printf("short file %s\n", path);
    std::exit(2);
  }
  std::fclose(f);
  return h;
}

// The task setup of the drivers: surface data in the workspace, the whole
// knot domain, the given nuv and side.
EvalAndConstructTask make_task(NurbsFace& obj, int nuv, int side) {
  int* d_ivec = (int*)allocate_from_workspace(7 * sizeof(int));
  float* d_data = (float*)allocate_from_workspace(obj.ndata * sizeof(float));
  if (!d_ivec || !d_data || cadMemcpy(d_ivec, obj.ivec, 7 * sizeof(int), CadMemcpyKind::HostToDevice) != 0 ||
      cadMemcpy(d_data, obj.data, obj.ndata * sizeof(float), CadMemcpyKind::HostToDevice) != 0) {
    std::printf("workspace setup failed\n");
    std::exit(2);
  }
  EvalAndConstructTask t;
  t.d_ivec = d_ivec;
  t.d_data = d_data;
  t.d_ndata = obj.ndata;
  t.ustart = obj.get_ubegin();
  t.vstart = obj.get_vbegin();
  t.ustop = obj.get_uend();
  t.vstop = obj.get_vend();
  t.nuv = nuv;
  t.eval_task_side = side;
  return t;
}

unsigned bits(float x) {
  unsigned b;
  std::memcpy(&b, &x, 4);
  return b;
}

}  // namespace

int main(int argc, char** argv) {
  init_cad();
  for (int a = 1; a < argc;) {
    const std::string op = argv[a];
    if (op == "dist" && a + 3 < argc) {
      HostFace h1 = read_face(argv[a + 1]), h2 = read_face(argv[a + 2]);
      NurbsFace f1 = h1.face(), f2 = h2.face();
      const int nuv = std::atoi(argv[a + 3]);
      EvalAndConstructTask t1 = make_task(f1, nuv, 1), t2 = make_task(f2, nuv, 2);
      const float d = minimum_distance(t1, t2);
      std::printf("RESULT dist %u\n", bits(d));
      a += 4;
    } else if (op == "isect" && a + 4 < argc) {
      HostFace h1 = read_face(argv[a + 1]), h2 = read_face(argv[a + 2]);
      NurbsFace f1 = h1.face(), f2 = h2.face();
      const int nuv = std::atoi(argv[a + 3]);
      setenv("CAD_INTERSECT_USE_K", std::atoi(argv[a + 4]) ? "1" : "0", 1);
      EvalAndConstructTask t1 = make_task(f1, nuv, 1), t2 = make_task(f2, nuv, 2);
      intersection(t1, t2);
      std::printf("RESULT isect\n");
      a += 5;
    } else if (op == "eval" && a + 3 < argc) {
      // evaluation_example.cu:10-19 with host pointers
      HostFace h = read_face(argv[a + 1]);
      NurbsFace obj = h.face();
      const int nuv = std::atoi(argv[a + 2]);
      const float ustep = (obj.get_uend() - obj.get_ubegin()) / nuv;
      const float vstep = (obj.get_vend() - obj.get_vbegin()) / nuv;
      std::vector<float> res((size_t)ONE_UV_SIZE * nuv * nuv);
      evalWithCuda(obj.ivec, obj.data, obj.ndata, ustep, vstep, obj.get_ubegin(), obj.get_vbegin(), nuv, res.data());
      FILE* f = std::fopen(argv[a + 3], "wb");
      if (!f || std::fwrite(res.data(), sizeof(float), res.size(), f) != res.size()) {
        std::printf("cannot write %s\n", argv[a + 3]);
        return 2;
      }
      std::fclose(f);
      std::vector<float> deriv((size_t)3 * nuv * nuv);
      float k = 0.0f;
      compute_K(res.data(), deriv.data(), &k, nuv);
      std::printf("RESULT eval %u %u %u\n", bits(ustep), bits(vstep), bits(k));
      a += 4;
    } else if (op == "prefix" && a + 1 < argc) {
      const int n = std::atoi(argv[a + 1]);
      int32_t* d_in = (int32_t*)allocate_from_workspace(n * sizeof(int32_t));
      int32_t* d_out = (int32_t*)allocate_from_workspace(n * sizeof(int32_t));
      std::vector<int32_t> h(n);
      for (int i = 0; i < n; ++i) h[i] = (int32_t)(((long long)i * 7919) % 23) - 11;
      cadMemcpy(d_in, h.data(), n * sizeof(int32_t), CadMemcpyKind::HostToDevice);
      const int last = PrefixSumAdd(d_in, d_out, (size_t)n);
      cadMemcpy(h.data(), d_out, n * sizeof(int32_t), CadMemcpyKind::DeviceToHost);
      long long sum = 0;
      for (int i = 0; i < n; ++i) sum += h[i];
      std::printf("RESULT prefix %d %lld\n", last, sum);
      a += 2;
    } else {
      std::printf("bad command at %s\n", argv[a]);
      return 2;
    }
  }
  std::fflush(stdout);
  deinit_cad();
  return 0;
}
