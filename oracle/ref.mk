# Compiles the reference's VERSION5 hot-path library from its own sources where
# they lie (read-only) into oracle/_ref/.  Nothing is copied into the repo.
# Flags follow JSI_CAD_NEXT_VERSION5/Makefile:9 (CXXFLAGS) plus -fPIC/-shared;
# the SVE flags are aarch64-only and leave x86 on the scalar branches.
# The intersection sample surfaces (examples/data/inter_case*.cu) are linked in
# so tests can read them through the NurbsFace symbols.
REF      ?= /root/reference/JSI_CAD_NEXT_VERSION5
CXX      ?= g++
HERE     := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT      := $(HERE)_ref
CXXFLAGS := -O3 -std=c++17 -fopenmp -DJSI_CAD_PRESERVE_PAIR_RELATION=1 -I $(REF)/include/ -I $(REF)/src/ -fPIC
SRCS     := src/common/common.cu src/evaluation/evaluation.cu src/distance_analysis/distance_analysis.cu \
            src/intersection/intersection.cu
DATA     := inter_case1-1 inter_case1-2 inter_case2-1 inter_case2-2 inter_case3-1 inter_case3-2 \
            inter_case4-1 inter_case4-2
OBJS     := $(addprefix $(OUT)/,$(notdir $(SRCS:.cu=.o))) $(addprefix $(OUT)/,$(addsuffix .o,$(DATA)))

all: $(OUT)/libjsi_v5_ref.so $(OUT)/libjsi_ascend_ref.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/common.o: $(REF)/src/common/common.cu | $(OUT)
	$(CXX) $(CXXFLAGS) -x c++ -c $< -o $@
$(OUT)/evaluation.o: $(REF)/src/evaluation/evaluation.cu | $(OUT)
	$(CXX) $(CXXFLAGS) -x c++ -c $< -o $@
$(OUT)/distance_analysis.o: $(REF)/src/distance_analysis/distance_analysis.cu | $(OUT)
	$(CXX) $(CXXFLAGS) -x c++ -c $< -o $@
$(OUT)/intersection.o: $(REF)/src/intersection/intersection.cu | $(OUT)
	$(CXX) $(CXXFLAGS) -x c++ -c $< -o $@
$(OUT)/inter_case%.o: $(REF)/examples/data/inter_case%_data.cu | $(OUT)
	$(CXX) $(CXXFLAGS) -x c++ -c $< -o $@

$(OUT)/libjsi_v5_ref.so: $(OBJS)
	$(CXX) -shared -fopenmp -o $@ $(OBJS)

.PHONY: all

# The reference's hotspot operator library (CPU implementation of the
# jsi_ascend_* ABI), src/ascend/hotspots_910b/src/jsi_ascend_hotspots.cpp.
$(OUT)/libjsi_ascend_ref.so: $(REF)/src/ascend/hotspots_910b/src/jsi_ascend_hotspots.cpp | $(OUT)
	$(CXX) -O3 -std=c++17 -fPIC -shared -I $(REF)/src/ascend/hotspots_910b/include $< -o $@
