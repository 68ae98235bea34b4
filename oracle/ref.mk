# Builds oracle/_ref/libref_distance.so from FullC0ntact's own sources where
# they lie (read-only) plus oracle/ref_shim.cpp.  Only the point-triangle
# distance builds standalone; the CPU DistanceMap includes the CUDA runtime
# header through grid/uniformgrid.h and the particle kernels are CUDA, so
# those are not built here.  TEST INFRASTRUCTURE ONLY.
REF     ?= /root/reference
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
CXX     ?= g++
INCDIRS := $(shell find $(REF)/inshape3dcore $(REF)/math $(REF)/util -type d 2>/dev/null)
CXXFLAGS := -std=c++11 -O2 -fPIC -ffp-contract=off $(addprefix -I,$(INCDIRS))
SRCS    := $(REF)/inshape3dcore/distance/distancetriangle.cpp \
           $(REF)/inshape3dcore/distance/distance.cpp \
           $(HERE)ref_shim.cpp

all: $(OUT)/libref_distance.so

$(OUT)/libref_distance.so: $(SRCS)
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -shared -o $@ $(SRCS)

clean:
	rm -rf $(OUT)

.PHONY: all clean
