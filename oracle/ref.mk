# Builds the reference CheMPS2 library from its own C++ sources, where they lie under
# /root/reference/CheMPS2 (read only), together with the driver ref_driver.cpp, into
# oracle/_ref/ (git-ignored). Test infrastructure only: the product never links or runs it.
#
#   make -f oracle/ref.mk -j8
#
# The reference's CMake (CMakeLists.txt, CheMPS2/CMakeLists.txt) is not used. The flags
# are the ones it would pass: -DH5_USE_110_API (CMakeLists.txt:43), -DCHEMPS2_MKL when
# linked against MKL (CMakeLists.txt:126-128), OpenMP on, C++11. BLAS/LAPACK come from
# MKL and HDF5 from /opt/conda, both part of this image. No file of the reference is
# copied, edited or generated: it has no generated headers (the only configure_file is
# Doxyfile.in).
REF ?= /root/reference/CheMPS2
CONDA ?= /opt/conda
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
CXX ?= g++

SRCS := $(filter-out $(REF)/executable.cpp,$(wildcard $(REF)/*.cpp))
OBJS := $(patsubst $(REF)/%.cpp,$(OUT)/obj/%.o,$(SRCS))
CXXFLAGS := -O3 -march=x86-64-v3 -fopenmp -fPIC -std=c++11 -w \
            -DH5_USE_110_API -DCHEMPS2_MKL -DCHEMPS2_VERSION=\"reference\" \
            -I$(REF)/include/chemps2 -I$(CONDA)/include
# MKL and HDF5 by path (no -L), so that the system libstdc++ (newer than the one beside
# them) is linked and found first at run time
LDLIBS := $(CONDA)/lib/libmkl_rt.so $(CONDA)/lib/libhdf5.so -Wl,-rpath-link,$(CONDA)/lib \
          -Wl,-rpath,/usr/lib/x86_64-linux-gnu -Wl,-rpath,$(CONDA)/lib

all: $(OUT)/libchemps2_ref.so $(OUT)/ref_driver

$(OUT)/obj/%.o: $(REF)/%.cpp
	@mkdir -p $(OUT)/obj
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libchemps2_ref.so: $(OBJS)
	$(CXX) -shared -fopenmp -o $@ $(OBJS) $(LDLIBS)

$(OUT)/ref_driver: $(HERE)ref_driver.cpp $(OUT)/libchemps2_ref.so
	$(CXX) $(CXXFLAGS) -o $@ $< -L$(OUT) -lchemps2_ref -Wl,-rpath,'$$ORIGIN' $(LDLIBS)

clean:
	rm -rf $(OUT)

.PHONY: all clean

# The level-2 wiring of INTEGRATION.md compiled against the reference's headers and library
# and against libchemps2_amd.so (tests/integration/level2_check.cpp); run on a GPU.
AMD := $(HERE)../chemps2_amd
$(OUT)/level2_check: $(HERE)../tests/integration/level2_check.cpp $(OUT)/libchemps2_ref.so $(AMD)/libchemps2_amd.so
	$(CXX) $(CXXFLAGS) -I$(HERE)../include -o $@ $< -L$(OUT) -lchemps2_ref -L$(AMD) -lchemps2_amd \
	  -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../chemps2_amd' -Wl,-rpath,/opt/rocm/lib $(LDLIBS)

level2: $(OUT)/level2_check
