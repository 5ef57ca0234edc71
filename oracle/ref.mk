# Builds the reference beagle-lib CPU path (front end + CPU and CPU-SSE plugins)
# directly from its sources under /root/reference, into oracle/_ref/.
#
# This is test infrastructure: the resulting libraries are the reference
# implementation that the oracle is checked against and the CPU baseline that
# bench.py times. Nothing here is shipped as product.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Only the reference's own sources are compiled; the preprocessor definitions
# below are the ones its top-level CMakeLists.txt:26-34 passes on the command
# line (it has no generated config.h on this path: HAVE_CONFIG_H stays unset).
# -march=x86-64-v3 instead of the reference's -march=native keeps the build
# portable to the GPU box's host CPU.

REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
ARCH     ?= -march=x86-64-v3
CXXFLAGS := -std=c++11 -O3 -pthread -fPIC $(ARCH) -w \
            -I$(REF) -I$(REF)/libhmsbeagle \
            -DPACKAGE_BUGREPORT=\"beagle-dev@googlegroups.com\" \
            -DPACKAGE_NAME=\"libhmsbeagle\" \
            -DPACKAGE_STRING=\"libhmsbeagle\ 4.0.1\" \
            -DPACKAGE_TARNAME=\"libhmsbeagle\" \
            -DPACKAGE_VERSION=\"4.0.1\" \
            -DPACKAGE_URL=\"\" \
            -DPLUGIN_VERSION=\"40\"

SRC      := $(REF)/libhmsbeagle
FRONT    := $(SRC)/beagle.cpp $(SRC)/plugin/Plugin.cpp $(SRC)/plugin/UnixSharedLibrary.cpp \
            $(SRC)/benchmark/BeagleBenchmark.cpp $(SRC)/benchmark/linalg.cpp

all: $(OUT)/libhmsbeagle.so $(OUT)/libhmsbeagle-cpu.so.40.0.0 $(OUT)/libhmsbeagle-cpu-sse.so.40.0.0

$(OUT):
	mkdir -p $(OUT)

# The front end dlopen()s plugins by bare file name; RUNPATH=$ORIGIN makes it
# find the plugins built next to it.
$(OUT)/libhmsbeagle.so: $(FRONT) | $(OUT)
	$(CXX) $(CXXFLAGS) -shared -o $@ $(FRONT) -Wl,-rpath,'$$ORIGIN' -ldl

$(OUT)/libhmsbeagle-cpu.so.40.0.0: $(SRC)/CPU/BeagleCPUPlugin.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -shared -o $@ $<

$(OUT)/libhmsbeagle-cpu-sse.so.40.0.0: $(SRC)/CPU/BeagleCPUSSEPlugin.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -shared -o $@ $<

clean:
	rm -f $(OUT)/libhmsbeagle*.so*

.PHONY: all clean
