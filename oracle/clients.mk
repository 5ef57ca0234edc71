# The reference's example clients (examples/hmctest/*.cpp,
# examples/tinytest/tinytest.cpp) compiled from their sources under
# /root/reference against this repository's include/beagle.h and linked to
# beagle-lib_amd/libhmsbeagle_amd.so, into oracle/_ref/clients/: the C-level
# drop-in check (a client of the reference API builds and runs unchanged on
# this library).  Test infrastructure only; tests/test_c_clients.py runs the
# binaries on the GPU and compares their printed results with the known
# answers the sources state.
#
#   make -f oracle/clients.mk            (from the repository root, after the
#                                         library is built)
#
# `#include "libhmsbeagle/beagle.h"` resolves to include/beagle.h (copied
# under the name clients include): that directory comes first on the include
# path; the clients' other include, libhmsbeagle/BeagleImpl.h, is read from
# the reference tree and itself includes libhmsbeagle/beagle.h, which again
# resolves to this repository's header.

REF      ?= /root/reference
OUT      ?= oracle/_ref/clients
INC      := $(OUT)/include
LIBDIR   := beagle-lib_amd
CXX      ?= g++
CXXFLAGS := -std=c++11 -O2 -w -I$(INC) -I$(REF)
LDFLAGS  := -L$(LIBDIR) -lhmsbeagle_amd -Wl,-rpath,'$$ORIGIN/../../../$(LIBDIR)'

HMC      := hmctest5 hmcGapTest hmctestMedium
all: $(addprefix $(OUT)/,$(HMC)) $(OUT)/tinytest

$(INC)/libhmsbeagle/beagle.h: include/beagle.h
	mkdir -p $(INC)/libhmsbeagle
	cp $< $@

$(OUT)/%: $(REF)/examples/hmctest/%.cpp $(INC)/libhmsbeagle/beagle.h $(LIBDIR)/libhmsbeagle_amd.so
	$(CXX) $(CXXFLAGS) -o $@ $< $(LDFLAGS)

$(OUT)/tinytest: $(REF)/examples/tinytest/tinytest.cpp $(INC)/libhmsbeagle/beagle.h $(LIBDIR)/libhmsbeagle_amd.so
	$(CXX) $(CXXFLAGS) -o $@ $< $(LDFLAGS)

clean:
	rm -rf $(OUT)

.PHONY: all clean
