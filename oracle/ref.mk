# Compiles the two reference translation units that build from their own
# sources (src/core/math/wrap.cpp, src/core/math/math.cpp) together with a C
# shim, into oracle/_ref/libref.so.  Reads /root/reference, writes only
# oracle/_ref/.  Same flags as the reference Release build (no -march, so no
# FMA) plus -ffp-contract=off.
REF ?= /root/reference
CXX ?= g++
CXXFLAGS = -std=c++17 -O3 -ffp-contract=off -fPIC -I$(REF)/include

all: _ref/libref.so

_ref/libref.so: ref_shim.cpp $(REF)/src/core/math/wrap.cpp $(REF)/src/core/math/math.cpp
	@mkdir -p _ref
	$(CXX) $(CXXFLAGS) -shared -o $@ ref_shim.cpp $(REF)/src/core/math/wrap.cpp $(REF)/src/core/math/math.cpp

.PHONY: all
