# Compiles the reference's self-contained integer sources (math.cpp, prime_factorization.cpp,
# header-only scrambler.hpp) with a C binding into oracle/_ref/libref.so. Test infrastructure;
# run from oracle/ by __graft_entry__.build() when /root/reference exists. The rest of the
# reference (generators, back ends) needs CMake-generated headers and is not built.
REF ?= /root/reference
CXX ?= g++
OUT := _ref

all: $(OUT)/libref.so

$(OUT)/libref.so: ref_binding.cpp $(REF)/src/base/math.cpp $(REF)/src/base/prime_factorization.cpp
	@mkdir -p $(OUT)
	$(CXX) -O2 -std=c++17 -fPIC -shared -I$(REF)/src/base -o $@.tmp $^ && mv $@.tmp $@

.PHONY: all
