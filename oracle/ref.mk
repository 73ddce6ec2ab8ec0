# Cross-check library from the reference's own CPU sources, compiled where
# they lie (read-only) into oracle/_ref/.  Test infrastructure only.
REF ?= /root/reference/src/CPU
CXX ?= g++
CXXFLAGS = -O2 -std=c++17 -fPIC -w -I$(REF)
all: _ref/libref_psm.so
_ref/convex_cell.o: $(REF)/convex_cell.cpp
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -c $< -o $@
_ref/ref_wrap.o: ref_wrap.cpp $(REF)/Delaunay_psm.cpp
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -c $< -o $@
_ref/libref_psm.so: _ref/convex_cell.o _ref/ref_wrap.o
	$(CXX) -shared -o $@ $^ -lpthread
.PHONY: all
