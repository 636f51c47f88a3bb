# Reference test programs compiled from their own sources where they lie
# (AMReX-free C++): Test/chemical_test.cpp holds the reference's Li-Dryer
# kinetics + thermo, Test/transport_test.cpp its mixture transport.  Each is
# built as a shared library so the tests can call its functions; `main` is
# renamed at compile time so loading the library runs nothing.  Output only
# into oracle/_ref/.
REF ?= /root/reference
CXX ?= g++
FLAGS = -O2 -std=c++17 -fPIC -shared -ffp-contract=off -fno-fast-math -w

all: _ref/libchemref.so _ref/libtransref.so

_ref/libchemref.so: $(REF)/Test/chemical_test.cpp
	mkdir -p _ref
	$(CXX) $(FLAGS) -Dmain=ocfd_ref_chemical_test_main -o $@ $<

_ref/libtransref.so: $(REF)/Test/transport_test.cpp
	mkdir -p _ref
	$(CXX) $(FLAGS) -Dmain=ocfd_ref_transport_test_main -o $@ $<
