# Builds oracle/_ref/libref_mat.so from the reference's standalone matrix
# headers where they lie (REF defaults to /root/reference), with the oracle's
# own flags so both follow the same FP64 evaluation order.
#   make -f oracle/ref.mk            (from the repository root)
REF ?= /root/reference
OUT := oracle/_ref
CXXFLAGS := -O2 -ffp-contract=off -fPIC -shared -std=c++11

$(OUT)/libref_mat.so: oracle/ref_mat_harness.cc $(REF)/Utilities/matlib.hh $(REF)/Utilities/mat3.hh $(REF)/Utilities/mat3.cc
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -I$(REF)/Utilities oracle/ref_mat_harness.cc $(REF)/Utilities/mat3.cc -o $@
