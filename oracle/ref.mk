# Builds oracle/_ref/libref_host.so from the reference's own host-side C++
# sources, compiled where they lie (nothing is copied), plus ref_shim.cc.
# Test infrastructure only; needs the reference tree at REF.
REF ?= /root/reference
SRC = $(REF)/lib/cusat_dsp
CXX ?= g++
CXXFLAGS = -O2 -fPIC -std=c++17 -I$(SRC) -I$(SRC)/include

_ref/libref_host.so: ref_shim.cc
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -shared -o $@ ref_shim.cc \
	    $(SRC)/modulation/constellation.cc $(SRC)/modulation/bit_interleaver.cc \
	    $(SRC)/modulation/dd_snr.cc \
	    $(SRC)/fec/bch_code.cc $(SRC)/fec/code_params.cc $(SRC)/ldpc_table_factory.cc \
	    $(SRC)/bb_scrambler.cc $(SRC)/bbframe_parser.cc $(SRC)/bb_header.cc
