// Constellations, bit interleaver and demapper geometry (host side).
#pragma once

#include <array>
#include <cstdint>

namespace cusat {

struct constellation {
    int bits_per_sym = 0;
    int n_points = 0;
    std::array<float, 512> xy{};  // (re, im) by label, first label bit = MSB
};

// Points of EN 302 307-1/-2 clause 5.4 for (modulation, rate, frame) as
// modulation/constellation.cc:461-493 selects them.  False if unsupported.
bool make_constellation(int modulation, int rate, int framesize, constellation* out);

struct interleaver {
    int n_cols = 0;                  // 0 = QPSK, no interleaving
    int rows = 0;
    std::array<uint32_t, 8> col{};   // column offsets (bit b of a symbol at col[b] + j)
};
// Column-row interleaver of EN 302 307 clause 5.3.3 / Table 9a
// (modulation/bit_interleaver.cc:14-95).
bool make_interleaver(int modulation, int rate, int framesize, interleaver* out);

// XFECFRAME length in symbols and the number of symbols carrying FECFRAME
// bits (128APSK: 6 stuffing bits in, 12 padding symbols out;
// demapper_kernel.cu:241-247).
void xfec_lengths(int modulation, int n_ldpc, int n_mod, int* xfec_len, int* demap_len);

// Max over (point, bit) of the squared distance to the nearest opposite-bit
// point (demapper_kernel.cu:216-237): QPSK 2, 8PSK 4.
float llr_dmax(int modulation, const constellation& c);

}  // namespace cusat
