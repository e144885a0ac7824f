// BCH outer code (host side): generator, byte tables, systematic encoder.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace cusat {

// GF(2) polynomial, bit i of word i/64 = coefficient of x^i (degree < 256)
using poly256 = std::array<uint64_t, 4>;

struct bch_code {
    int n_bits = 0, k_bits = 0, degree = 0;
    poly256 g{};
    // lut[b] = b(x) x^degree mod g(x): one byte of remainder update
    std::vector<poly256> lut;
};

// g = product of the first t minimal polynomials of EN 302 307-1 Table 6a
// (normal) / 6b (short) or EN 302 307-2 (medium) (fec/bch_code.cc:14-19,99-122).
bool make_bch(int framesize, int n_bits, int t, bch_code* out);

// Systematic encode of k_bits/8 message bytes -> n_bits/8 codeword bytes,
// MSB first (fec/bch_code.cc:124-143).
void bch_encode(const bch_code& c, const uint8_t* msg, uint8_t* cw);

// x^e mod g for e = 0..n_pow-1 (lowest coefficient in bit 0 of word 0)
std::vector<poly256> bch_powers(const bch_code& c, size_t n_pow);

}  // namespace cusat
