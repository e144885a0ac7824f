// BCH outer code (host side).
#include "bch.h"

#include <cstring>

namespace cusat {

namespace {

const uint32_t MINPOLY[3][12] = {
    {0x1002D, 0x10173, 0x10FBD, 0x15A55, 0x11F2F, 0x1F7B5, 0x1AF65, 0x17367, 0x10EA1,
     0x175A7, 0x13A2D, 0x11AE3},
    {0x402B, 0x4941, 0x4647, 0x5591, 0x6B55, 0x6389, 0x6CE5, 0x4F21, 0x460F, 0x5A49, 0x5811,
     0x65EF},
    {0x802D, 0x8C93, 0xB555, 0x8D6D, 0x94D7, 0xB0D1, 0xD8B5, 0xCB55, 0xBAB7, 0xBC9F, 0x8A17,
     0xED15},
};

bool bit(const poly256& p, int i) { return (p[i >> 6] >> (i & 63)) & 1; }

int deg(const poly256& p)
{
    for (int i = 255; i >= 0; i--)
        if (bit(p, i)) return i;
    return -1;
}

poly256 shl(const poly256& p, int s)
{
    poly256 r{};
    const int ws = s >> 6, bs = s & 63;
    for (int w = 3; w >= ws; w--) {
        uint64_t v = p[w - ws] << bs;
        if (bs && w - ws - 1 >= 0) v |= p[w - ws - 1] >> (64 - bs);
        r[w] = v;
    }
    return r;
}

void xor_in(poly256& a, const poly256& b)
{
    for (int w = 0; w < 4; w++) a[w] ^= b[w];
}

poly256 mod(poly256 a, const poly256& g, int dg)
{
    for (int i = deg(a); i >= dg; i--)
        if (bit(a, i)) xor_in(a, shl(g, i - dg));
    return a;
}

}  // namespace

bool make_bch(int framesize, int n_bits, int t, bch_code* c)
{
    if (framesize < 0 || framesize > 2 || t < 1 || t > 12) return false;
    poly256 g{};
    g[0] = 1;
    for (int i = 0; i < t; i++) {
        poly256 r{};
        for (int s = 0; s < 32; s++)
            if ((MINPOLY[framesize][i] >> s) & 1) xor_in(r, shl(g, s));
        g = r;
    }
    c->g = g;
    c->degree = deg(g);
    c->n_bits = n_bits;
    c->k_bits = n_bits - c->degree;
    if (c->k_bits <= 0 || (c->degree & 7)) return false;
    c->lut.assign(256, poly256{});
    for (int b = 0; b < 256; b++) {
        poly256 p{};
        p[0] = (uint64_t)b;
        c->lut[b] = mod(shl(p, c->degree), g, c->degree);
    }
    return true;
}

void bch_encode(const bch_code& c, const uint8_t* msg, uint8_t* cw)
{
    const int p = c.degree;
    const int kb = c.k_bits / 8;
    poly256 reg{};
    auto byte_at = [](const poly256& r, int lo) {
        const int w = lo >> 6, b = lo & 63;
        uint64_t v = r[w] >> b;
        if (b > 56 && w < 3) v |= r[w + 1] << (64 - b);
        return (uint8_t)v;
    };
    for (int i = 0; i < kb; i++) {
        const uint8_t top = byte_at(reg, p - 8) ^ msg[i];
        reg = shl(reg, 8);
        // keep x^0 .. x^(p-1)
        for (int w = 0; w < 4; w++) {
            const int lo = 64 * w;
            if (p <= lo) reg[w] = 0;
            else if (p < lo + 64) reg[w] &= (1ull << (p - lo)) - 1;
        }
        xor_in(reg, c.lut[top]);
    }
    memcpy(cw, msg, kb);
    for (int j = 0; j < p / 8; j++) cw[kb + j] = byte_at(reg, p - 8 * (j + 1));
}

std::vector<poly256> bch_powers(const bch_code& c, size_t n_pow)
{
    std::vector<poly256> pw(n_pow);
    const int p = c.degree;
    pw[0] = poly256{};
    pw[0][0] = 1;
    for (size_t e = 1; e < n_pow; e++) {
        poly256 a = shl(pw[e - 1], 1);
        if (bit(a, p)) xor_in(a, c.g);
        pw[e] = a;
    }
    return pw;
}

}  // namespace cusat
