// Test traffic of the C++ GPU tests: XFECFRAMEs carrying GSE data fields
// (BBHEADER + payload, BB-scrambled), BCH + LDPC encoded, interleaved,
// mapped and noised by the library's seeded host transmitter.
#pragma once

#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "cusat_dsp/compat.h"
#include "cusat_dsp/modcod_types.h"
#include "cusat_fec.h"

namespace {

using cusat_dsp::code_rate_t;
using cusat_dsp::fecframe_size_t;
using cusat_dsp::modulation_t;

struct traffic {
    std::vector<std::vector<uint8_t>> payloads;
    std::vector<complexf> symbols;  // n_frames x xfec
    int xfec = 0;
};

// n frames of GSE payloads (random lengths, one empty, one full) for the MODCOD
inline traffic make_traffic(fecframe_size_t fs, code_rate_t rate, modulation_t mod, int n, float es, uint32_t seed)
{
    cusat_fec_info inf;
    if (cusat_fec_query(fs, rate, mod, 0, &inf) != CUSAT_OK) throw std::runtime_error(cusat_last_error());
    const int kbch = inf.k_bch / 8;
    std::vector<uint8_t> prbs(kbch);
    cusat_bb_prbs(prbs.data(), kbch);
    std::mt19937 rng(seed);
    traffic t;
    t.xfec = inf.xfec_len;
    std::vector<uint8_t> frames((size_t)n * kbch);
    for (int b = 0; b < n; b++) {
        int len = (int)(rng() % (kbch - 10));
        if (b == 1) len = 0;
        if (b == 2) len = kbch - 10;
        std::vector<uint8_t> p(len);
        for (auto& x : p) x = (uint8_t)rng();
        uint8_t* f = &frames[(size_t)b * kbch];
        for (int i = 0; i < kbch; i++) f[i] = (uint8_t)rng();
        const int dfl = 8 * len;
        f[0] = (1 << 5) | (1 << 4) | 2;  // TS/GS generic packetized, SIS, CCM, roll-off 0.20
        f[1] = f[2] = f[3] = 0;
        f[4] = (uint8_t)(dfl >> 8);
        f[5] = (uint8_t)dfl;
        f[6] = 0;
        f[7] = f[8] = 0xff;
        f[9] = (uint8_t)cusat_bb_crc8(f, 9);
        if (len) memcpy(f + 10, p.data(), len);
        for (int i = 0; i < kbch; i++) f[i] ^= prbs[i];
        t.payloads.push_back(std::move(p));
    }
    t.symbols.resize((size_t)n * inf.xfec_len);
    if (cusat_encode_frames(fs, rate, mod, seed, n, es, frames.data(), reinterpret_cast<float*>(t.symbols.data()),
                            nullptr) != CUSAT_OK)
        throw std::runtime_error(cusat_last_error());
    return t;
}

}  // namespace
