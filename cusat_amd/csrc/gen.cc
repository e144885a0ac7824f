// Seeded test-traffic generator (host CPU): info bits -> BCH -> IRA LDPC ->
// interleave + map -> AWGN.  Frames are independent: frame b draws from its
// own std::mt19937 seeded with (seed, b), so results do not depend on the
// number of worker threads.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "bch.h"
#include "codes.h"
#include "cusat_fec.h"
#include "modulation.h"

namespace cusat {
void set_last_error(const char* fmt, ...);
}

using namespace cusat;

namespace {

inline int get_bit(const uint8_t* b, unsigned i) { return (b[i >> 3] >> (7 - (i & 7))) & 1; }

// IRA encode (ldpc_encode_kernel.cu:1-17): parity accumulators
// p[(a + (i mod 360) q) mod R] ^= u_i, then p[c] ^= p[c-1].
void ldpc_encode(const ldpc_table& t, const uint8_t* info, uint8_t* cw)
{
    const int K = t.K, R = t.N - t.K, q = R / CIRC;
    std::vector<uint8_t> p(R, 0);
    for (int i = 0; i < K; i++) {
        if (!get_bit(info, i)) continue;
        const int jj = i % CIRC;
        for (int a : t.rows[i / CIRC]) p[(a + jj * q) % R] ^= 1;
    }
    for (int c = 1; c < R; c++) p[c] ^= p[c - 1];
    memcpy(cw, info, K / 8);
    memset(cw + K / 8, 0, R / 8);
    for (int c = 0; c < R; c++)
        if (p[c]) cw[(K + c) >> 3] |= (uint8_t)(0x80 >> ((K + c) & 7));
}

}  // namespace

namespace {

int gen_impl(int framesize, int rate, int modulation, uint64_t seed, int batch, float es_n0_db,
             const uint8_t* info_in, float* sym_out, uint8_t* info_out, uint8_t* cw_out)
{
    fec_params fp;
    const char* tname = ldpc_table_name(framesize, rate);
    const ldpc_table* tab = tname ? find_table(tname) : nullptr;
    constellation con;
    interleaver il;
    if (batch < 0 || !sym_out || !get_fec_params(framesize, rate, &fp) || !tab ||
        !make_constellation(modulation, rate, framesize, &con) ||
        !make_interleaver(modulation, rate, framesize, &il)) {
        set_last_error("cusat_gen_frames: unsupported (frame %d, rate %d, modulation %d)", framesize,
                       rate, modulation);
        return CUSAT_E_INVALID;
    }
    bch_code bch;
    if (!make_bch(framesize, fp.n_bch, fp.t, &bch) || bch.k_bits != fp.k_bch) {
        set_last_error("cusat_gen_frames: BCH geometry");
        return CUSAT_E_INVALID;
    }
    const int N = tab->N, n_mod = con.bits_per_sym;
    int xfec, dlen;
    xfec_lengths(modulation, N, n_mod, &xfec, &dlen);
    const bool noisy = es_n0_db < 1000.0f;
    const float n0 = std::pow(10.0f, -es_n0_db / 10.0f);
    const float sigma = std::sqrt(n0 / 2.0f);
    uint32_t cols[8];
    unsigned stride = 1;
    if (il.n_cols == 0) { cols[0] = 0; cols[1] = 1; stride = 2; }
    else for (int b = 0; b < 8; b++) cols[b] = il.col[b];
    const unsigned stuffed = modulation == CUSAT_MOD_128APSK ? (unsigned)N + 6 : (unsigned)N;

    auto work = [&](int b0, int b1) {
        std::vector<uint8_t> info(fp.k_bch / 8), bcw(fp.n_bch / 8), cw((stuffed + 7) / 8 + 1);
        for (int b = b0; b < b1; b++) {
            std::seed_seq ss{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)b};
            std::mt19937 rng(ss);
            if (info_in)
                memcpy(info.data(), info_in + (size_t)b * info.size(), info.size());
            else
                for (auto& x : info) x = (uint8_t)(rng() & 0xff);
            bch_encode(bch, info.data(), bcw.data());
            std::fill(cw.begin(), cw.end(), 0);
            ldpc_encode(*tab, bcw.data(), cw.data());
            if (info_out) memcpy(info_out + (size_t)b * info.size(), info.data(), info.size());
            if (cw_out) memcpy(cw_out + (size_t)b * (N / 8), cw.data(), N / 8);
            std::normal_distribution<float> nd(0.0f, sigma);
            float* s = sym_out + (size_t)b * xfec * 2;
            for (int j = 0; j < xfec; j++) {
                int label;
                if (j < dlen) {
                    label = 0;
                    for (int k = 0; k < n_mod; k++) {
                        const unsigned pos = cols[k] + (unsigned)j * stride;
                        const int bit = pos < stuffed && pos < (unsigned)N ? get_bit(cw.data(), pos) : 0;
                        label |= bit << (n_mod - 1 - k);
                    }
                } else {
                    label = con.n_points - 1;  // 128APSK padding: all-ones label
                }
                float re = con.xy[2 * label], im = con.xy[2 * label + 1];
                if (noisy) {
                    re += nd(rng);
                    im += nd(rng);
                }
                s[2 * j] = re;
                s[2 * j + 1] = im;
            }
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nt = std::min<unsigned>(nt, (unsigned)std::max(batch, 1));
    std::vector<std::thread> th;
    const int per = (batch + (int)nt - 1) / (int)nt;
    for (unsigned t = 0; t < nt; t++) {
        const int b0 = (int)t * per, b1 = std::min(batch, b0 + per);
        if (b0 < b1) th.emplace_back(work, b0, b1);
    }
    for (auto& x : th) x.join();
    return CUSAT_OK;
}

}  // namespace

extern "C" int cusat_gen_frames(int framesize, int rate, int modulation, uint64_t seed, int batch,
                                float es_n0_db, float* sym_out, uint8_t* info_out, uint8_t* cw_out)
{
    return gen_impl(framesize, rate, modulation, seed, batch, es_n0_db, nullptr, sym_out, info_out, cw_out);
}

extern "C" int cusat_encode_frames(int framesize, int rate, int modulation, uint64_t seed, int batch,
                                   float es_n0_db, const uint8_t* info_in, float* sym_out, uint8_t* cw_out)
{
    if (!info_in) {
        set_last_error("cusat_encode_frames: null input");
        return CUSAT_E_INVALID;
    }
    return gen_impl(framesize, rate, modulation, seed, batch, es_n0_db, info_in, sym_out, nullptr, cw_out);
}
