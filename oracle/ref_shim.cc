// C entry points over the reference's own host-side C++ sources (compiled
// where they lie under the reference tree by oracle/ref.mk) -- TEST
// INFRASTRUCTURE ONLY.  Tests compare this build's constellations,
// interleaver, BCH encoder, FEC parameters, LDPC table selection,
// decision-directed SNR estimates and baseband step (scrambler, BBHEADER
// CRC-8, parser, GSE BBFRAME builder) with these, value for value.
#include <cstdint>
#include <cstring>
#include <stdexcept>

#include "bb_header.h"
#include "bb_scrambler.h"
#include "bbframe_parser.h"
#include "fec/bch_code.h"
#include "fec/code_params.h"
#include "fec/crc8.h"
#include "ldpc_table_factory.h"
#include "modulation/bit_interleaver.h"
#include "modulation/constellation.h"
#include "modulation/dd_snr.h"

using namespace cusat_dsp;

extern "C" {

int ref_constellation(int mod, int rate, int framesize, float* pts, int* bits)
{
    try {
        const constellation_t c =
            make_constellation((modulation_t)mod, (code_rate_t)rate, (fecframe_size_t)framesize);
        for (int i = 0; i < c.n_points; i++) {
            pts[2 * i] = c.points[i].real();
            pts[2 * i + 1] = c.points[i].imag();
        }
        *bits = c.bits_per_sym;
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_interleaver(int mod, int rate, int framesize, uint32_t* col, int* n_cols, int* rows)
{
    try {
        const bit_interleaver_t il =
            make_bit_interleaver((modulation_t)mod, (code_rate_t)rate, (fecframe_size_t)framesize);
        for (int c = 0; c < 8; c++) col[c] = il.col_offset[c];
        *n_cols = il.n_cols;
        *rows = (int)il.rows;
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_fec_params(int framesize, int rate, int* k_bch, int* n_bch, int* t)
{
    if (!has_fec_params((fecframe_size_t)framesize, (code_rate_t)rate)) return -1;
    const fec_params_t p = get_fec_params((fecframe_size_t)framesize, (code_rate_t)rate);
    *k_bch = (int)p.k_bch;
    *n_bch = (int)p.n_bch;
    *t = (int)p.t_bch;
    return 0;
}

int ref_bch(int framesize, int n_bits, int t, const uint8_t* msg, uint8_t* cw, uint64_t* gen)
{
    try {
        const bch_code_t c = make_bch_code((fecframe_size_t)framesize, n_bits, t);
        for (int w = 0; w < 4; w++) gen[w] = c.generator[w];
        if (msg && cw) bch_encode(c, msg, cw);
        return c.parity_bits;
    } catch (...) {
        return -1;
    }
}

// Table the reference selects for (frame, rate): N, K, rows, address count
// and the addresses themselves (up to cap entries).
int ref_ldpc_table(int framesize, int rate, int* n, int* k, int* n_rows, int* links_total,
                   int* links_max_cn, uint16_t* addr, int cap)
{
    ldpc_code* c = create_ldpc_table((fecframe_size_t)framesize, (code_rate_t)rate);
    if (!c) return -1;
    *n = c->n;
    *k = c->k;
    *n_rows = c->n_rows;
    *links_total = c->links_total;
    *links_max_cn = c->links_max_cn;
    const int total = (int)c->row_off[c->n_rows];
    for (int i = 0; i < total && i < cap; i++) addr[i] = c->addr[i];
    delete c;
    return total;
}

// modulation/dd_snr.cc: QPSK sign decisions; any constellation of the
// MODCOD every stride-th symbol (sym: interleaved re, im floats)
float ref_qpsk_snr_dd(const float* sym, int n_sym)
{
    return qpsk_snr_dd(reinterpret_cast<const complexf*>(sym), n_sym);
}

float ref_constellation_snr_dd(int mod, int rate, int framesize, const float* sym, int n_sym, int stride)
{
    try {
        const constellation_t c =
            make_constellation((modulation_t)mod, (code_rate_t)rate, (fecframe_size_t)framesize);
        return constellation_snr_dd(c, reinterpret_cast<const complexf*>(sym), n_sym, stride);
    } catch (...) {
        return -1.0f;
    }
}

// bb_scrambler.cc, fec/crc8.h, bbframe_parser.cc, bb_header.cc
int ref_bb_scramble(const uint8_t* in, uint8_t* out, int n_bytes)
{
    if (n_bytes < 0 || (size_t)n_bytes > BB_SCRAMBLER_MAX_BYTES) return -1;
    bb_scramble(in, out, (size_t)n_bytes);
    return 0;
}

int ref_crc8(const uint8_t* data, int len) { return crc8_bbheader(data, len); }

// data-field bytes of a descrambled BBFRAME of kbch_bytes bytes (0 = none)
int ref_extract_gse_datafield(const uint8_t* bbframe, int kbch_bytes, uint8_t* out, int max_out)
{
    return extract_gse_datafield(bbframe, (size_t)kbch_bytes, out, max_out);
}

// one BBFRAME (not scrambled) carrying len bytes of GSE data, CCM, roll-off
// 0.20: the reference transmitter's header and zero padding; returns the
// data-field bytes it took
int ref_build_gse_bbframe(int framesize, int rate, const uint8_t* data, int len, uint8_t* out)
{
    try {
        bb_header h((fecframe_size_t)framesize, (code_rate_t)rate);
        return (int)h.build_gse_bbframe(data, (unsigned)len, out);
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
