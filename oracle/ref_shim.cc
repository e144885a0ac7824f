// C entry points over the reference's own host-side C++ sources (compiled
// where they lie under the reference tree by oracle/ref.mk) -- TEST
// INFRASTRUCTURE ONLY.  Tests compare this build's constellations,
// interleaver, BCH encoder, FEC parameters and LDPC table selection with
// these, value for value.
#include <cstdint>
#include <cstring>
#include <stdexcept>

#include "fec/bch_code.h"
#include "fec/code_params.h"
#include "ldpc_table_factory.h"
#include "modulation/bit_interleaver.h"
#include "modulation/constellation.h"

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

}  // extern "C"
