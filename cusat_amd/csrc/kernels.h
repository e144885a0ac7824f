// Launchers of the gfx950 kernels (host side declarations).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "codes.h"

namespace cusat {

struct ldpc_args {
    const int16_t* llr;        // [B][N]
    int8_t* msg;               // [B][ldpc_msg_bytes] check-to-variable message records
    uint8_t* hard;             // [B][K/8]
    int32_t* result;           // [B]
    const uint32_t* circ;      // circulant records (4 dwords each), per layer
    const int* lay_cnt;        // [q]
    const int* lay_off;        // [q]
    const int* lay_moff;       // [q]
    const int* steps;          // [n_steps][2], n_steps even (padded with idle steps)
    const int* grp_off;        // [n_steps + 1]
    const int* groups;         // [n_groups][8] int32 collision group records
    const int* step_defer;     // [n_steps] 1 = the step's groups are applied during the next step
    const int* beta;           // [n_param_iters][q] offsets, or null (plain: 1)
    const int* alpha;          // [n_param_iters][q] scales in 64ths, or null (off)
    int n_param_iters;
    const int32_t* active;     // [B] nonzero = decode the frame, or null (all)
    bool split;                // use the split-edge kernel (ldpc_kernel_split.h) when it exists
    int N, K, R, q, n_steps, lanes, n_slots, n_groups, lt_alloc, max_trials, cmax;
    bool exact;
};

// dynamic LDS of the decode kernel: APP + parity + delta slots + schedule,
// group offsets and group records (n_steps as launched)
int ldpc_lds_bytes(int K, int R, int n_slots, int n_steps, int n_groups);
// Kernel variant for a code: CMAX info circulants per layer (exact when every
// layer has cnt == CMAX, else guarded), lanes 1|2.  Returns false when no
// variant fits.
bool ldpc_select_variant(int cnt_min, int cnt_max, int lanes, int* cmax, bool* exact);
// message record bytes per frame for a variant
size_t ldpc_msg_bytes(int q, int cmax);
hipError_t launch_ldpc_decode(const ldpc_args& a, int batch, hipStream_t stream);
}  // namespace cusat
struct cusat_fec_ctx;
namespace cusat {
// the decode arguments of a context (tables, geometry, its message
// workspace) for the given LLR and output buffers
ldpc_args make_ldpc_args(const cusat_fec_ctx* c, const int16_t* llr, uint8_t* hard, int32_t* trials);
// look up (once per device, with the device current) the address of the
// decoder's static message scratch; 0 on success
int ldpc_scratch_setup();
// split-edge kernel: exists for every variant; its LDS adds the partial
// exchange area, its message records are [layer][half][pair]
bool ldpc_split_available(int cmax, bool exact, int lanes);
int ldpc_lds_bytes_split(int K, int R, int n_slots, int n_steps, int n_groups, int lanes);
size_t ldpc_msg_bytes_split(int q, int cmax);
// split-edge kernel decoding two frames of a one-lane code at once (one
// per lane): two frame regions, then the shared tables and wg_any2 flags
int ldpc_lds_bytes_pair(int K, int R, int n_slots, int n_steps, int n_groups);
template <int CMAX, bool EXACT, int LANES>
hipError_t launch_ldpc_split_variant(const ldpc_args& a, int batch, hipStream_t stream);
template <int CMAX, bool EXACT, int LANES>
hipError_t launch_ldpc_variant(const ldpc_args& a, int batch, hipStream_t stream);

struct demap_args {
    const float* sym;          // [B][xfec][2]
    int16_t* llr;              // [B][N]
    const float* pts;          // [2 n_points] device (APSK), or null
    int xfec, dlen, N, n_mod, n_points;
    uint32_t col[8];
    float qpsk_scale, psk8_scale, gain;
    float grid_scale;          // sorted APSK demap: cells per unit of the 32 x 32 symbol grid / 2
};
hipError_t launch_demap(const demap_args& a, int batch, hipStream_t stream);
// Constellations demapped by the sorted path (one workgroup per frame, its
// symbols counting-sorted over a 32 x 32 grid so that each wave holds
// neighbouring symbols and skips the points none of them needs): APSK
// tables of at least SORTED_DEMAP_MIN_POINTS points and at most
// SORTED_DEMAP_MAX_SYMBOLS demapped symbols.
constexpr int SORTED_DEMAP_MIN_POINTS = 64;
constexpr int DEMAP_BLOCK = 512;  // threads of a mixed-batch demap workgroup
constexpr int SORTED_DEMAP_MAX_SYMBOLS = 16200;
inline bool demap_sorted(bool table, int n_points, int dlen)
{
    return table && n_points >= SORTED_DEMAP_MIN_POINTS && dlen <= SORTED_DEMAP_MAX_SYMBOLS;
}
// A sorted-path frame is demapped by demap_parts workgroups: each sorts the
// whole frame and demaps one 64-aligned share of the sorted symbols.
// Redoing the sort is cheap next to the demap, and a frame then no longer
// sets the launch's length by itself.
constexpr int DEMAP_MAX_PARTS = 4;
constexpr int64_t DEMAP_PART_COST = 1 << 19;  // symbols x points per part
inline int demap_parts(int dlen, int n_points)
{
    const int64_t p = ((int64_t)dlen * n_points + DEMAP_PART_COST / 2) / DEMAP_PART_COST;
    return p < 1 ? 1 : p > DEMAP_MAX_PARTS ? DEMAP_MAX_PARTS : (int)p;
}
// Symbols per work item of a mixed batch's frames outside the sorted path:
// enough work per item that the queue fetch and the frame lookup stay
// small beside it (QPSK / 8PSK closed forms cost a few operations per
// symbol, a table of 2^n_mod points about n_mod 2^n_mod).
inline int demap_span(bool table, int n_mod)
{
    return !table ? 8 * DEMAP_BLOCK : n_mod <= 5 ? 2 * DEMAP_BLOCK : DEMAP_BLOCK;
}
// grid scale of a constellation of maximum radius rmax: the grid spans
// [-1.05 rmax, 1.05 rmax] in both axes
inline float demap_grid_scale(float rmax) { return rmax > 0.0f ? 16.0f / (1.05f * rmax) : 0.0f; }

// Mixed-MODCOD receive batch (cusat_stream_rx): per code and per frame
// records in device memory.
struct rx_code {
    const float* pts;          // constellation (APSK), null for the QPSK / 8PSK formulas
    const uint8_t* prbs;       // BB scrambler sequence, kbch_bytes
    const float* dd_pts;       // constellation of every code (decision-directed SNR)
    int xfec, dlen, N, n_mod, n_points;
    int kbch_bytes, hard_bytes;  // k_bch / 8, k_ldpc / 8
    int has_cap;               // the demapper gain is capped at gain_cap
    float gain_cap;
    float grid_scale;          // demap_grid_scale of the constellation (sorted path)
    int span;                  // demap_span: symbols per mixed-batch work item
    uint32_t col[8];
};
struct rx_rec {
    int code, blk0;            // blk0: the frame's first demap block (-1: sorted path)
    int64_t sym_off;           // complex symbols
    int64_t llr_off;           // LLRs
    int64_t hard_off;          // bytes
    int64_t out_off;           // data-field bytes
    float qpsk_scale, psk8_scale, gain;
    int flags;                 // RX_SCALE_STATE: scales from the batch's state snapshot
};
constexpr int RX_SCALE_STATE = 1;
// SNR events of a receive batch, applied in order to the carried per-code
// state: frame >= 0 = the decision-directed estimate from that frame's
// symbols (applied when above 1), frame < 0 = the value snr_lin (snr_db
// kept for reporting)
struct rx_event {
    int code, frame;
    float snr_lin, snr_db;
};
// per-code demapper scales of a batch (state after its events)
struct rx_scale {
    float qpsk_scale, psk8_scale, gain, snr_lin;
};
// hdr[0] = frames, hdr[1] = work items (blocks of a code's span of
// symbols) of the frames outside the sorted path, hdr[2] = SNR events, hdr[3] = work items of the sorted
// path, hdr[4] = dlist entries, hdr[8, 9] = the demap launch's work queue
// and exit count (0 in a fresh plan; the launch leaves them 0) -- device
// memory.  dlist: the sorted-path items first, costliest frame first, each
// (frame number, part | parts << 8) for the demap_parts of its frame; then
// per other frame, costliest first, (frame number, first block), blocks
// numbered in that order.  One launch: a fixed grid takes work items from
// the queue -- a sorted-path item per workgroup, then the blocks of the
// others.
hipError_t launch_demap_mixed(const rx_code* codes, const rx_rec* recs, const int2* dlist, const rx_scale* snap,
                              int* hdr, const float* sym, int16_t* llr, int n_cu, hipStream_t stream);
// est[e] = decision-directed SNR of the frame of every estimate event
hipError_t launch_dd_estimate(const rx_code* codes, const rx_rec* recs, const rx_event* ev, const int* hdr,
                              const float* sym, float* est, int grid, hipStream_t stream);
// apply the events in order to state[n_codes] (noise_var, llr_gain, snr_lin,
// snr_db); write the batch's snapshot and, when state_out is not null, a copy
// of the new state
hipError_t launch_dd_resolve(const rx_code* codes, const rx_event* ev, const float* est, const int* hdr,
                             int n_codes, float4* state, rx_scale* snap, float4* state_out, hipStream_t stream);
hipError_t launch_bb_extract_mixed(const rx_code* codes, const rx_rec* recs, const int* hdr, const uint8_t* hard,
                                   const int32_t* trials, const int32_t* valid, uint8_t* out, int32_t* dfl_out,
                                   int grid, hipStream_t stream);

struct bch_args {
    const uint8_t* hard;       // [B][n_bytes]
    int32_t* valid;            // [B]
    const uint64_t* lut;       // [256][W]
    const uint64_t* pw;        // [64][P][W]
    const int32_t* active;     // [B] nonzero = check the frame, or null (all)
    int n_bytes, degree, chunk;
};
hipError_t launch_bch_check(const bch_args& a, int batch, hipStream_t stream);
// mask[b] = 1 unless trials[b] >= 0 and valid[b] == 1 (a frame the plain
// decoder must redo); optional per-frame flags (trials / valid may be pinned host)
hipError_t launch_fallback_mask(const int32_t* trials, const int32_t* valid, int32_t* mask,
                                int batch, hipStream_t stream);

}  // namespace cusat
