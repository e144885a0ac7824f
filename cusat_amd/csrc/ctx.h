// Private: the context behind the opaque cusat_fec_ctx handle of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "bch.h"
#include "codes.h"
#include "modulation.h"

namespace cusat {
void set_last_error(const char* fmt, ...);
}

struct cusat_fec_ctx {
    int framesize, rate, modulation, max_trials, max_batch;
    unsigned flags = 0;
    cusat::fec_params fp{};
    cusat::decoder_geometry geo;
    int cmax = 0;      // kernel variant: info circulants per layer
    bool exact = false;
    cusat::constellation con;
    cusat::interleaver il;
    int xfec = 0, dlen = 0;
    float dmax = 0, llr_target = 320.0f;
    float grid_scale = 0;  // symbol grid of the sorted APSK demap (demap_grid_scale)
    cusat::bch_code bch;
    int bch_chunk = 0;
    int device = 0;
    // device state
    void* d_tables = nullptr;  // one allocation holding every table below
    uint32_t* d_circ = nullptr;
    int *d_lay_cnt = nullptr, *d_lay_off = nullptr, *d_lay_moff = nullptr, *d_grp_off = nullptr;
    int* d_steps = nullptr;
    int* d_step_defer = nullptr;  // [kern_steps] deferred collision groups (1) per step
    int kern_steps = 0;  // schedule steps as launched (padded to even)
    int* d_groups = nullptr;
    float* d_pts = nullptr;
    uint64_t *d_bch_lut = nullptr, *d_bch_pw = nullptr;
    uint8_t* d_prbs = nullptr;  // BB scrambler PRBS, k_bch/8 bytes
    int8_t* d_msg = nullptr;  // max_batch x ldpc_msg_bytes(q, cmax)
    int32_t* d_mask = nullptr;  // [max_batch] fallback mask of cusat_rx_fec_learned
    // learned check-node parameters [n_param_iters][q] (int32 on the device)
    int* d_beta = nullptr;
    int* d_alpha = nullptr;
    int n_param_iters = 0, param_cap = 0;
    // bumped by every cusat_ldpc_set_params: a stream context re-uploads
    // its copy of the parameter pointers and counts when it differs
    unsigned param_gen = 0;
    // parameter buffers outgrown by a larger table; a decode enqueued (or a
    // graph captured) before the change may still read them, so they are
    // freed with the context, not when replaced
    std::vector<int*> retired_params;
    // test-traffic generator tables (cusat_gen_frames_dev), made on first use
    void* d_gen = nullptr;
    uint64_t* d_gen_pw = nullptr;
    int *d_gen_off = nullptr, *d_gen_addr = nullptr;
    int gen_chunk = 0;
};
