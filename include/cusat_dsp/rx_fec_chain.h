/*
 * RX FEC chain: demap -> deinterleave -> LDPC -> BCH check -> BB descramble
 * -> BBHEADER / GSE data field, with the public C++ API of
 * lib/cusat_dsp/rx_fec_chain.h:53-196 (gpu_decode_budget, rx_fec_chain:
 * same constructor, methods, statistics and static acceptance test).
 *
 * Underneath, every micro-batch is one call sequence of the C ABI
 * (include/cusat_fec.h) on one of the chain's HIP streams: host->device copy
 * of the symbols, cusat_stream_rx_plan_async, one replay of a captured
 * cusat_stream_rx graph (demap, LDPC + BCH, BB descramble + header + GSE
 * data field, all on the GPU), one device->host copy of the per-frame
 * results into the flight slot.
 * Frames are submitted into a ring of pinned flight slots, launched when a
 * micro-batch is full (subject to the shared decode budget) and drained in
 * submission order (rx_fec_chain.cc:180-392).
 *
 * Differences a caller sees: HIP errors throw std::runtime_error (the
 * reference aborts); streams are created at the default priority; the
 * diagnostic switches (set_diag_*) are accepted and write nothing.
 */
#ifndef CUSAT_AMD_DSP_RX_FEC_CHAIN_H
#define CUSAT_AMD_DSP_RX_FEC_CHAIN_H

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "compat.h"
#include "modcod_types.h"

struct cusat_fec_ctx;
struct cusat_stream_ctx;
struct ihipStream_t;
struct ihipEvent_t;
struct hipGraphExec;

namespace cusat_dsp {

/** Budget of LDPC decode workgroups in flight on the GPU, shared by the
 *  chains of an ACM pool (rx_fec_chain.h:53-60). */
struct gpu_decode_budget {
    std::atomic<int> in_flight{0};  // frames launched, not drained
    int cap = 0;                    // 0 = unlimited

    /** cap = one decode workgroup per compute unit of the current device;
     *  CUSAT_LDPC_INFLIGHT_CAP overrides (0 disables the cap). */
    static std::shared_ptr<gpu_decode_budget> make_default();
};

class rx_fec_chain
{
public:
    rx_fec_chain(fecframe_size_t framesize, code_rate_t rate, modulation_t constellation, int max_trials = 25,
                 int cuda_max_batch = 64, int n_gpu_streams = 8, int micro_batch = 4);
    ~rx_fec_chain();
    rx_fec_chain(const rx_fec_chain&) = delete;
    rx_fec_chain& operator=(const rx_fec_chain&) = delete;

    /** Queue one XFECFRAME (copied into a pinned slot; the caller's buffer
     *  is free on return).  Returns true when completed frames produced
     *  output (get_gse_output).  A frame of the wrong length is counted as
     *  an error and dropped. */
    bool submit_xfecframe(const complexf* symbols, int n_symbols);

    /** GSE data fields of the frames drained since the last call, in order. */
    int get_gse_output(uint8_t* out, int max_bytes);
    const std::vector<int>& gse_dfl_sizes() const { return d_gse_dfl_sizes; }

    void reset();
    void flush();
    bool drain_completed();

    float snr_db() const { return d_snr_db; }
    uint64_t frame_count() const { return d_frame_cnt; }
    uint64_t frame_error_count() const { return d_frame_error_cnt; }
    int batch_size() const { return d_batch_size; }
    int max_output_bytes() const { return d_kbch_bytes * d_batch_size; }
    float data_snr_db() const { return d_snr_db; }

    void set_diag_fec_bench(bool on) { d_diag[0] = on; }
    void set_diag_llr_histogram(bool on) { d_diag[1] = on; }
    void set_diag_ldpc_convergence(bool on) { d_diag[2] = on; }
    void set_block_when_full(bool on) { d_block_when_full = on; }

    /** Demapper noise variance / LLR gain from an external (pilot) SNR
     *  estimate (rx_fec_chain.h:148-156). */
    void update_snr(float snr_db)
    {
        if (snr_db > 0.0f) {
            d_pilot_snr_db = snr_db;
            const float snr_lin = std::pow(10.0f, snr_db / 10.0f);
            d_snr_db = snr_db;
            d_noise_var = 1.0f / snr_lin;
            d_llr_gain = std::fmin(4.0f / d_noise_var, 1000.0f);
        }
    }

    float precision() const { return d_llr_gain; }
    int pending_count() const { return d_batch_idx; }
    void launch_partial();
    void warmup();
    std::chrono::steady_clock::time_point batch_start_time() const { return d_batch_start_time; }
    uint64_t gpu_drop_count() const { return d_gpu_drop_cnt; }
    int gpu_inflight() const { return d_launch_idx - d_drain_idx; }
    int gpu_pending_launch() const { return d_submit_idx - d_launch_idx; }
    uint64_t force_drain_count() const { return d_force_drain_cnt; }
    void set_cuda_mutex(std::mutex* mtx) { d_mtx = mtx; }
    void set_decode_budget(std::shared_ptr<gpu_decode_budget> b)
    {
        if (b) d_budget = std::move(b);
    }

    /** A frame is usable only if LDPC converged and BCH validated it. */
    static bool frame_passes_fec_checks(int ldpc_trials_remaining, bool bch_valid)
    {
        return ldpc_trials_remaining >= 0 && bch_valid;
    }

private:
    static constexpr int N_FLIGHT = 128;

    struct flight_slot {
        complexf* syms = nullptr;   // pinned, micro_batch x xfecframe (64-byte aligned frames)
        // one pinned block, the layout of stream_res::blk
        unsigned char* blk = nullptr;
        int32_t* trials = nullptr;  // micro_batch each
        int32_t* valid = nullptr;
        int32_t* dfl = nullptr;     // data-field bytes (-1 = FEC failure)
        uint8_t* gse = nullptr;     // micro_batch x max data field
        ihipEvent_t* event = nullptr;
        bool in_flight = false;
        int stream_idx = -1;
        int n_frames = 0;
    };
    struct stream_res {
        ihipStream_t* stream = nullptr;
        cusat_stream_ctx* sc = nullptr;  // over the chain's one context
        complexf* sym_dev = nullptr;
        unsigned char* blk = nullptr;    // trials | valid | dfl | data fields
        hipGraphExec* graph = nullptr;   // cusat_stream_rx
    };

    modulation_t d_modulation;
    int d_n_mod = 0;
    std::vector<complexf> d_points;  // constellation, for decision-directed SNR
    int d_kbch_bytes = 0, d_max_dfl = 0, d_kldpc_bytes = 0, d_nldpc = 0, d_xfecframe_len = 0;
    int d_sym_stride = 0;     // symbols per frame slot (xfecframe rounded up to 8)
    size_t d_res_bytes = 0;   // trials | valid | dfl in a result block
    cusat_fec_ctx* d_ctx = nullptr;
    int d_batch_size = 0, d_batch_idx = 0, d_micro_batch = 4, d_n_streams = 8;
    std::vector<uint8_t> d_output;
    int d_output_len = 0;
    std::vector<int> d_gse_dfl_sizes;
    float d_snr_db = 0, d_pilot_snr_db = 0, d_noise_var = 1.0f, d_llr_gain = 4.0f;
    uint64_t d_frame_cnt = 0, d_frame_error_cnt = 0;
    int d_snr_skip_count = 0;
    bool d_diag[3] = {false, false, false};
    bool d_block_when_full = false;
    std::chrono::steady_clock::time_point d_batch_start_time;
    std::mutex* d_mtx = nullptr;
    std::vector<flight_slot> d_flight;
    std::vector<stream_res> d_streams;
    int d_submit_idx = 0, d_launch_idx = 0, d_drain_idx = 0, d_next_stream = 0;
    uint64_t d_force_drain_cnt = 0, d_gpu_drop_cnt = 0;
    std::shared_ptr<gpu_decode_budget> d_budget;

    void launch_slot(flight_slot& slot);
    void pump_launches(bool bypass_cap = false);
    void release_budget(int n) { if (d_budget) d_budget->in_flight.fetch_sub(n, std::memory_order_relaxed); }
    void force_drain_oldest();
    void post_process_slot(flight_slot& slot);
    void update_dd_snr(const complexf* frame);
    void free_all();
};

/** Decision-directed SNR estimates (modulation/dd_snr.cc:23-51): signal /
 *  error energy of hard decisions, 1e6 for error-free input, 0 for
 *  non-finite input.  QPSK: every symbol, sign decisions; others: nearest
 *  constellation point of every stride-th symbol. */
float qpsk_snr_dd(const complexf* sym, int n_sym);
float constellation_snr_dd(const complexf* points, int n_points, const complexf* sym, int n_sym, int stride);

}  // namespace cusat_dsp

#endif
