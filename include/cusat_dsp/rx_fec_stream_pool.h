/*
 * ACM receive pool on the persistent mixed-MODCOD decoder: the public API of
 * rx_fec_pool (lib/cusat_dsp/rx_fec_pool.h:26-156: constructor, submit,
 * get_gse_output, launch_stale, drain / finish, update_snr, counters) with
 * the per-MODCOD chains replaced by ONE flight ring of mixed micro-batches.
 *
 * The reference routes each frame to its MODCOD's rx_fec_chain and, when the
 * MODCOD changes, flushes the previous chain (rx_fec_pool.cc:55-77) so that
 * output stays in arrival order -- a host wait on every MODCOD switch.  Here
 * a micro-batch holds frames of any enabled MODCOD in arrival order and goes
 * to the GPU as one graph replay of cusat_stream_rx (demap, persistent
 * LDPC + BCH, BB extraction: three launches, whatever the mix), after an
 * asynchronous plan upload.  Output order is arrival order by construction,
 * with no flush barrier; launch_stale sends a partial micro-batch exactly as
 * a chain's.
 *
 * What a caller sees, compared with rx_fec_pool:
 *  - the concatenated output bytes, the data-field sizes and the per-MODCOD
 *    frame and error counters are the same for the same frames (each frame
 *    is decoded bit-identically to its MODCOD's chain at the same scaling);
 *  - an output block holds one MODCOD's consecutive data fields of one drain
 *    (block boundaries may fall elsewhere than rx_fec_pool's);
 *  - the demapper scaling of a frame is its MODCOD's SNR state when the
 *    micro-batch is closed (a chain reads it when it launches one; the two
 *    agree whenever the budget lets a micro-batch launch at once).  The
 *    decision-directed SNR is refreshed every 32nd frame of that MODCOD
 *    (rx_fec_chain.cc:321-342).  By default the estimate runs on the host
 *    at submit, as a chain does.  set_gpu_dd_snr(true) runs it on the GPU
 *    in the micro-batch's own launch sequence instead
 *    (cusat_stream_rx_dd_estimate and _resolve: bit-identical sums, no host
 *    pass over the symbols; only the resolve steps wait for the previous
 *    micro-batch's); update_snr values then travel with the micro-batch as
 *    events, in submit order, and data_snr_db() follows once a micro-batch
 *    is drained;
 *  - micro_batch = frames per launch across all MODCODs; n_gpu_streams
 *    streams (each with its own stream context and captured graph) overlap
 *    copies and launches; the decode budget caps frames in flight as
 *    gpu_decode_budget does; frames are dropped (and counted) only when the
 *    ring is full: max(8, 512 / micro_batch) micro-batches, about the 512
 *    frames of a chain's 128 micro-batches of 4;
 *  - set_learned_params gives a MODCOD the learned offset / normalised
 *    min-sum with in-launch plain fallback (cusat_stream_rx in
 *    CUSAT_STREAM_LEARNED mode); the equality with rx_fec_pool above holds
 *    for the MODCODs without learned parameters;
 *  - HIP errors throw std::runtime_error.
 */
#ifndef CUSAT_AMD_DSP_RX_FEC_STREAM_POOL_H
#define CUSAT_AMD_DSP_RX_FEC_STREAM_POOL_H

#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "acm_config.h"
#include "cusat_fec.h"
#include "plframe.h"
#include "rx_fec_chain.h"

struct cusat_fec_ctx;
struct cusat_stream_ctx;
struct ihipStream_t;
struct ihipEvent_t;
struct hipGraphExec;

namespace cusat_dsp {

class rx_fec_stream_pool
{
public:
    /** An unknown MODCOD index throws std::invalid_argument. */
    rx_fec_stream_pool(fecframe_size_t framesize, const std::vector<uint8_t>& enabled_modcods, int max_trials = 25,
                       int cuda_max_batch = 64, int n_gpu_streams = 8, int micro_batch = 4);
    ~rx_fec_stream_pool();
    rx_fec_stream_pool(const rx_fec_stream_pool&) = delete;
    rx_fec_stream_pool& operator=(const rx_fec_stream_pool&) = delete;

    /** Queue a frame of MODCOD pls.modcod (ignored when not enabled: returns
     *  false).  Returns true when an output block is waiting. */
    bool submit(const complexf* symbols, int n_symbols, const pl_layout& pls);

    int get_gse_output(uint8_t* out, int max_bytes);
    const std::vector<int>& gse_dfl_sizes() const { return d_last_sizes; }

    void reset_all();
    void warmup_all();
    bool launch_stale(std::chrono::milliseconds max_age);
    bool drain_all_completed();
    void finish_all_pending();

    void set_diag_fec_bench(bool on) { d_diag[0] = on; }
    void set_diag_llr_histogram(bool on) { d_diag[1] = on; }
    void set_diag_ldpc_convergence(bool on) { d_diag[2] = on; }
    void set_cuda_mutex(std::mutex* mtx) { d_mtx = mtx; }
    void set_decode_budget(std::shared_ptr<gpu_decode_budget> b)
    {
        if (b) d_budget = std::move(b);
    }

    void update_snr(uint8_t modcod, float snr_db);

    int max_output_bytes() const { return d_max_output_bytes; }
    uint64_t total_frame_count() const;
    uint64_t total_error_count() const;
    bool modcod_counters(uint8_t modcod, uint64_t& frames, uint64_t& errors) const;
    uint64_t total_gpu_drop_count() const { return d_gpu_drop_cnt; }
    float data_snr_db() const;

    /** true: decision-directed SNR on the GPU; false (default): on the host,
     *  at submit, as a chain does (the faster of the two on a 12-MODCOD
     *  stream of 6144 frames: 128-151 k vs 112-126 k frames/s).  Set before
     *  the first submit. */
    void set_gpu_dd_snr(bool on) { d_gpu_dd = on; }
    /** Learned offset / normalised min-sum for one enabled MODCOD (no
     *  counterpart in rx_fec_pool): beta and alpha are int16 [n_iters][q]
     *  tables (q = layer_count(modcod)) of offsets in LLR units and scales in
     *  64ths (0 = no scale), as cusat_ldpc_set_params takes them.  Frames of
     *  that MODCOD launched afterwards decode n_iters learned iterations, and
     *  any frame failing frame_passes_fec_checks is decoded again by the plain
     *  decoder (max_trials) in the same launch; n_iters = 0 goes back to
     *  plain.  Waits for the micro-batches in flight.  Returns false for a
     *  MODCOD that is not enabled; bad tables throw std::runtime_error. */
    bool set_learned_params(uint8_t modcod, const int16_t* beta, const int16_t* alpha, int n_iters);
    /** Layers (check-row groups of 360) of an enabled MODCOD's code: the row
     *  length of the learned tables; -1 when not enabled. */
    int layer_count(uint8_t modcod) const;
    /** Frames whose learned pass was redone by the plain decoder (not in
     *  rx_fec_pool: it has no learned decoder). */
    uint64_t learned_fallback_count() const;
    /** Wait for the oldest micro-batch instead of dropping a frame when the
     *  ring is full (rx_fec_chain::set_block_when_full). */
    void set_block_when_full(bool on) { d_block_when_full = on; }
    int micro_batch() const { return d_micro_batch; }
    int pending_count() const { return d_batch_idx; }

private:
    static constexpr int N_FRAMES_BUFFERED = 512;

    struct modcod_state {
        uint8_t idx = 0;
        int code = 0;  // context index in the stream contexts
        modulation_t modulation = MOD_QPSK;
        int n_mod = 0, xfec = 0, max_dfl = 0, q = 0;
        std::vector<complexf> points;
        float snr_db = 0, pilot_snr_db = 0, noise_var = 1.0f, llr_gain = 4.0f;
        int snr_skip = 0;
        int host_event_batch = -1;  // micro-batch that carries the latest update_snr
        uint64_t frames = 0, errors = 0, fallbacks = 0;
    };
    struct flight_slot {
        complexf* syms = nullptr;  // pinned, frames at 64-byte aligned offsets
        // one pinned block, the layout of stream_res::blk: results, SNR state, data fields
        unsigned char* blk = nullptr;
        int32_t* res = nullptr;    // trials | valid | dfl, micro_batch each
        float* state = nullptr;    // SNR state per MODCOD after the micro-batch
        uint8_t* out = nullptr;    // data fields, packed
        ihipEvent_t* event = nullptr;
        std::vector<cusat_stream_rx_event> events;  // SNR events, submit order
        std::vector<int> mc;       // modcod_state index per frame
        std::vector<int64_t> sym_off, out_off;
        int64_t n_syms = 0, n_out = 0;
        int n_frames = 0;
        int batch = 0;             // micro-batch number (submit order)
        bool in_flight = false;
    };
    struct stream_res {
        ihipStream_t* stream = nullptr;
        cusat_stream_ctx* sc = nullptr;
        complexf* sym_dev = nullptr;
        // one device block: results | SNR state | data fields, copied back whole
        unsigned char* blk = nullptr;
        int32_t* res_dev = nullptr;
        float* state_dev = nullptr;      // state after the stream's latest micro-batch
        uint8_t* out_dev = nullptr;
        hipGraphExec* graph = nullptr;   // cusat_stream_rx
        hipGraphExec* graph_dd_est = nullptr;  // cusat_stream_rx_dd_estimate
        hipGraphExec* graph_dd = nullptr;      // cusat_stream_rx_dd_resolve
    };
    struct block {
        std::vector<uint8_t> bytes;
        std::vector<int> sizes;
        int mc = -1;
    };

    std::vector<cusat_fec_ctx*> d_ctx;       // one per enabled MODCOD (ascending index)
    std::vector<modcod_state> d_mc;          // same order
    std::map<uint8_t, int> d_by_idx;         // MODCOD index -> d_mc position
    std::vector<stream_res> d_streams;
    std::vector<flight_slot> d_flight;
    std::deque<block> d_ready;
    std::vector<int> d_last_sizes;
    int d_last_output = -1;                  // d_mc position of the latest output block
    int d_micro_batch = 4, d_n_flight = 0, d_max_xfec = 0, d_max_dfl = 0, d_max_output_bytes = 0;
    size_t d_res_bytes = 0, d_state_bytes = 0;  // offsets of the state and data fields in a block
    int d_batch_idx = 0, d_submit_idx = 0, d_launch_idx = 0, d_drain_idx = 0, d_next_stream = 0;
    uint64_t d_gpu_drop_cnt = 0;
    bool d_diag[3] = {false, false, false};
    std::chrono::steady_clock::time_point d_batch_start_time;
    std::mutex* d_mtx = nullptr;
    std::shared_ptr<gpu_decode_budget> d_budget;
    bool d_gpu_dd = false;
    bool d_block_when_full = false;
    float* d_state = nullptr;                 // device SNR state, float4 per MODCOD
    ihipEvent_t* d_dd_done = nullptr;         // recorded after the latest SNR launch
    bool d_dd_recorded = false;
    std::vector<cusat_stream_rx_event> d_events;  // events of the micro-batch being filled

    void update_dd_snr(modcod_state& m, const complexf* frame);
    void add_host_event(const modcod_state& m, float snr_lin, float snr_db);
    void reset_state();
    void close_batch();
    void pump_launches(bool bypass_cap = false);
    void launch_slot(flight_slot& s);
    bool drain(bool wait_all, bool oldest_only = false);
    void post_process(flight_slot& s, block*& cur);
    void release_budget(int n)
    {
        if (d_budget) d_budget->in_flight.fetch_sub(n, std::memory_order_relaxed);
    }
    void free_all();
};

}  // namespace cusat_dsp

#endif
