/*
 * RX FEC pool for ACM / VCM reception: one rx_fec_chain per enabled MODCOD,
 * frames routed by the MODCOD index of their PL layout, with the public C++
 * API of lib/cusat_dsp/rx_fec_pool.h:26-156 (same constructor, methods and
 * counters).
 *
 * Behaviour kept from the reference (rx_fec_pool.cc:15-295):
 *  - every chain draws on ONE gpu_decode_budget, so the in-flight cap
 *    bounds the decode work of all MODCODs together;
 *  - a change of MODCOD between two submitted frames is an ordering
 *    barrier: the previous MODCOD's chain is flushed (its frames decoded
 *    and post-processed) before the new frame is queued, so output blocks
 *    leave the pool in the order their frames arrived;
 *  - every completed output block is copied at once into a pool-owned FIFO
 *    (a chain's own output buffer only lives until its next poll), and
 *    get_gse_output hands the blocks out one at a time;
 *  - frames of a MODCOD index the pool was not built for are ignored.
 * Chains are kept in ascending MODCOD order, so polls, flushes and counters
 * visit them in a fixed order.
 */
#ifndef CUSAT_AMD_DSP_RX_FEC_POOL_H
#define CUSAT_AMD_DSP_RX_FEC_POOL_H

#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "acm_config.h"
#include "plframe.h"
#include "rx_fec_chain.h"

namespace cusat_dsp {

class rx_fec_pool
{
public:
    /** One chain per enabled MODCOD index; an unknown index throws
     *  std::invalid_argument (as the reference). */
    rx_fec_pool(fecframe_size_t framesize, const std::vector<uint8_t>& enabled_modcods, int max_trials = 25,
                int cuda_max_batch = 64, int n_gpu_streams = 8, int micro_batch = 4);

    /** Route a frame to the chain of pls.modcod.  Returns true when an
     *  output block is waiting in the pool (get_gse_output). */
    bool submit(const complexf* symbols, int n_symbols, const pl_layout& pls);

    /** The oldest waiting output block (at most max_bytes of it); 0 when none. */
    int get_gse_output(uint8_t* out, int max_bytes);
    /** Data-field sizes of the block last returned by get_gse_output. */
    const std::vector<int>& gse_dfl_sizes() const { return d_last_sizes; }

    void reset_all();
    void warmup_all();

    /** Launch every chain's partial micro-batch older than max_age; no
     *  waiting (the results are collected by drain_all_completed). */
    bool launch_stale(std::chrono::milliseconds max_age);
    /** Non-blocking: collect every chain's completed work. */
    bool drain_all_completed();
    /** Blocking: flush every chain and collect all output. */
    void finish_all_pending();

    void set_diag_fec_bench(bool on);
    void set_diag_llr_histogram(bool on);
    void set_diag_ldpc_convergence(bool on);
    void set_cuda_mutex(std::mutex* mtx);

    /** Demapper SNR of one MODCOD's chain. */
    void update_snr(uint8_t modcod, float snr_db);

    int max_output_bytes() const { return d_max_output_bytes; }
    uint64_t total_frame_count() const;
    uint64_t total_error_count() const;
    bool modcod_counters(uint8_t modcod, uint64_t& frames, uint64_t& errors) const;
    uint64_t total_gpu_drop_count() const;
    /** Data SNR of the chain that produced the latest output. */
    float data_snr_db() const;

private:
    struct block {
        std::vector<uint8_t> bytes;
        std::vector<int> sizes;
        uint8_t modcod = 0;
    };

    std::map<uint8_t, std::unique_ptr<rx_fec_chain>> d_chains;
    int d_max_output_bytes = 0;
    bool d_have_last_submit = false;
    uint8_t d_last_submit = 0;  // MODCOD of the previous submitted frame
    uint8_t d_last_output = 0;  // MODCOD of the latest output
    std::deque<block> d_ready;
    std::vector<block> d_spare;  // emptied blocks kept for their capacity
    std::vector<int> d_last_sizes;

    bool collect(uint8_t modcod, rx_fec_chain& chain);
    void drop_ready();
    void forget_order();
};

}  // namespace cusat_dsp

#endif
