// rx_fec_pool and the MODCOD table of the cusat_dsp C++ API
// (include/cusat_dsp/rx_fec_pool.h, include/cusat_dsp/acm_config.h).
#include "cusat_dsp/rx_fec_pool.h"

#include <algorithm>
#include <stdexcept>
#include <string>

#include "cusat_fec.h"

namespace cusat_dsp {

// ---- MODCOD table (lib/cusat_dsp/acm_config.h:50-117) ----

const std::vector<modcod_def_t>& get_supported_modcods()
{
    static const std::vector<modcod_def_t> t = {
        {1, MOD_QPSK, RATE_1_4, "qpsk1/4", 0.49f, -0.05f},
        {33, MOD_QPSK, RATE_13_45, "qpsk13/45", 0.57f, 0.07f},
        {2, MOD_QPSK, RATE_1_3, "qpsk1/3", 0.66f, 0.56f},
        {3, MOD_QPSK, RATE_2_5, "qpsk2/5", 0.79f, 1.00f},
        {43, MOD_QPSK, RATE_9_20, "qpsk9/20", 0.89f, 1.42f},
        {4, MOD_QPSK, RATE_1_2, "qpsk1/2", 0.99f, 1.80f},
        {44, MOD_QPSK, RATE_11_20, "qpsk11/20", 1.09f, 2.25f},
        {5, MOD_QPSK, RATE_3_5, "qpsk3/5", 1.19f, 2.73f},
        {6, MOD_QPSK, RATE_2_3, "qpsk2/3", 1.32f, 3.50f},
        {7, MOD_QPSK, RATE_3_4, "qpsk3/4", 1.49f, 4.43f},
        {8, MOD_QPSK, RATE_4_5, "qpsk4/5", 1.59f, 5.08f},
        {9, MOD_QPSK, RATE_5_6, "qpsk5/6", 1.65f, 5.48f},
        {10, MOD_QPSK, RATE_8_9, "qpsk8/9", 1.77f, 6.50f},
        {11, MOD_QPSK, RATE_9_10, "qpsk9/10", 1.79f, 6.72f},
        {56, MOD_8APSK, RATE_100_180, "8apsk5/9", 1.65f, 4.93f},
        {57, MOD_8APSK, RATE_104_180, "8apsk26/45", 1.71f, 5.33f},
        {12, MOD_8PSK, RATE_3_5, "8psk3/5", 1.78f, 5.80f},
        {45, MOD_8PSK, RATE_23_36, "8psk23/36", 1.90f, 7.22f},
        {13, MOD_8PSK, RATE_2_3, "8psk2/3", 1.98f, 6.92f},
        {46, MOD_8PSK, RATE_25_36, "8psk25/36", 2.06f, 7.22f},
        {47, MOD_8PSK, RATE_13_18, "8psk13/18", 2.15f, 7.69f},
        {14, MOD_8PSK, RATE_3_4, "8psk3/4", 2.23f, 8.11f},
        {15, MOD_8PSK, RATE_5_6, "8psk5/6", 2.48f, 9.55f},
        {16, MOD_8PSK, RATE_8_9, "8psk8/9", 2.65f, 10.79f},
        {17, MOD_8PSK, RATE_9_10, "8psk9/10", 2.68f, 11.08f},
        {58, MOD_16APSK, RATE_90_180, "16apsk1/2l", 1.97f, 6.27f},
        {59, MOD_16APSK, RATE_96_180, "16apsk8/15l", 2.10f, 6.85f},
        {60, MOD_16APSK, RATE_100_180, "16apsk5/9l", 2.19f, 7.04f},
        {61, MOD_16APSK, RATE_18_30, "16apsk3/5l", 2.37f, 7.71f},
        {62, MOD_16APSK, RATE_20_30, "16apsk2/3l", 2.64f, 8.73f},
        {48, MOD_16APSK, RATE_26_45, "16apsk26/45", 2.28f, 7.81f},
        {49, MOD_16APSK, RATE_3_5, "16apsk3/5", 2.37f, 8.10f},
        {50, MOD_16APSK, RATE_28_45, "16apsk28/45", 2.46f, 8.30f},
        {51, MOD_16APSK, RATE_23_36, "16apsk23/36", 2.52f, 8.58f},
        {18, MOD_16APSK, RATE_2_3, "16apsk2/3", 2.64f, 9.17f},
        {52, MOD_16APSK, RATE_25_36, "16apsk25/36", 2.75f, 9.47f},
        {53, MOD_16APSK, RATE_13_18, "16apsk13/18", 2.86f, 9.91f},
        {19, MOD_16APSK, RATE_3_4, "16apsk3/4", 2.97f, 10.41f},
        {54, MOD_16APSK, RATE_140_180, "16apsk7/9", 3.08f, 10.85f},
        {20, MOD_16APSK, RATE_4_5, "16apsk4/5", 3.17f, 11.13f},
        {21, MOD_16APSK, RATE_5_6, "16apsk5/6", 3.30f, 11.71f},
        {55, MOD_16APSK, RATE_154_180, "16apsk77/90", 3.39f, 12.19f},
        {22, MOD_16APSK, RATE_8_9, "16apsk8/9", 3.52f, 12.99f},
        {23, MOD_16APSK, RATE_9_10, "16apsk9/10", 3.57f, 13.23f},
        {63, MOD_32APSK, RATE_2_3, "32apsk2/3l", 3.29f, 11.30f},
        {64, MOD_32APSK, RATE_128_180, "32apsk32/45", 3.51f, 12.05f},
        {65, MOD_32APSK, RATE_132_180, "32apsk11/15", 3.62f, 12.47f},
        {24, MOD_32APSK, RATE_3_4, "32apsk3/4", 3.70f, 12.93f},
        {66, MOD_32APSK, RATE_140_180, "32apsk7/9", 3.84f, 13.35f},
        {25, MOD_32APSK, RATE_4_5, "32apsk4/5", 3.95f, 13.84f},
        {26, MOD_32APSK, RATE_5_6, "32apsk5/6", 4.12f, 14.48f},
        {27, MOD_32APSK, RATE_8_9, "32apsk8/9", 4.40f, 15.79f},
        {28, MOD_32APSK, RATE_9_10, "32apsk9/10", 4.47f, 16.05f},
        {29, MOD_64APSK, RATE_132_180, "64apsk11/15", 4.40f, 15.00f},
        {30, MOD_64APSK, RATE_7_9, "64apsk7/9", 4.67f, 15.60f},
        {31, MOD_64APSK, RATE_4_5, "64apsk4/5", 4.80f, 16.10f},
        {32, MOD_128APSK, RATE_135_180, "128apsk3/4", 5.25f, 18.00f},
        {34, MOD_128APSK, RATE_140_180, "128apsk7/9", 5.44f, 18.89f},
        {35, MOD_64APSK, RATE_5_6, "64apsk5/6", 4.98f, 17.23f},
        {36, MOD_64APSK, RATE_128_180, "64apsk32/45", 4.27f, 14.18f},
        {37, MOD_256APSK, RATE_135_180, "256apsk3/4", 6.00f, 23.34f},
        {38, MOD_256APSK, RATE_128_180, "256apsk32/45", 5.69f, 22.40f},
        {39, MOD_256APSK, RATE_116_180, "256apsk29/45", 5.16f, 20.20f},
        {40, MOD_256APSK, RATE_124_180, "256apsk31/45", 5.51f, 21.50f},
        {41, MOD_256APSK, RATE_20_30, "256apsk2/3", 5.06f, 18.10f},
        {42, MOD_256APSK, RATE_22_30, "256apsk11/15", 5.50f, 19.80f},
    };
    return t;
}

const modcod_def_t* find_modcod(uint8_t modcod_idx)
{
    const auto& t = get_supported_modcods();
    auto it = std::find_if(t.begin(), t.end(), [&](const modcod_def_t& m) { return m.modcod_idx == modcod_idx; });
    return it == t.end() ? nullptr : &*it;
}

const modcod_def_t* find_modcod_by_name(const char* name)
{
    if (!name) return nullptr;
    const auto& t = get_supported_modcods();
    auto it = std::find_if(t.begin(), t.end(), [&](const modcod_def_t& m) { return std::strcmp(m.name, name) == 0; });
    return it == t.end() ? nullptr : &*it;
}

bool has_fec_params(fecframe_size_t framesize, code_rate_t rate)
{
    cusat_fec_info info;
    return cusat_fec_query((int)framesize, (int)rate, CUSAT_MOD_QPSK, 0, &info) == CUSAT_OK;
}

// ---- rx_fec_pool (rx_fec_pool.cc:15-295) ----

rx_fec_pool::rx_fec_pool(fecframe_size_t framesize, const std::vector<uint8_t>& enabled_modcods, int max_trials,
                         int cuda_max_batch, int n_gpu_streams, int micro_batch)
{
    auto budget = gpu_decode_budget::make_default();
    for (uint8_t idx : enabled_modcods) {
        if (d_chains.count(idx)) continue;
        const modcod_def_t* m = find_modcod(idx);
        if (!m) throw std::invalid_argument("Unknown modcod_idx: " + std::to_string(idx));
        auto chain = std::make_unique<rx_fec_chain>(framesize, m->code_rate, m->constellation, max_trials,
                                                    cuda_max_batch, n_gpu_streams, micro_batch);
        chain->set_decode_budget(budget);
        d_max_output_bytes = std::max(d_max_output_bytes, chain->max_output_bytes());
        d_chains.emplace(idx, std::move(chain));
    }
}

bool rx_fec_pool::collect(uint8_t modcod, rx_fec_chain& chain)
{
    block b;
    if (!d_spare.empty()) {
        b = std::move(d_spare.back());
        d_spare.pop_back();
    }
    b.bytes.resize(std::max(chain.max_output_bytes(), 1));
    const int n = chain.get_gse_output(b.bytes.data(), (int)b.bytes.size());
    if (n <= 0) {
        b.bytes.clear();
        b.sizes.clear();
        d_spare.push_back(std::move(b));
        return false;
    }
    b.bytes.resize(n);
    b.sizes = chain.gse_dfl_sizes();
    b.modcod = modcod;
    d_last_output = modcod;
    d_ready.push_back(std::move(b));
    return true;
}

bool rx_fec_pool::submit(const complexf* symbols, int n_symbols, const pl_layout& pls)
{
    auto it = d_chains.find(pls.modcod);
    if (it == d_chains.end()) return false;
    bool got = false;
    // MODCOD change: finish the previous MODCOD's frames first, so that its
    // output is queued before anything of the new run
    if (d_have_last_submit && pls.modcod != d_last_submit) {
        auto prev = d_chains.find(d_last_submit);
        if (prev != d_chains.end()) {
            rx_fec_chain& pc = *prev->second;
            if (pc.pending_count() > 0 || pc.gpu_inflight() > 0 || pc.gpu_pending_launch() > 0) {
                pc.flush();
                got |= collect(prev->first, pc);
            }
        }
    }
    d_last_submit = pls.modcod;
    d_have_last_submit = true;
    if (it->second->submit_xfecframe(symbols, n_symbols)) got |= collect(it->first, *it->second);
    return got || !d_ready.empty();
}

int rx_fec_pool::get_gse_output(uint8_t* out, int max_bytes)
{
    d_last_sizes.clear();
    if (d_ready.empty()) return 0;
    block b = std::move(d_ready.front());
    d_ready.pop_front();
    const int n = std::min<int>((int)b.bytes.size(), std::max(max_bytes, 0));
    if (n > 0) std::copy_n(b.bytes.data(), n, out);
    d_last_output = b.modcod;
    d_last_sizes.swap(b.sizes);
    b.bytes.clear();
    b.sizes.clear();
    d_spare.push_back(std::move(b));
    return n;
}

void rx_fec_pool::drop_ready()
{
    while (!d_ready.empty()) {
        block b = std::move(d_ready.front());
        d_ready.pop_front();
        b.bytes.clear();
        b.sizes.clear();
        d_spare.push_back(std::move(b));
    }
    d_last_sizes.clear();
}

void rx_fec_pool::forget_order()
{
    d_have_last_submit = false;
    d_last_submit = 0;
    d_last_output = 0;
    drop_ready();
}

void rx_fec_pool::reset_all()
{
    for (auto& kv : d_chains) kv.second->reset();
    forget_order();
}

void rx_fec_pool::warmup_all()
{
    for (auto& kv : d_chains) kv.second->warmup();
    forget_order();
}

bool rx_fec_pool::launch_stale(std::chrono::milliseconds max_age)
{
    const auto now = std::chrono::steady_clock::now();
    bool launched = false;
    for (auto& kv : d_chains) {
        rx_fec_chain& c = *kv.second;
        if (c.pending_count() > 0 && now - c.batch_start_time() >= max_age) {
            c.launch_partial();
            launched = true;
        }
    }
    return launched;
}

bool rx_fec_pool::drain_all_completed()
{
    bool got = false;
    for (auto& kv : d_chains)
        if (kv.second->drain_completed()) got |= collect(kv.first, *kv.second);
    return got || !d_ready.empty();
}

void rx_fec_pool::finish_all_pending()
{
    for (auto& kv : d_chains) {
        kv.second->flush();
        collect(kv.first, *kv.second);
    }
}

void rx_fec_pool::set_diag_fec_bench(bool on)
{
    for (auto& kv : d_chains) kv.second->set_diag_fec_bench(on);
}

void rx_fec_pool::set_diag_llr_histogram(bool on)
{
    for (auto& kv : d_chains) kv.second->set_diag_llr_histogram(on);
}

void rx_fec_pool::set_diag_ldpc_convergence(bool on)
{
    for (auto& kv : d_chains) kv.second->set_diag_ldpc_convergence(on);
}

void rx_fec_pool::set_cuda_mutex(std::mutex* mtx)
{
    for (auto& kv : d_chains) kv.second->set_cuda_mutex(mtx);
}

void rx_fec_pool::update_snr(uint8_t modcod, float snr_db)
{
    auto it = d_chains.find(modcod);
    if (it != d_chains.end()) it->second->update_snr(snr_db);
}

uint64_t rx_fec_pool::total_frame_count() const
{
    uint64_t n = 0;
    for (const auto& kv : d_chains) n += kv.second->frame_count();
    return n;
}

uint64_t rx_fec_pool::total_error_count() const
{
    uint64_t n = 0;
    for (const auto& kv : d_chains) n += kv.second->frame_error_count();
    return n;
}

bool rx_fec_pool::modcod_counters(uint8_t modcod, uint64_t& frames, uint64_t& errors) const
{
    auto it = d_chains.find(modcod);
    if (it == d_chains.end()) return false;
    frames = it->second->frame_count();
    errors = it->second->frame_error_count();
    return true;
}

uint64_t rx_fec_pool::total_gpu_drop_count() const
{
    uint64_t n = 0;
    for (const auto& kv : d_chains) n += kv.second->gpu_drop_count();
    return n;
}

float rx_fec_pool::data_snr_db() const
{
    auto it = d_chains.find(d_last_output);
    return it == d_chains.end() ? 0.0f : it->second->data_snr_db();
}

}  // namespace cusat_dsp
