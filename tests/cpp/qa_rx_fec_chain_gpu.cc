// GPU checks of the C++ cusat_dsp API: GSE payloads survive the whole
// chain (BBHEADER + BB scrambling + BCH + LDPC + mapping on the host
// transmitter, rx_fec_chain on the GPU) in submission order; wrong-length
// frames and a full flight ring are counted as errors and dropped; warmup
// leaves clean counters; a shared decode budget with block-when-full
// delivers every frame; ldpc_decode decodes host and device LLRs.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <random>

#include "cusat_dsp/ldpc_decode.h"
#include "cusat_dsp/rx_fec_chain.h"
#include "cusat_fec.h"
#include "qa_traffic.h"
#include "qa_util.h"

using namespace cusat_dsp;

namespace {

struct collected {
    std::vector<uint8_t> bytes;
    std::vector<int> sizes;
    void take(rx_fec_chain& c)
    {
        std::vector<uint8_t> buf(c.max_output_bytes());
        const int n = c.get_gse_output(buf.data(), (int)buf.size());
        bytes.insert(bytes.end(), buf.begin(), buf.begin() + n);
        sizes.insert(sizes.end(), c.gse_dfl_sizes().begin(), c.gse_dfl_sizes().end());
    }
};

void expect_payloads(const collected& got, const traffic& t, const std::vector<int>& which)
{
    std::vector<uint8_t> want;
    std::vector<int> sizes;
    for (int b : which) {
        want.insert(want.end(), t.payloads[b].begin(), t.payloads[b].end());
        if (!t.payloads[b].empty()) sizes.push_back((int)t.payloads[b].size());
    }
    QA_CHECK_EQ(got.bytes.size(), want.size());
    QA_CHECK(got.bytes == want);
    QA_CHECK(got.sizes == sizes);
}

void run_chain(rx_fec_chain& c, const traffic& t, int n, collected& out, int n_symbols_override = -1)
{
    for (int b = 0; b < n; b++) {
        const int ns = n_symbols_override >= 0 && b == 0 ? n_symbols_override : t.xfec;
        if (c.submit_xfecframe(&t.symbols[(size_t)b * t.xfec], ns)) out.take(c);
    }
    c.flush();
    out.take(c);
}

std::vector<int> all(int n)
{
    std::vector<int> v(n);
    for (int i = 0; i < n; i++) v[i] = i;
    return v;
}

}  // namespace

int main()
{
    return qa_run({
        {"gse_roundtrip_8psk34",
         [] {
             const int n = 40;
             traffic t = make_traffic(FECFRAME_NORMAL, RATE_3_4, MOD_8PSK, n, 10.0f, 11);
             rx_fec_chain c(FECFRAME_NORMAL, RATE_3_4, MOD_8PSK, 50, 64, 3, 4);
             c.update_snr(10.0f);
             collected got;
             run_chain(c, t, n, got);
             QA_CHECK_EQ(c.frame_count(), (uint64_t)n);
             QA_CHECK_EQ(c.frame_error_count(), (uint64_t)0);
             QA_CHECK_EQ(c.gpu_drop_count(), (uint64_t)0);
             expect_payloads(got, t, all(n));
             QA_CHECK(c.snr_db() > 8.0f && c.snr_db() < 12.0f);  // decision-directed re-estimate
         }},
        {"gse_roundtrip_short_16apsk",
         [] {
             const int n = 17;  // a partial last micro-batch
             traffic t = make_traffic(FECFRAME_SHORT, RATE_2_3, MOD_16APSK, n, 13.0f, 12);
             rx_fec_chain c(FECFRAME_SHORT, RATE_2_3, MOD_16APSK, 50, 64, 2, 4);
             c.update_snr(13.0f);
             collected got;
             run_chain(c, t, n, got);
             QA_CHECK_EQ(c.frame_count(), (uint64_t)n);
             QA_CHECK_EQ(c.frame_error_count(), (uint64_t)0);
             expect_payloads(got, t, all(n));
         }},
        {"wrong_length_frame_is_dropped",
         [] {
             const int n = 6;
             traffic t = make_traffic(FECFRAME_NORMAL, RATE_1_2, MOD_QPSK, n, 6.0f, 13);
             rx_fec_chain c(FECFRAME_NORMAL, RATE_1_2, MOD_QPSK, 50, 64, 2, 4);
             c.update_snr(6.0f);
             collected got;
             run_chain(c, t, n, got, t.xfec - 1);
             QA_CHECK_EQ(c.frame_count(), (uint64_t)n);
             QA_CHECK_EQ(c.frame_error_count(), (uint64_t)1);
             QA_CHECK_EQ(c.gpu_drop_count(), (uint64_t)1);
             std::vector<int> kept = all(n);
             kept.erase(kept.begin());
             expect_payloads(got, t, kept);
         }},
        {"noise_frames_fail_fec",
         [] {
             const int n = 8;
             traffic t = make_traffic(FECFRAME_NORMAL, RATE_9_10, MOD_8PSK, n, 5.0f, 14);  // far below threshold
             rx_fec_chain c(FECFRAME_NORMAL, RATE_9_10, MOD_8PSK, 10, 64, 2, 4);
             c.update_snr(5.0f);
             collected got;
             run_chain(c, t, n, got);
             QA_CHECK_EQ(c.frame_count(), (uint64_t)n);
             QA_CHECK_EQ(c.frame_error_count(), (uint64_t)n);
             QA_CHECK(got.bytes.empty());
         }},
        {"warmup_leaves_clean_counters",
         [] {
             rx_fec_chain c(FECFRAME_SHORT, RATE_1_2, MOD_QPSK, 25, 64, 2, 4);
             c.warmup();
             QA_CHECK_EQ(c.frame_count(), (uint64_t)0);
             QA_CHECK_EQ(c.frame_error_count(), (uint64_t)0);
             QA_CHECK_EQ(c.gpu_drop_count(), (uint64_t)0);
             QA_CHECK_EQ(c.force_drain_count(), (uint64_t)0);
             QA_CHECK_EQ(c.pending_count(), 0);
             QA_CHECK_EQ(c.gpu_inflight(), 0);
         }},
        {"shared_budget_block_when_full_delivers_all",
         [] {
             const int n = 300;  // > 126 one-frame slots: the ring fills
             traffic t = make_traffic(FECFRAME_SHORT, RATE_1_2, MOD_QPSK, n, 6.0f, 15);
             auto budget = std::make_shared<gpu_decode_budget>();
             budget->cap = 2;
             rx_fec_chain c(FECFRAME_SHORT, RATE_1_2, MOD_QPSK, 50, 64, 4, 1);
             c.set_decode_budget(budget);
             c.set_block_when_full(true);
             c.update_snr(6.0f);
             collected got;
             run_chain(c, t, n, got);
             QA_CHECK_EQ(c.frame_count(), (uint64_t)n);
             QA_CHECK_EQ(c.frame_error_count(), (uint64_t)0);
             QA_CHECK_EQ(c.gpu_drop_count(), (uint64_t)0);
             QA_CHECK(c.force_drain_count() > 0);
             QA_CHECK_EQ(budget->in_flight.load(), 0);
             expect_payloads(got, t, all(n));
         }},
        {"full_ring_drops_without_blocking",
         [] {
             const int n = 200;
             traffic t = make_traffic(FECFRAME_SHORT, RATE_1_2, MOD_QPSK, n, 6.0f, 16);
             auto budget = std::make_shared<gpu_decode_budget>();
             budget->cap = 1;
             rx_fec_chain c(FECFRAME_SHORT, RATE_1_2, MOD_QPSK, 50, 64, 2, 1);
             c.set_decode_budget(budget);
             c.update_snr(6.0f);
             collected got;
             run_chain(c, t, n, got);
             QA_CHECK_EQ(c.frame_count(), (uint64_t)n);
             QA_CHECK(c.gpu_drop_count() > 0);
             QA_CHECK_EQ(c.frame_error_count(), c.gpu_drop_count());
             QA_CHECK_EQ(budget->in_flight.load(), 0);
         }},
        {"ldpc_decode_host_and_device_llrs",
         [] {
             cusat_fec_info inf;
             cusat_fec_query(FECFRAME_SHORT, RATE_1_2, CUSAT_MOD_QPSK, 0, &inf);
             const int B = 3, N = inf.n_ldpc, K = inf.k_ldpc;
             std::vector<float> sym((size_t)B * inf.xfec_len * 2);
             std::vector<uint8_t> cw((size_t)B * N / 8);
             cusat_gen_frames(FECFRAME_SHORT, RATE_1_2, CUSAT_MOD_QPSK, 5, B, 1000.0f, sym.data(), nullptr,
                              cw.data());
             std::vector<int16_t> llr((size_t)B * N);
             for (int b = 0; b < B; b++)
                 for (int v = 0; v < N; v++)
                     llr[(size_t)b * N + v] = ((cw[(size_t)b * N / 8 + v / 8] >> (7 - v % 8)) & 1) ? -100 : 100;
             ldpc_decode d(FECFRAME_SHORT, RATE_1_2, 25, 8);
             QA_CHECK_EQ(d.n_bits(), (unsigned)N);
             QA_CHECK_EQ(d.k_bytes(), (unsigned)K / 8);
             std::vector<uint8_t> hard((size_t)B * K / 8, 0xaa);
             std::vector<int> trials(B, -7);
             d.decode_batch_cuda(llr.data(), hard.data(), trials.data(), B);
             for (int b = 0; b < B; b++) {
                 QA_CHECK_EQ(trials[b], 25);
                 QA_CHECK(memcmp(&hard[(size_t)b * K / 8], &cw[(size_t)b * N / 8], K / 8) == 0);
             }
             int16_t* dev = nullptr;
             QA_CHECK(hipMalloc(&dev, llr.size() * 2) == hipSuccess);
             QA_CHECK(hipMemcpy(dev, llr.data(), llr.size() * 2, hipMemcpyHostToDevice) == hipSuccess);
             std::fill(hard.begin(), hard.end(), 0x55);
             std::fill(trials.begin(), trials.end(), -7);
             d.decode_batch_cuda_d2h(dev, hard.data(), trials.data(), B);
             for (int b = 0; b < B; b++) {
                 QA_CHECK_EQ(trials[b], 25);
                 QA_CHECK(memcmp(&hard[(size_t)b * K / 8], &cw[(size_t)b * N / 8], K / 8) == 0);
             }
             // a batch outside 1..max returns without decoding (ldpc_decode_kernel.cu:1291-1292)
             std::fill(trials.begin(), trials.end(), -7);
             d.decode_batch_cuda(llr.data(), hard.data(), trials.data(), 9);
             QA_CHECK_EQ(trials[0], -7);
             hipFree(dev);
         }},
    });
}
