// GPU checks of cusat_dsp::rx_fec_pool (include/cusat_dsp/rx_fec_pool.h):
// an interleaved ACM stream of three MODCODs comes out in arrival order
// (a MODCOD change flushes the previous chain first, rx_fec_pool.cc:55-77);
// per-MODCOD and total counters; frames of a MODCOD the pool was not built
// for are ignored; launch_stale sends a partial micro-batch and
// drain_all_completed collects it (rx_fec_pool.cc:179-206); reset_all drops
// queued output; an unknown MODCOD index in the constructor throws.
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "cusat_dsp/rx_fec_pool.h"
#include "qa_traffic.h"
#include "qa_util.h"

using namespace cusat_dsp;

namespace {

struct pool_output {
    std::vector<uint8_t> bytes;
    std::vector<int> sizes;
    int blocks = 0;
    void take_all(rx_fec_pool& p)
    {
        std::vector<uint8_t> buf(p.max_output_bytes());
        for (;;) {
            const int n = p.get_gse_output(buf.data(), (int)buf.size());
            if (n <= 0) break;
            blocks++;
            bytes.insert(bytes.end(), buf.begin(), buf.begin() + n);
            sizes.insert(sizes.end(), p.gse_dfl_sizes().begin(), p.gse_dfl_sizes().end());
        }
    }
};

pl_layout layout_of(uint8_t modcod)
{
    pl_layout l;
    l.modcod = modcod;
    return l;
}

}  // namespace

int main()
{
    return qa_run({
        {"acm_stream_in_arrival_order",
         [] {
             // MODCOD 4 = QPSK 1/2, 14 = 8PSK 3/4, 18 = 16APSK 2/3 (normal frames)
             const uint8_t ids[3] = {4, 14, 18};
             const float es[3] = {4.0f, 10.0f, 12.0f};
             std::vector<traffic> t;
             for (int k = 0; k < 3; k++) {
                 const modcod_def_t* m = find_modcod(ids[k]);
                 QA_CHECK(m != nullptr);
                 t.push_back(make_traffic(FECFRAME_NORMAL, m->code_rate, m->constellation, 12, es[k], 40 + k));
             }
             // runs of (MODCOD slot, frames): changes inside partial micro-batches
             const int runs[][2] = {{0, 5}, {1, 3}, {0, 2}, {2, 6}, {1, 4}, {2, 1}, {0, 5}, {2, 5}, {1, 5}};
             rx_fec_pool pool(FECFRAME_NORMAL, {4, 14, 18}, 50, 64, 3, 4);
             for (int k = 0; k < 3; k++) pool.update_snr(ids[k], es[k]);
             pool_output got;
             std::vector<uint8_t> want;
             std::vector<int> want_sizes;
             int next[3] = {0, 0, 0};
             for (const auto& r : runs)
                 for (int j = 0; j < r[1]; j++) {
                     const int k = r[0], b = next[k]++;
                     const auto& p = t[k].payloads[b];
                     want.insert(want.end(), p.begin(), p.end());
                     if (!p.empty()) want_sizes.push_back((int)p.size());
                     if (pool.submit(&t[k].symbols[(size_t)b * t[k].xfec], t[k].xfec, layout_of(ids[k])))
                         got.take_all(pool);
                 }
             pool.finish_all_pending();
             got.take_all(pool);
             QA_CHECK_EQ(got.bytes.size(), want.size());
             QA_CHECK(got.bytes == want);
             QA_CHECK(got.sizes == want_sizes);
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)36);
             QA_CHECK_EQ(pool.total_error_count(), (uint64_t)0);
             QA_CHECK_EQ(pool.total_gpu_drop_count(), (uint64_t)0);
             for (int k = 0; k < 3; k++) {
                 uint64_t f = 0, e = 0;
                 QA_CHECK(pool.modcod_counters(ids[k], f, e));
                 QA_CHECK_EQ(f, (uint64_t)12);
                 QA_CHECK_EQ(e, (uint64_t)0);
             }
             uint64_t f = 0, e = 0;
             QA_CHECK(!pool.modcod_counters(7, f, e));
             QA_CHECK(pool.data_snr_db() > 0.0f);
         }},
        {"unknown_modcod_frames_are_ignored",
         [] {
             const modcod_def_t* m = find_modcod(4);
             traffic t = make_traffic(FECFRAME_SHORT, m->code_rate, m->constellation, 3, 5.0f, 50);
             rx_fec_pool pool(FECFRAME_SHORT, {4}, 50, 64, 2, 4);
             pool.update_snr(4, 5.0f);
             QA_CHECK(!pool.submit(&t.symbols[0], t.xfec, layout_of(14)));  // not enabled
             for (int b = 0; b < 3; b++) pool.submit(&t.symbols[(size_t)b * t.xfec], t.xfec, layout_of(4));
             pool.finish_all_pending();
             pool_output got;
             got.take_all(pool);
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)3);
             std::vector<uint8_t> want;
             for (const auto& p : t.payloads) want.insert(want.end(), p.begin(), p.end());
             QA_CHECK(got.bytes == want);
         }},
        {"launch_stale_then_drain",
         [] {
             const modcod_def_t* m = find_modcod(14);
             traffic t = make_traffic(FECFRAME_NORMAL, m->code_rate, m->constellation, 1, 10.0f, 51);
             rx_fec_pool pool(FECFRAME_NORMAL, {14, 4}, 50, 64, 2, 4);
             pool.update_snr(14, 10.0f);
             QA_CHECK(!pool.submit(&t.symbols[0], t.xfec, layout_of(14)));  // micro-batch of 4 not full
             QA_CHECK(!pool.launch_stale(std::chrono::milliseconds(60000)));
             QA_CHECK(pool.launch_stale(std::chrono::milliseconds(0)));
             pool_output got;
             for (int tries = 0; tries < 2000 && got.blocks == 0; tries++) {
                 if (pool.drain_all_completed()) got.take_all(pool);
                 else std::this_thread::sleep_for(std::chrono::milliseconds(1));
             }
             QA_CHECK_EQ(got.blocks, 1);
             QA_CHECK(got.bytes == t.payloads[0]);
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)1);
         }},
        {"reset_all_drops_queued_output",
         [] {
             const modcod_def_t* m = find_modcod(4);
             traffic t = make_traffic(FECFRAME_SHORT, m->code_rate, m->constellation, 4, 5.0f, 52);
             rx_fec_pool pool(FECFRAME_SHORT, {4}, 50, 64, 2, 4);
             pool.update_snr(4, 5.0f);
             bool ready = false;
             for (int b = 0; b < 4; b++) ready |= pool.submit(&t.symbols[(size_t)b * t.xfec], t.xfec, layout_of(4));
             pool.finish_all_pending();
             std::vector<uint8_t> buf(pool.max_output_bytes());
             pool.reset_all();
             QA_CHECK_EQ(pool.get_gse_output(buf.data(), (int)buf.size()), 0);
             QA_CHECK(pool.gse_dfl_sizes().empty());
         }},
        {"unknown_index_in_constructor_throws",
         [] {
             bool threw = false;
             try {
                 rx_fec_pool pool(FECFRAME_NORMAL, {14, 200});
             } catch (const std::invalid_argument&) {
                 threw = true;
             }
             QA_CHECK(threw);
         }},
    });
}
