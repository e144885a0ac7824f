// GPU checks of cusat_dsp::rx_fec_stream_pool (include/cusat_dsp/rx_fec_stream_pool.h)
// against cusat_dsp::rx_fec_pool, the per-MODCOD-chain pool of the
// reference's design (rx_fec_pool.cc:15-295):
//  - an interleaved ACM stream (four MODCODs, one of them at its waterfall so
//    that some frames fail, one far below threshold, a frame of the wrong
//    length, frames of a MODCOD the pools were not built for) gives the same
//    output bytes, data-field sizes and per-MODCOD frame / error counters from
//    both pools, for several micro-batch sizes and stream counts; the payloads
//    of the frames that pass come out in arrival order;
//  - launch_stale sends a partial micro-batch and drain_all_completed
//    collects it; reset_all drops queued output; warmup_all leaves clean
//    counters; a full ring drops frames (counted as errors) while the rest
//    still leave in arrival order, or blocks when set_block_when_full is on;
//    an unknown MODCOD index in the constructor throws;
//  - learned parameters per MODCOD (set_learned_params): tables that restate
//    the plain check node give the plain pool's output, tables that stop
//    every learned pass send each frame to the in-launch plain fallback
//    (learned_fallback_count) and still give it.
// Fewer than 32 frames of each MODCOD are sent, so the decision-directed SNR
// re-estimate (every 32nd frame of a MODCOD) never fires and both pools scale
// every frame by the update_snr value.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <thread>

#include "cusat_dsp/rx_fec_pool.h"
#include "cusat_dsp/rx_fec_stream_pool.h"
#include "qa_traffic.h"
#include "qa_util.h"

using namespace cusat_dsp;

namespace {

struct pool_output {
    std::vector<uint8_t> bytes;
    std::vector<int> sizes;
    int blocks = 0;
    template <class P>
    void take_all(P& p)
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

struct source {
    uint8_t modcod;
    float es_tx;    // channel Es/N0
    float es_rx;    // the receiver's SNR (update_snr)
    traffic t;
};

struct arrival {
    int src, frame;  // frame -1: a frame of the wrong length
};

// MODCOD 4 = QPSK 1/2 at its waterfall (some frames fail), 14 = 8PSK 3/4,
// 18 = 16APSK 2/3, 26 = 32APSK 5/6 far below threshold (every frame fails)
std::vector<source> make_sources()
{
    const uint8_t ids[4] = {4, 14, 18, 26};
    const float es_tx[4] = {1.55f, 9.0f, 10.5f, 9.0f};
    const float es_rx[4] = {1.55f, 9.0f, 10.5f, 14.0f};
    std::vector<source> s;
    for (int k = 0; k < 4; k++) {
        const modcod_def_t* m = find_modcod(ids[k]);
        if (!m) throw std::runtime_error("no MODCOD");
        s.push_back({ids[k], es_tx[k], es_rx[k],
                     make_traffic(FECFRAME_NORMAL, m->code_rate, m->constellation, 24, es_tx[k], 70 + k)});
    }
    return s;
}

std::vector<arrival> make_arrivals(const std::vector<source>& src, uint32_t seed)
{
    std::mt19937 rng(seed);
    std::vector<int> left(src.size());
    for (size_t k = 0; k < src.size(); k++) left[k] = (int)src[k].t.payloads.size();
    std::vector<arrival> a;
    int next[8] = {};
    for (;;) {
        std::vector<int> open;
        for (size_t k = 0; k < src.size(); k++)
            if (left[k]) open.push_back((int)k);
        if (open.empty()) break;
        const int k = open[rng() % open.size()];
        const int run = std::min(left[k], 1 + (int)(rng() % 7));  // runs of 1..7 frames
        for (int j = 0; j < run; j++) a.push_back({k, next[k]++});
        left[k] -= run;
        if (rng() % 13 == 0) a.push_back({k, -1});
    }
    return a;
}

template <class P>
void feed(P& pool, const std::vector<source>& src, const std::vector<arrival>& arr, pool_output& got)
{
    for (const auto& s : src) pool.update_snr(s.modcod, s.es_rx);
    int n = 0;
    for (const auto& a : arr) {
        const traffic& t = src[a.src].t;
        const uint8_t mc = src[a.src].modcod;
        bool ready;
        if (a.frame < 0)
            ready = pool.submit(&t.symbols[0], t.xfec - 1, layout_of(mc));
        else
            ready = pool.submit(&t.symbols[(size_t)a.frame * t.xfec], t.xfec, layout_of(mc));
        if (ready) got.take_all(pool);
        if (++n % 17 == 0) QA_CHECK(!pool.submit(&t.symbols[0], t.xfec, layout_of(7)));  // not enabled
    }
    pool.finish_all_pending();
    got.take_all(pool);
}

}  // namespace

int main()
{
    return qa_run({
        {"same_output_as_rx_fec_pool",
         [] {
             const std::vector<source> src = make_sources();
             const std::vector<uint8_t> ids = {4, 14, 18, 26};
             const std::vector<arrival> arr = make_arrivals(src, 5);
             rx_fec_pool ref(FECFRAME_NORMAL, ids, 50, 64, 3, 4);
             pool_output want;
             feed(ref, src, arr, want);
             // the reference pool itself: arrival order of the payloads that passed
             QA_CHECK(ref.total_error_count() > 0);
             uint64_t f4 = 0, e4 = 0;
             QA_CHECK(ref.modcod_counters(4, f4, e4));
             std::printf("MODCOD 4 at its waterfall: %llu frames, %llu errors (drops included)\n",
                         (unsigned long long)f4, (unsigned long long)e4);
             QA_CHECK(e4 < f4);
             const int shapes[][2] = {{4, 3}, {16, 2}, {64, 4}, {1, 1}};  // micro-batch, streams
             for (const auto& sh : shapes) {
                 rx_fec_stream_pool pool(FECFRAME_NORMAL, ids, 50, 64, sh[1], sh[0]);
                 pool_output got;
                 feed(pool, src, arr, got);
                 QA_CHECK(got.bytes == want.bytes);
                 QA_CHECK(got.sizes == want.sizes);
                 QA_CHECK_EQ(pool.total_frame_count(), ref.total_frame_count());
                 QA_CHECK_EQ(pool.total_error_count(), ref.total_error_count());
                 QA_CHECK_EQ(pool.total_gpu_drop_count(), ref.total_gpu_drop_count());
                 for (uint8_t id : ids) {
                     uint64_t f = 0, e = 0, rf = 0, re = 0;
                     QA_CHECK(pool.modcod_counters(id, f, e));
                     QA_CHECK(ref.modcod_counters(id, rf, re));
                     QA_CHECK_EQ(f, rf);
                     QA_CHECK_EQ(e, re);
                 }
                 uint64_t f = 0, e = 0;
                 QA_CHECK(!pool.modcod_counters(7, f, e));
                 QA_CHECK(pool.data_snr_db() > 0.0f);
             }
         }},
        {"payloads_in_arrival_order",
         [] {
             // every frame above threshold: the output is every payload, in order
             const uint8_t ids[3] = {14, 18, 4};
             const float es[3] = {10.0f, 12.0f, 4.0f};
             std::vector<source> src;
             for (int k = 0; k < 3; k++) {
                 const modcod_def_t* m = find_modcod(ids[k]);
                 src.push_back({ids[k], es[k], es[k],
                                make_traffic(FECFRAME_SHORT, m->code_rate, m->constellation, 30, es[k], 90 + k)});
             }
             std::vector<arrival> arr = make_arrivals(src, 6);
             std::vector<arrival> clean;
             for (const auto& a : arr)
                 if (a.frame >= 0) clean.push_back(a);
             std::vector<uint8_t> want;
             std::vector<int> want_sizes;
             for (const auto& a : clean) {
                 const auto& p = src[a.src].t.payloads[a.frame];
                 want.insert(want.end(), p.begin(), p.end());
                 if (!p.empty()) want_sizes.push_back((int)p.size());
             }
             rx_fec_stream_pool pool(FECFRAME_SHORT, {4, 14, 18}, 50, 64, 2, 8);
             pool_output got;
             feed(pool, src, clean, got);
             QA_CHECK(got.bytes == want);
             QA_CHECK(got.sizes == want_sizes);
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)90);
             QA_CHECK_EQ(pool.total_error_count(), (uint64_t)0);
         }},
        {"launch_stale_then_drain",
         [] {
             const modcod_def_t* m = find_modcod(14);
             traffic t = make_traffic(FECFRAME_NORMAL, m->code_rate, m->constellation, 1, 10.0f, 51);
             rx_fec_stream_pool pool(FECFRAME_NORMAL, {14, 4}, 50, 64, 2, 4);
             pool.update_snr(14, 10.0f);
             QA_CHECK(!pool.submit(&t.symbols[0], t.xfec, layout_of(14)));  // micro-batch of 4 not full
             QA_CHECK_EQ(pool.pending_count(), 1);
             QA_CHECK(!pool.launch_stale(std::chrono::milliseconds(60000)));
             QA_CHECK(pool.launch_stale(std::chrono::milliseconds(0)));
             QA_CHECK_EQ(pool.pending_count(), 0);
             pool_output got;
             for (int tries = 0; tries < 2000 && got.blocks == 0; tries++) {
                 if (pool.drain_all_completed()) got.take_all(pool);
                 else std::this_thread::sleep_for(std::chrono::milliseconds(1));
             }
             QA_CHECK_EQ(got.blocks, 1);
             QA_CHECK(got.bytes == t.payloads[0]);
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)1);
         }},
        {"reset_and_warmup",
         [] {
             const modcod_def_t* m = find_modcod(4);
             traffic t = make_traffic(FECFRAME_SHORT, m->code_rate, m->constellation, 4, 5.0f, 52);
             rx_fec_stream_pool pool(FECFRAME_SHORT, {4, 14}, 50, 64, 2, 4);
             pool.warmup_all();
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)0);
             QA_CHECK_EQ(pool.total_error_count(), (uint64_t)0);
             pool.update_snr(4, 5.0f);
             for (int b = 0; b < 4; b++) pool.submit(&t.symbols[(size_t)b * t.xfec], t.xfec, layout_of(4));
             pool.finish_all_pending();
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)4);
             std::vector<uint8_t> buf(pool.max_output_bytes());
             pool.reset_all();
             QA_CHECK_EQ(pool.get_gse_output(buf.data(), (int)buf.size()), 0);
             QA_CHECK(pool.gse_dfl_sizes().empty());
             QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)0);
             // usable again after the reset
             pool.update_snr(4, 5.0f);
             for (int b = 0; b < 4; b++) pool.submit(&t.symbols[(size_t)b * t.xfec], t.xfec, layout_of(4));
             pool.finish_all_pending();
             pool_output got;
             got.take_all(pool);
             std::vector<uint8_t> want;
             for (const auto& p : t.payloads) want.insert(want.end(), p.begin(), p.end());
             QA_CHECK(got.bytes == want);
         }},
        {"full_ring_drops_in_order_or_blocks",
         [] {
             // micro-batches of one frame, a ring of 512; the shared decode
             // budget is held by someone else (in_flight above a cap of 1), so
             // nothing launches on its own: the ring fills after 510 frames
             // and the rest are dropped, or, with block-when-full, each
             // submit waits for the oldest micro-batch (launched regardless)
             const modcod_def_t* m = find_modcod(4);
             traffic t = make_traffic(FECFRAME_SHORT, m->code_rate, m->constellation, 50, 6.0f, 60);
             const int n = 600;
             for (int block = 0; block < 2; block++) {
                 rx_fec_stream_pool pool(FECFRAME_SHORT, {4}, 50, 64, 2, 1);
                 auto held = std::make_shared<gpu_decode_budget>();
                 held->cap = 1;
                 held->in_flight = 1000;
                 pool.set_decode_budget(held);
                 pool.set_block_when_full(block == 1);
                 pool.update_snr(4, 6.0f);
                 pool_output got;
                 std::vector<int> sent;
                 for (int k = 0; k < n; k++) {
                     sent.push_back(k % 50);
                     if (pool.submit(&t.symbols[(size_t)(k % 50) * t.xfec], t.xfec, layout_of(4))) got.take_all(pool);
                 }
                 pool.finish_all_pending();
                 got.take_all(pool);
                 QA_CHECK_EQ(pool.total_frame_count(), (uint64_t)n);
                 QA_CHECK_EQ(pool.total_error_count(), pool.total_gpu_drop_count());  // only drops fail
                 QA_CHECK_EQ(pool.total_gpu_drop_count(), (uint64_t)(block ? 0 : n - 510));
                 std::printf("block_when_full %d: %llu of %d frames dropped\n", block,
                             (unsigned long long)pool.total_gpu_drop_count(), n);
                 // the output is the payloads of the kept frames, in arrival order
                 size_t pos = 0, byte = 0;
                 int kept = 0;
                 bool in_order = true;
                 for (int sz : got.sizes) {
                     auto matches = [&](size_t p) {
                         const auto& pl = t.payloads[sent[p]];
                         return (int)pl.size() == sz && std::equal(pl.begin(), pl.end(), got.bytes.begin() + byte);
                     };
                     while (pos < sent.size() && !matches(pos)) pos++;
                     if (pos == sent.size()) {
                         in_order = false;
                         break;
                     }
                     byte += sz;
                     pos++;
                     kept++;
                 }
                 QA_CHECK(in_order);
                 QA_CHECK_EQ(byte, got.bytes.size());
                 int want_kept = 0;  // non-empty payloads among the frames that were kept
                 for (int k = 0; k < (block ? n : 510); k++) want_kept += t.payloads[sent[k]].empty() ? 0 : 1;
                 QA_CHECK_EQ(kept, want_kept);
             }
         }},
        {"learned_parameters",
         [] {
             // learned tables for MODCODs 4 and 14 that restate the plain check
             // node (offset 1, no scale): every output byte, size and counter
             // equals the plain pool's, the fallback frames being redone by
             // the plain decoder in the same launch; tables that silence the
             // messages (offset 127, scale 1/64): with thousands of channel
             // bit errors per frame no learned pass converges, every decoded
             // frame of 4 and 14 falls back, the output is still the plain
             // pool's
             const std::vector<source> src = make_sources();
             const std::vector<uint8_t> ids = {4, 14, 18, 26};
             const std::vector<arrival> arr = make_arrivals(src, 5);
             rx_fec_stream_pool plain(FECFRAME_NORMAL, ids, 50, 64, 3, 4);
             pool_output want;
             feed(plain, src, arr, want);
             QA_CHECK_EQ(plain.learned_fallback_count(), (uint64_t)0);
             uint64_t decoded = 0;  // frames of MODCODs 4 and 14 of the right length
             for (const auto& a : arr) decoded += (a.frame >= 0 && (a.src == 0 || a.src == 1)) ? 1 : 0;
             for (int kind = 0; kind < 2; kind++) {
                 rx_fec_stream_pool pool(FECFRAME_NORMAL, ids, 50, 64, 3, 4);
                 QA_CHECK_EQ(pool.layer_count(7), -1);
                 QA_CHECK(!pool.set_learned_params(7, nullptr, nullptr, 0));
                 const int T = 20;
                 for (uint8_t id : {(uint8_t)4, (uint8_t)14}) {
                     const int q = pool.layer_count(id);
                     QA_CHECK(q > 0);
                     std::vector<int16_t> beta((size_t)T * q, (int16_t)(kind == 0 ? 1 : 127));
                     std::vector<int16_t> alpha((size_t)T * q, (int16_t)(kind == 0 ? 0 : 1));
                     QA_CHECK(pool.set_learned_params(id, beta.data(), alpha.data(), T));
                 }
                 pool_output got;
                 feed(pool, src, arr, got);
                 QA_CHECK(got.bytes == want.bytes);
                 QA_CHECK(got.sizes == want.sizes);
                 QA_CHECK_EQ(pool.total_frame_count(), plain.total_frame_count());
                 QA_CHECK_EQ(pool.total_error_count(), plain.total_error_count());
                 const uint64_t fb = pool.learned_fallback_count();
                 std::printf("learned tables %s: %llu of %llu frames of MODCODs 4 and 14 fell back\n",
                             kind == 0 ? "offset 1" : "offset 127, scale 1/64", (unsigned long long)fb,
                             (unsigned long long)decoded);
                 if (kind == 0) QA_CHECK(fb < decoded);
                 else QA_CHECK_EQ(fb, decoded);
                 // back to plain: no more fallbacks
                 pool.reset_all();
                 for (uint8_t id : {(uint8_t)4, (uint8_t)14}) QA_CHECK(pool.set_learned_params(id, nullptr, nullptr, 0));
                 pool_output again;
                 feed(pool, src, arr, again);
                 QA_CHECK(again.bytes == want.bytes);
                 QA_CHECK_EQ(pool.learned_fallback_count(), (uint64_t)0);
             }
         }},
        {"unknown_index_in_constructor_throws",
         [] {
             bool threw = false;
             try {
                 rx_fec_stream_pool pool(FECFRAME_NORMAL, {14, 200});
             } catch (const std::invalid_argument&) {
                 threw = true;
             }
             QA_CHECK(threw);
         }},
    });
}
