// GPU checks of the LLR-scale front door (cusat_stream_rx_dd): the
// decision-directed SNR estimated on the GPU equals, bit for bit, the host
// estimate of the cusat_dsp API (qpsk_snr_dd / constellation_snr_dd, after
// modulation/dd_snr.cc:23-51) on noisy, noiseless, all-zero, non-finite and
// huge-valued frames of QPSK, 8PSK, 16APSK, 32APSK and 64APSK codes; the
// events of a batch apply in order (a later host value wins over an earlier
// estimate, an estimate of at most 1 changes nothing); noise variance and
// LLR gain follow the host formulas exactly; the state carries from batch to
// batch; the rx_fec_stream_pool with the GPU estimate gives the same output
// and counters as with the host estimate when every micro-batch launches at
// once.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "cusat_dsp/rx_fec_chain.h"
#include "cusat_dsp/rx_fec_stream_pool.h"
#include "cusat_fec.h"
#include "qa_traffic.h"
#include "qa_util.h"

using namespace cusat_dsp;

namespace {

struct code_setup {
    int fs, rate, mod;
    cusat_fec_ctx* ctx = nullptr;
    cusat_fec_info inf{};
    std::vector<complexf> points;
};

uint32_t bits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

float host_estimate(const code_setup& c, const complexf* sym)
{
    if (c.mod == CUSAT_MOD_QPSK) return qpsk_snr_dd(sym, c.inf.xfec_len);
    return constellation_snr_dd(c.points.data(), (int)c.points.size(), sym, c.inf.xfec_len, 8);
}

// frame kinds: 0 noisy (sigma), 1 noiseless, 2 all zero, 3 one NaN symbol,
// 4 one infinite symbol, 5 huge values
std::vector<complexf> make_frame(const code_setup& c, int kind, float sigma, std::mt19937& rng)
{
    std::vector<complexf> s(c.inf.xfec_len);
    std::normal_distribution<float> g(0.0f, sigma);
    for (auto& x : s) {
        const complexf p = c.points[rng() % c.points.size()];
        x = kind == 1 ? p : p + complexf(g(rng), g(rng));
        if (kind == 5) x *= 1e19f;
    }
    if (kind == 2) std::fill(s.begin(), s.end(), complexf(0.0f, 0.0f));
    if (kind == 3) s[8 * 17] = complexf(std::numeric_limits<float>::quiet_NaN(), 0.5f);
    if (kind == 4) s[8 * 3] = complexf(0.1f, std::numeric_limits<float>::infinity());
    return s;
}

}  // namespace

int main()
{
    return qa_run({
        {"gpu_estimate_equals_host_estimate",
         [] {
             std::vector<code_setup> codes = {{0, CUSAT_RATE_1_2, CUSAT_MOD_QPSK},
                                              {0, CUSAT_RATE_3_4, CUSAT_MOD_8PSK},
                                              {1, CUSAT_RATE_2_3, CUSAT_MOD_16APSK},
                                              {0, CUSAT_RATE_3_4, CUSAT_MOD_32APSK},
                                              {0, CUSAT_RATE_4_5, CUSAT_MOD_64APSK},
                                              {1, CUSAT_RATE_1_4, CUSAT_MOD_QPSK}};
             std::vector<cusat_fec_ctx*> ctxs;
             for (auto& c : codes) {
                 QA_CHECK_EQ(cusat_fec_ctx_create(&c.ctx, c.fs, c.rate, c.mod, 10, 1, 0), CUSAT_OK);
                 cusat_fec_ctx_info(c.ctx, &c.inf);
                 std::vector<float> pts(2u << c.inf.n_mod);
                 uint32_t cols[8];
                 int nm = 0, rows = 0;
                 cusat_modcod_constellation(c.fs, c.rate, c.mod, pts.data(), cols, &nm, &rows);
                 for (size_t k = 0; k < pts.size() / 2; k++) c.points.emplace_back(pts[2 * k], pts[2 * k + 1]);
                 ctxs.push_back(c.ctx);
             }
             const int nc = (int)codes.size();
             const int kinds = 6;
             const float sigmas[3] = {0.05f, 0.3f, 1.5f};
             const int per_code = kinds + 2;  // every kind + two more noise levels
             const int n = nc * per_code;
             cusat_stream_ctx* sc = nullptr;
             QA_CHECK_EQ(cusat_stream_create(&sc, ctxs.data(), nc, n, 0), CUSAT_OK);
             std::mt19937 rng(11);
             std::vector<complexf> all;
             std::vector<cusat_stream_rx_frame> fr;
             std::vector<float> want;  // host estimate per frame
             for (int k = 0; k < nc; k++)
                 for (int j = 0; j < per_code; j++) {
                     const int kind = j < kinds ? j : 0;
                     const float sigma = j < kinds ? sigmas[1] : sigmas[j == kinds ? 0 : 2];
                     std::vector<complexf> s = make_frame(codes[k], kind, sigma, rng);
                     cusat_stream_rx_frame f{};
                     f.code = k;
                     f.flags = CUSAT_RX_SCALE_STATE;
                     f.sym_offset = (int64_t)all.size();
                     f.out_offset = 0;
                     fr.push_back(f);
                     want.push_back(host_estimate(codes[k], s.data()));
                     all.insert(all.end(), s.begin(), s.end());
                 }
             float *sym = nullptr, *state = nullptr, *state_out = nullptr;
             QA_CHECK_EQ(hipMalloc(&sym, all.size() * sizeof(complexf)), hipSuccess);
             QA_CHECK_EQ(hipMalloc(&state, nc * 16), hipSuccess);
             QA_CHECK_EQ(hipHostMalloc((void**)&state_out, nc * 16, hipHostMallocDefault), hipSuccess);
             hipMemcpy(sym, all.data(), all.size() * sizeof(complexf), hipMemcpyHostToDevice);
             // one estimate per batch and code: frame j of every code in batch j
             // (the state starts at snr_lin 0, so it shows the estimate or, at
             // most 1, keeps the previous value)
             std::vector<float> prev(nc, 0.0f);
             int mism = 0, checked = 0;
             for (int j = 0; j < per_code; j++) {
                 std::vector<float> init(4 * nc, 0.0f);
                 for (int k = 0; k < nc; k++) {
                     init[4 * k] = 1.0f;
                     init[4 * k + 1] = 4.0f;
                 }
                 hipMemcpy(state, init.data(), init.size() * 4, hipMemcpyHostToDevice);
                 std::vector<cusat_stream_rx_event> ev;
                 for (int k = 0; k < nc; k++) ev.push_back({k, k * per_code + j, 0.0f, 0.0f});
                 QA_CHECK_EQ(cusat_stream_rx_plan_async(sc, fr.data(), n, ev.data(), (int)ev.size(), nullptr),
                             CUSAT_OK);
                 QA_CHECK_EQ(cusat_stream_rx_dd(sc, sym, state, state_out, nullptr), CUSAT_OK);
                 hipDeviceSynchronize();
                 for (int k = 0; k < nc; k++) {
                     const float w = want[k * per_code + j];
                     const float* st = state_out + 4 * k;
                     checked++;
                     if (w > 1.0f) {
                         const float nv = 1.0f / w;
                         const float lg = std::fmin(4.0f / nv, 1000.0f);
                         if (bits(st[2]) != bits(w) || bits(st[0]) != bits(nv) || bits(st[1]) != bits(lg) ||
                             !std::isnan(st[3])) {
                             std::printf("code %d frame kind %d: gpu %.9g host %.9g\n", k, j, st[2], w);
                             mism++;
                         }
                     } else if (st[2] != 0.0f || st[0] != 1.0f || st[1] != 4.0f) {
                         std::printf("code %d frame kind %d: host estimate %.9g (no update), gpu state %.9g\n", k, j,
                                     w, st[2]);
                         mism++;
                     }
                 }
             }
             QA_CHECK_EQ(mism, 0);
             QA_CHECK_EQ(checked, nc * per_code);
             // events in order: host value, estimate, host value for another code
             {
                 std::vector<float> init(4 * nc, 0.0f);
                 for (int k = 0; k < nc; k++) {
                     init[4 * k] = 1.0f;
                     init[4 * k + 1] = 4.0f;
                 }
                 hipMemcpy(state, init.data(), init.size() * 4, hipMemcpyHostToDevice);
                 const float v5 = std::pow(10.0f, 0.5f), v7 = std::pow(10.0f, 0.7f);
                 std::vector<cusat_stream_rx_event> ev = {
                     {0, -1, v5, 5.0f},          // code 0: host 5 dB ...
                     {0, 0, 0.0f, 0.0f},         // ... then the estimate of frame 0 (noisy)
                     {1, per_code + 0, 0, 0},    // code 1: estimate ...
                     {1, -1, v7, 7.0f},          // ... then host 7 dB wins
                     {2, -1, v5, 5.0f},          // code 2: host 5 dB ...
                     {2, 2 * per_code + 2, 0, 0}  // ... an all-zero frame: estimate 1, no change
                 };
                 QA_CHECK_EQ(cusat_stream_rx_plan_async(sc, fr.data(), n, ev.data(), (int)ev.size(), nullptr),
                             CUSAT_OK);
                 QA_CHECK_EQ(cusat_stream_rx_dd(sc, sym, state, state_out, nullptr), CUSAT_OK);
                 hipDeviceSynchronize();
                 QA_CHECK_EQ(bits(state_out[2]), bits(want[0]));
                 QA_CHECK(std::isnan(state_out[3]));
                 QA_CHECK_EQ(bits(state_out[4 + 2]), bits(v7));
                 QA_CHECK_EQ(bits(state_out[4 + 0]), bits(1.0f / v7));
                 QA_CHECK(state_out[4 + 3] == 7.0f);
                 QA_CHECK_EQ(bits(state_out[8 + 2]), bits(v5));
                 QA_CHECK(state_out[8 + 3] == 5.0f);
                 QA_CHECK(state_out[12 + 2] == 0.0f);  // no event: unchanged
                 // the next batch starts from this state
                 QA_CHECK_EQ(cusat_stream_rx_plan_async(sc, fr.data(), n, nullptr, 0, nullptr), CUSAT_OK);
                 QA_CHECK_EQ(cusat_stream_rx_dd(sc, sym, state, state_out, nullptr), CUSAT_OK);
                 hipDeviceSynchronize();
                 QA_CHECK_EQ(bits(state_out[4 + 2]), bits(v7));
             }
             // bad events are refused
             {
                 cusat_stream_rx_event bad[1] = {{0, n, 0.0f, 0.0f}};
                 QA_CHECK(cusat_stream_rx_plan_async(sc, fr.data(), n, bad, 1, nullptr) != CUSAT_OK);
                 bad[0] = {0, -1, -1.0f, 0.0f};
                 QA_CHECK(cusat_stream_rx_plan_async(sc, fr.data(), n, bad, 1, nullptr) != CUSAT_OK);
                 bad[0] = {nc, 0, 0.0f, 0.0f};
                 QA_CHECK(cusat_stream_rx_plan_async(sc, fr.data(), n, bad, 1, nullptr) != CUSAT_OK);
             }
             hipFree(sym);
             hipFree(state);
             hipHostFree(state_out);
             cusat_stream_destroy(sc);
             for (auto& c : codes) cusat_fec_ctx_destroy(c.ctx);
         }},
        {"pool_gpu_estimate_equals_host_estimate",
         [] {
             // 40 frames of each of three MODCODs near their waterfall, so the
             // estimate fires and the scaling decides some frames
             const uint8_t ids[3] = {4, 14, 18};
             const float es[3] = {1.6f, 8.2f, 9.3f};
             std::vector<traffic> t;
             for (int k = 0; k < 3; k++) {
                 const modcod_def_t* m = find_modcod(ids[k]);
                 t.push_back(make_traffic(FECFRAME_NORMAL, m->code_rate, m->constellation, 40, es[k], 300 + k));
             }
             std::mt19937 rng(3);
             std::vector<std::pair<int, int>> arr;
             int next[3] = {0, 0, 0};
             while (arr.size() < 120) {
                 const int k = (int)(rng() % 3);
                 const int run = 1 + (int)(rng() % 5);
                 for (int j = 0; j < run && next[k] < 40; j++) arr.push_back({k, next[k]++});
                 if (next[0] == 40 && next[1] == 40 && next[2] == 40) break;
             }
             std::vector<uint8_t> out[2];
             std::vector<int> sizes[2];
             uint64_t frames[2][3], errors[2][3];
             float snr[2][3];
             for (int mode = 0; mode < 2; mode++) {
                 rx_fec_stream_pool pool(FECFRAME_NORMAL, {4, 14, 18}, 50, 64, 3, 16);
                 pool.set_gpu_dd_snr(mode == 1);
                 auto budget = std::make_shared<gpu_decode_budget>();
                 budget->cap = 0;  // every micro-batch launches when it closes
                 pool.set_decode_budget(budget);
                 for (int k = 0; k < 3; k++) pool.update_snr(ids[k], es[k] + 0.4f);
                 std::vector<uint8_t> buf(pool.max_output_bytes());
                 auto take = [&] {
                     for (;;) {
                         const int nb = pool.get_gse_output(buf.data(), (int)buf.size());
                         if (nb <= 0) break;
                         out[mode].insert(out[mode].end(), buf.begin(), buf.begin() + nb);
                         sizes[mode].insert(sizes[mode].end(), pool.gse_dfl_sizes().begin(),
                                            pool.gse_dfl_sizes().end());
                     }
                 };
                 pl_layout l;
                 int n = 0;
                 for (const auto& a : arr) {
                     l.modcod = ids[a.first];
                     if (pool.submit(&t[a.first].symbols[(size_t)a.second * t[a.first].xfec], t[a.first].xfec, l))
                         take();
                     if (++n == 50) pool.update_snr(ids[1], es[1] + 1.0f);  // a pilot value mid-stream
                 }
                 pool.finish_all_pending();
                 take();
                 for (int k = 0; k < 3; k++) pool.modcod_counters(ids[k], frames[mode][k], errors[mode][k]);
                 snr[mode][0] = pool.data_snr_db();  // the latest output's MODCOD
                 std::printf("mode %s: errors %llu %llu %llu, data snr %.6f dB\n", mode ? "gpu" : "host",
                             (unsigned long long)errors[mode][0], (unsigned long long)errors[mode][1],
                             (unsigned long long)errors[mode][2], snr[mode][0]);
             }
             QA_CHECK(out[0] == out[1]);
             QA_CHECK(sizes[0] == sizes[1]);
             for (int k = 0; k < 3; k++) {
                 QA_CHECK_EQ(frames[0][k], frames[1][k]);
                 QA_CHECK_EQ(errors[0][k], errors[1][k]);
             }
             QA_CHECK_EQ(bits(snr[0][0]), bits(snr[1][0]));
         }},
    });
}
