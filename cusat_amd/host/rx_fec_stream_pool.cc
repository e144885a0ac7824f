// cusat_dsp::rx_fec_stream_pool: the rx_fec_pool API over one flight ring of
// mixed-MODCOD micro-batches, each decoded by a graph replay of
// cusat_stream_rx.  Ring, budget and drop rules follow rx_fec_chain.cc
// (submit / launch / drain order, the decode budget, drops when the ring is
// full, the SNR re-estimate every 32 frames of a MODCOD); routing, counters
// and output FIFO follow rx_fec_pool.cc:15-295 without the flush on MODCOD
// changes, which one ring in arrival order makes unnecessary.
#include "cusat_dsp/rx_fec_stream_pool.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "copy_frame.h"
#include "cusat_fec.h"

namespace cusat_dsp {

namespace {

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string("rx_fec_stream_pool: ") + what + ": " + hipGetErrorString(e));
}

void abi_check(int rc, const char* what)
{
    if (rc != CUSAT_OK)
        throw std::runtime_error(std::string("rx_fec_stream_pool: ") + what + ": " + cusat_last_error());
}

size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

rx_fec_stream_pool::rx_fec_stream_pool(fecframe_size_t framesize, const std::vector<uint8_t>& enabled_modcods,
                                       int max_trials, int cuda_max_batch, int n_gpu_streams, int micro_batch)
    : d_micro_batch(micro_batch)
{
    (void)cuda_max_batch;  // sized the reference's spare decoder; micro_batch bounds a launch here
    if (micro_batch < 1 || n_gpu_streams < 1 || max_trials < 0)
        throw std::invalid_argument("rx_fec_stream_pool: micro_batch and n_gpu_streams must be >= 1");
    std::vector<uint8_t> ids(enabled_modcods);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    for (uint8_t idx : ids)
        if (!find_modcod(idx)) throw std::invalid_argument("Unknown modcod_idx: " + std::to_string(idx));
    try {
        for (uint8_t idx : ids) {
            const modcod_def_t* m = find_modcod(idx);
            cusat_fec_info inf;
            abi_check(cusat_fec_query((int)framesize, (int)m->code_rate, (int)m->constellation, 0, &inf), "query");
            cusat_fec_ctx* c = nullptr;
            abi_check(cusat_fec_ctx_create(&c, (int)framesize, (int)m->code_rate, (int)m->constellation, max_trials,
                                           1, 0),
                      "context");
            d_ctx.push_back(c);
            modcod_state s;
            s.idx = idx;
            s.code = (int)d_mc.size();
            s.modulation = m->constellation;
            s.n_mod = inf.n_mod;
            s.xfec = inf.xfec_len;
            s.max_dfl = inf.k_bch / 8 - 10;
            s.q = inf.q;
            std::vector<float> pts(2u << inf.n_mod);
            uint32_t cols[8];
            int nm = 0, rows = 0;
            abi_check(cusat_modcod_constellation((int)framesize, (int)m->code_rate, (int)m->constellation,
                                                 pts.data(), cols, &nm, &rows),
                      "constellation");
            for (size_t k = 0; k < pts.size() / 2; k++) s.points.emplace_back(pts[2 * k], pts[2 * k + 1]);
            d_max_xfec = std::max(d_max_xfec, s.xfec);
            d_max_dfl = std::max(d_max_dfl, s.max_dfl);
            d_by_idx[idx] = (int)d_mc.size();
            d_mc.push_back(std::move(s));
        }
        const size_t mb = (size_t)micro_batch;
        const size_t state_bytes = std::max<size_t>(d_mc.size(), 1) * 4 * sizeof(float);
        d_res_bytes = round256(4 * mb * sizeof(int32_t));  // trials | valid | dfl | fallback
        d_state_bytes = round256(state_bytes);
        hip_check(hipMalloc(&d_state, state_bytes), "state");
        {
            hipEvent_t ev = nullptr;
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "events");
            d_dd_done = ev;
        }
        reset_state();
        d_streams.resize(n_gpu_streams);
        for (auto& r : d_streams) {
            hipStream_t st = nullptr;
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
            r.stream = st;
            if (d_ctx.empty()) continue;
            // learned mode: a MODCOD decodes learned only once set_learned_params
            // gave its context parameters; until then it decodes plain
            abi_check(cusat_stream_create(&r.sc, d_ctx.data(), (int)d_ctx.size(), micro_batch, CUSAT_STREAM_LEARNED),
                      "stream context");
            hip_check(hipMalloc(&r.sym_dev, mb * (d_max_xfec + 8) * sizeof(complexf)), "device buffers");
            hip_check(hipMalloc(&r.blk, d_res_bytes + d_state_bytes + std::max<size_t>(mb * d_max_dfl, 1)),
                      "device buffers");
            r.res_dev = reinterpret_cast<int32_t*>(r.blk);
            r.state_dev = reinterpret_cast<float*>(r.blk + d_res_bytes);
            r.out_dev = r.blk + d_res_bytes + d_state_bytes;
            // capture the SNR launch and the three receive launches once;
            // every micro-batch replays them
            abi_check(cusat_stream_rx_plan_async(r.sc, nullptr, 0, nullptr, 0, st), "plan");
            int32_t* res = r.res_dev;
            const float* sym = reinterpret_cast<const float*>(r.sym_dev);
            // cusat_stream_rx's three launches, with the decode's fallback flags kept
            auto rx = [&] {
                int rc = cusat_stream_rx_demap(r.sc, sym, st);
                if (rc == CUSAT_OK) rc = cusat_stream_rx_decode(r.sc, res, res + mb, res + 3 * mb, st);
                if (rc == CUSAT_OK) rc = cusat_stream_rx_bb(r.sc, r.out_dev, res, res + mb, res + 2 * mb, st);
                return rc;
            };
            auto est = [&] { return cusat_stream_rx_dd_estimate(r.sc, sym, st); };
            auto resolve = [&] { return cusat_stream_rx_dd_resolve(r.sc, d_state, r.state_dev, st); };
            abi_check(est(), "snr");
            abi_check(resolve(), "snr");
            abi_check(rx(), "rx");
            hip_check(hipStreamSynchronize(st), "warm launch");
            for (int k = 0; k < 3; k++) {
                hipGraph_t g = nullptr;
                hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture");
                const int rc = k == 0 ? est() : k == 1 ? resolve() : rx();
                const hipError_t ec = hipStreamEndCapture(st, &g);
                abi_check(rc, "capture");
                hip_check(ec, "capture");
                hipGraphExec_t ge = nullptr;
                const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                hip_check(ei, "graph instantiate");
                (k == 0 ? r.graph_dd_est : k == 1 ? r.graph_dd : r.graph) = ge;
            }
        }
        // a ring of about N_FRAMES_BUFFERED frames (128 micro-batches of 4 in
        // the reference's chain), at least 8 micro-batches
        d_n_flight = std::max(8, (N_FRAMES_BUFFERED + micro_batch - 1) / micro_batch);
        d_flight.resize(d_n_flight);
        // an output block never exceeds the ring's frames of the largest code
        d_max_output_bytes = (d_max_dfl + 10) * d_n_flight * micro_batch;
        for (auto& f : d_flight) {
            const size_t sym_bytes = mb * (d_max_xfec + 8) * sizeof(complexf);
            const size_t blk_bytes = d_res_bytes + d_state_bytes + std::max<size_t>(mb * d_max_dfl, 1);
            hip_check(hipHostMalloc((void**)&f.syms, sym_bytes, hipHostMallocDefault), "pinned slots");
            hip_check(hipHostMalloc((void**)&f.blk, blk_bytes, hipHostMallocDefault), "pinned slots");
            f.res = reinterpret_cast<int32_t*>(f.blk);
            f.state = reinterpret_cast<float*>(f.blk + d_res_bytes);
            f.out = f.blk + d_res_bytes + d_state_bytes;
            // touch every page now: a first write into pinned memory costs a
            // fault per page, which would land in submit
            memset(f.syms, 0, sym_bytes);
            memset(f.blk, 0, blk_bytes);
            f.events.reserve(2 * mb + d_mc.size());
            hipEvent_t ev = nullptr;
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "events");
            f.event = ev;
            f.mc.reserve(mb);
            f.sym_off.reserve(mb);
            f.out_off.reserve(mb);
        }
    } catch (...) {
        free_all();
        throw;
    }
    d_budget = gpu_decode_budget::make_default();
}

void rx_fec_stream_pool::free_all()
{
    for (auto& r : d_streams)
        if (r.stream) hipStreamSynchronize(r.stream);
    for (auto& f : d_flight) {
        if (f.syms) hipHostFree(f.syms);
        if (f.blk) hipHostFree(f.blk);
        if (f.event) hipEventDestroy(f.event);
        f = flight_slot{};
    }
    for (auto& r : d_streams) {
        if (r.graph) hipGraphExecDestroy(r.graph);
        if (r.graph_dd) hipGraphExecDestroy(r.graph_dd);
        if (r.graph_dd_est) hipGraphExecDestroy(r.graph_dd_est);
        if (r.sc) cusat_stream_destroy(r.sc);
        if (r.sym_dev) hipFree(r.sym_dev);
        if (r.blk) hipFree(r.blk);
        if (r.stream) hipStreamDestroy(r.stream);
        r = stream_res{};
    }
    for (cusat_fec_ctx* c : d_ctx) cusat_fec_ctx_destroy(c);
    d_ctx.clear();
    if (d_state) hipFree(d_state);
    d_state = nullptr;
    if (d_dd_done) hipEventDestroy(d_dd_done);
    d_dd_done = nullptr;
}

rx_fec_stream_pool::~rx_fec_stream_pool()
{
    // frames still in flight hold budget that may be shared
    for (; d_drain_idx < d_launch_idx; d_drain_idx++) {
        flight_slot& s = d_flight[d_drain_idx % d_n_flight];
        if (s.in_flight) {
            hipEventSynchronize(s.event);
            release_budget(s.n_frames);
            s.in_flight = false;
        }
    }
    free_all();
}

void rx_fec_stream_pool::update_dd_snr(modcod_state& m, const complexf* frame)
{
    float snr_lin;
    if (m.n_mod >= 7 && m.pilot_snr_db > 5.0f)  // dense APSK: trust the pilot estimate
        snr_lin = std::pow(10.0f, m.pilot_snr_db / 10.0f);
    else if (m.modulation == MOD_QPSK)
        snr_lin = qpsk_snr_dd(frame, m.xfec);
    else
        snr_lin = constellation_snr_dd(m.points.data(), (int)m.points.size(), frame, m.xfec, 8);
    if (snr_lin > 1.0f) {
        m.snr_db = 10.0f * std::log10(snr_lin);
        m.noise_var = 1.0f / snr_lin;
        m.llr_gain = std::fmin(4.0f / m.noise_var, 1000.0f);
    }
}

void rx_fec_stream_pool::update_snr(uint8_t modcod, float snr_db)
{
    auto it = d_by_idx.find(modcod);
    if (it == d_by_idx.end() || !(snr_db > 0.0f)) return;
    modcod_state& m = d_mc[it->second];
    m.pilot_snr_db = snr_db;
    const float snr_lin = std::pow(10.0f, snr_db / 10.0f);
    m.snr_db = snr_db;
    m.noise_var = 1.0f / snr_lin;
    m.llr_gain = std::fmin(4.0f / m.noise_var, 1000.0f);
    if (d_gpu_dd) add_host_event(m, snr_lin, snr_db);
}

void rx_fec_stream_pool::add_host_event(const modcod_state& m, float snr_lin, float snr_db)
{
    d_mc[m.code].host_event_batch = d_submit_idx;
    // a host value that no estimate of the MODCOD follows yet is replaced,
    // which keeps a micro-batch's events within 2 x frames + MODCODs
    for (auto e = d_events.rbegin(); e != d_events.rend(); ++e) {
        if (e->code != m.code) continue;
        if (e->frame < 0) {
            e->snr_lin = snr_lin;
            e->snr_db = snr_db;
            return;
        }
        break;
    }
    d_events.push_back(cusat_stream_rx_event{m.code, -1, snr_lin, snr_db});
}

void rx_fec_stream_pool::reset_state()
{
    std::vector<float> init(std::max<size_t>(d_mc.size(), 1) * 4, 0.0f);
    for (size_t k = 0; k < d_mc.size(); k++) {
        init[4 * k] = 1.0f;  // noise_var, llr_gain of a fresh chain
        init[4 * k + 1] = 4.0f;
    }
    hip_check(hipMemcpy(d_state, init.data(), init.size() * sizeof(float), hipMemcpyHostToDevice), "state");
    d_events.clear();
    d_dd_recorded = false;
    for (auto& m : d_mc) m.host_event_batch = -1;
}

bool rx_fec_stream_pool::submit(const complexf* symbols, int n_symbols, const pl_layout& pls)
{
    auto it = d_by_idx.find(pls.modcod);
    if (it == d_by_idx.end()) return false;
    modcod_state& m = d_mc[it->second];
    // drain at micro-batch boundaries, when the ring fills up and while a
    // closed micro-batch waits for decode budget (a drain releases budget)
    if (d_batch_idx == 0 || d_launch_idx < d_submit_idx || d_submit_idx - d_drain_idx >= d_n_flight / 2)
        drain(false);
    while (d_block_when_full && d_submit_idx - d_drain_idx >= d_n_flight - 2) drain(true, true);
    if (d_submit_idx - d_drain_idx >= d_n_flight - 2 || n_symbols != m.xfec || !symbols) {
        d_gpu_drop_cnt++;
        m.frames++;
        m.errors++;
        return !d_ready.empty();
    }
    flight_slot& s = d_flight[d_submit_idx % d_n_flight];
    if (d_batch_idx == 0) {
        s.mc.clear();
        s.sym_off.clear();
        s.out_off.clear();
        s.n_syms = s.n_out = 0;
        d_batch_start_time = std::chrono::steady_clock::now();
    }
    s.n_syms = (s.n_syms + 7) & ~(int64_t)7;  // 64-byte aligned frame starts
    complexf* dst = s.syms + s.n_syms;
    detail::copy_frame(dst, symbols, (size_t)n_symbols * sizeof(complexf));
    if (++m.snr_skip >= 32) {
        m.snr_skip = 0;
        if (!d_gpu_dd) {
            update_dd_snr(m, symbols);  // the caller's copy: pinned memory is slow to read from the host
        } else if (m.n_mod >= 7 && m.pilot_snr_db > 5.0f) {  // dense APSK: the pilot estimate, on the host
            const float snr_lin = std::pow(10.0f, m.pilot_snr_db / 10.0f);
            add_host_event(m, snr_lin, 10.0f * std::log10(snr_lin));
        } else {
            d_events.push_back(cusat_stream_rx_event{m.code, d_batch_idx, 0.0f, 0.0f});
        }
    }
    s.mc.push_back(it->second);
    s.sym_off.push_back(s.n_syms);
    s.out_off.push_back(s.n_out);
    s.n_syms += n_symbols;
    s.n_out += m.max_dfl;
    if (++d_batch_idx >= d_micro_batch) close_batch();
    return !d_ready.empty();
}

void rx_fec_stream_pool::close_batch()
{
    if (d_batch_idx > 0) {
        flight_slot& s = d_flight[d_submit_idx % d_n_flight];
        s.events.swap(d_events);
        d_events.clear();
        s.n_frames = d_batch_idx;
        s.batch = d_submit_idx;
        d_submit_idx++;
        d_batch_idx = 0;
    }
    pump_launches();
}

void rx_fec_stream_pool::pump_launches(bool bypass_cap)
{
    while (d_launch_idx < d_submit_idx) {
        flight_slot& s = d_flight[d_launch_idx % d_n_flight];
        if (!bypass_cap && d_budget && d_budget->cap > 0) {
            const int cur = d_budget->in_flight.load(std::memory_order_relaxed);
            // an idle GPU always admits, so an oversized micro-batch cannot stall
            if (cur > 0 && cur + s.n_frames > d_budget->cap) break;
        }
        if (d_budget) d_budget->in_flight.fetch_add(s.n_frames, std::memory_order_relaxed);
        launch_slot(s);
        d_launch_idx++;
    }
}

void rx_fec_stream_pool::launch_slot(flight_slot& s)
{
    stream_res& r = d_streams[d_next_stream++ % d_streams.size()];
    const int n = s.n_frames;
    cusat_stream_rx_frame recs[256];
    std::vector<cusat_stream_rx_frame> big;
    cusat_stream_rx_frame* rec = recs;
    if (n > 256) {
        big.resize(n);
        rec = big.data();
    }
    for (int i = 0; i < n; i++) {
        const modcod_state& m = d_mc[s.mc[i]];
        rec[i].code = m.code;
        rec[i].flags = d_gpu_dd ? CUSAT_RX_SCALE_STATE : 0;  // the state after the batch's events
        rec[i].sym_offset = s.sym_off[i];
        rec[i].out_offset = s.out_off[i];
        rec[i].noise_var = m.noise_var;  // host mode: the MODCOD's scaling at launch, as a chain's
        rec[i].llr_gain = m.llr_gain;
    }
    {
        std::unique_lock<std::mutex> lk;
        if (d_mtx) lk = std::unique_lock<std::mutex>(*d_mtx);
        hipStream_t st = r.stream;
        hip_check(hipMemcpyAsync(r.sym_dev, s.syms, (size_t)s.n_syms * sizeof(complexf), hipMemcpyHostToDevice, st),
                  "copy in");
        abi_check(cusat_stream_rx_plan_async(r.sc, rec, n, s.events.data(), (int)s.events.size(), st), "plan");
        if (d_gpu_dd) {
            // the estimate needs only this micro-batch's symbols; the SNR
            // state passes from micro-batch to micro-batch in launch order
            hip_check(hipGraphLaunch(r.graph_dd_est, st), "graph launch");
            if (d_dd_recorded) hip_check(hipStreamWaitEvent(st, d_dd_done, 0), "wait");
            hip_check(hipGraphLaunch(r.graph_dd, st), "graph launch");
            hip_check(hipEventRecord(d_dd_done, st), "event");
            d_dd_recorded = true;
        }
        hip_check(hipGraphLaunch(r.graph, st), "graph launch");
        // results, SNR state and data fields in one copy
        hip_check(hipMemcpyAsync(s.blk, r.blk, d_res_bytes + d_state_bytes + (size_t)s.n_out, hipMemcpyDeviceToHost,
                                 st),
                  "copy out");
        hip_check(hipEventRecord(s.event, st), "event");
    }
    s.in_flight = true;
}

void rx_fec_stream_pool::post_process(flight_slot& s, block*& cur)
{
    const size_t mb = (size_t)d_micro_batch;
    const int32_t* trials = s.res;
    const int32_t* valid = s.res + mb;
    const int32_t* dfl = s.res + 2 * mb;
    const int32_t* fallback = s.res + 3 * mb;
    if (d_gpu_dd)
        for (auto& m : d_mc) {
            // the state after this micro-batch, unless an update_snr came later
            if (s.batch < m.host_event_batch) continue;
            const float* st = s.state + 4 * m.code;
            m.noise_var = st[0];
            m.llr_gain = st[1];
            m.snr_db = std::isnan(st[3]) ? 10.0f * std::log10(st[2]) : st[3];
        }
    for (int i = 0; i < s.n_frames; i++) {
        modcod_state& m = d_mc[s.mc[i]];
        m.frames++;
        if (fallback[i] == 1) m.fallbacks++;
        if (!rx_fec_chain::frame_passes_fec_checks(trials[i], valid[i] == 1)) {
            m.errors++;
            continue;
        }
        const int n = dfl[i];  // data field of the descrambled BBFRAME (0: none)
        if (n <= 0) continue;
        if (!cur || cur->mc != s.mc[i]) {
            d_ready.emplace_back();  // deque: earlier blocks stay in place
            cur = &d_ready.back();
            cur->mc = s.mc[i];
            d_last_output = s.mc[i];
        }
        const uint8_t* p = s.out + s.out_off[i];
        cur->bytes.insert(cur->bytes.end(), p, p + n);
        cur->sizes.push_back(n);
    }
}

bool rx_fec_stream_pool::drain(bool wait_all, bool oldest_only)
{
    bool any = false;
    block* cur = nullptr;  // consecutive data fields of one MODCOD share a block
    for (;;) {
        if (oldest_only && any) break;
        if (wait_all && d_drain_idx >= d_launch_idx && d_drain_idx < d_submit_idx)
            pump_launches(true);  // the oldest micro-batch may still wait for budget
        if (d_drain_idx >= d_launch_idx) break;
        flight_slot& s = d_flight[d_drain_idx % d_n_flight];
        if (s.in_flight) {
            if (wait_all) {
                hip_check(hipEventSynchronize(s.event), "drain");  // not under the launch mutex
            } else {
                const hipError_t q = hipEventQuery(s.event);
                if (q == hipErrorNotReady) break;  // keep the order
                hip_check(q, "event");
            }
            post_process(s, cur);
            s.in_flight = false;
            release_budget(s.n_frames);
            any = true;
        }
        d_drain_idx++;
    }
    pump_launches();
    return any;
}

int rx_fec_stream_pool::get_gse_output(uint8_t* out, int max_bytes)
{
    d_last_sizes.clear();
    if (d_ready.empty()) return 0;
    block b = std::move(d_ready.front());
    d_ready.pop_front();
    const int n = std::min<int>((int)b.bytes.size(), std::max(max_bytes, 0));
    if (n > 0) std::copy_n(b.bytes.data(), n, out);
    d_last_output = b.mc;
    d_last_sizes.swap(b.sizes);
    return n;
}

bool rx_fec_stream_pool::launch_stale(std::chrono::milliseconds max_age)
{
    if (d_batch_idx == 0 || std::chrono::steady_clock::now() - d_batch_start_time < max_age) return false;
    close_batch();
    return true;
}

bool rx_fec_stream_pool::drain_all_completed()
{
    drain(false);
    return !d_ready.empty();
}

void rx_fec_stream_pool::finish_all_pending()
{
    close_batch();
    drain(true);
}

void rx_fec_stream_pool::reset_all()
{
    {
        std::unique_lock<std::mutex> lk;
        if (d_mtx) lk = std::unique_lock<std::mutex>(*d_mtx);
        for (; d_drain_idx < d_launch_idx; d_drain_idx++) {
            flight_slot& s = d_flight[d_drain_idx % d_n_flight];
            if (s.in_flight) {
                hip_check(hipEventSynchronize(s.event), "reset");
                release_budget(s.n_frames);
            }
            s.in_flight = false;
        }
    }
    d_submit_idx = d_launch_idx = d_drain_idx = d_next_stream = 0;
    d_batch_idx = 0;
    d_ready.clear();
    d_last_sizes.clear();
    d_last_output = -1;
    for (auto& m : d_mc) {
        m.snr_db = 0;
        m.noise_var = 1.0f;
        m.llr_gain = 4.0f;
        m.frames = m.errors = m.fallbacks = 0;
    }
    reset_state();  // every launch has completed
}

void rx_fec_stream_pool::warmup_all()
{
    for (const auto& m : d_mc) {
        std::vector<complexf> zeros(m.xfec, complexf(0.0f, 0.0f));
        pl_layout l;
        l.modcod = m.idx;
        submit(zeros.data(), (int)zeros.size(), l);
    }
    finish_all_pending();
    reset_all();
    d_gpu_drop_cnt = 0;
}

uint64_t rx_fec_stream_pool::total_frame_count() const
{
    uint64_t n = 0;
    for (const auto& m : d_mc) n += m.frames;
    return n;
}

uint64_t rx_fec_stream_pool::total_error_count() const
{
    uint64_t n = 0;
    for (const auto& m : d_mc) n += m.errors;
    return n;
}

bool rx_fec_stream_pool::set_learned_params(uint8_t modcod, const int16_t* beta, const int16_t* alpha, int n_iters)
{
    const auto it = d_by_idx.find(modcod);
    if (it == d_by_idx.end()) return false;
    // micro-batches in flight may still read the parameter buffer that the
    // call rewrites in place; the next plan uploads the new descriptors
    for (auto& r : d_streams) hip_check(hipStreamSynchronize(r.stream), "sync");
    abi_check(cusat_ldpc_set_params(d_ctx[it->second], n_iters > 0 ? beta : nullptr, n_iters > 0 ? alpha : nullptr,
                                    n_iters),
              "learned parameters");
    return true;
}

int rx_fec_stream_pool::layer_count(uint8_t modcod) const
{
    const auto it = d_by_idx.find(modcod);
    return it == d_by_idx.end() ? -1 : d_mc[it->second].q;
}

uint64_t rx_fec_stream_pool::learned_fallback_count() const
{
    uint64_t n = 0;
    for (const auto& m : d_mc) n += m.fallbacks;
    return n;
}

bool rx_fec_stream_pool::modcod_counters(uint8_t modcod, uint64_t& frames, uint64_t& errors) const
{
    auto it = d_by_idx.find(modcod);
    if (it == d_by_idx.end()) return false;
    frames = d_mc[it->second].frames;
    errors = d_mc[it->second].errors;
    return true;
}

float rx_fec_stream_pool::data_snr_db() const
{
    return d_last_output < 0 ? 0.0f : d_mc[d_last_output].snr_db;
}

}  // namespace cusat_dsp
