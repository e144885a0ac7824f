// cusat_dsp::rx_fec_chain over the C ABI.  The flight-ring behaviour
// follows lib/cusat_dsp/rx_fec_chain.cc (submit / launch / drain order,
// decode budget, drops when the ring is full, block-when-full, the SNR
// re-estimate every 32 frames); the device work of a micro-batch is one
// replay of a captured cusat_stream_rx graph on the slot's stream.
#include "cusat_dsp/rx_fec_chain.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
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
        throw std::runtime_error(std::string("rx_fec_chain: ") + what + ": " + hipGetErrorString(e));
}

void abi_check(int rc, const char* what)
{
    if (rc != CUSAT_OK) throw std::runtime_error(std::string("rx_fec_chain: ") + what + ": " + cusat_last_error());
}

size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

float snr_ratio(float signal, float error)
{
    if (!std::isfinite(signal) || !std::isfinite(error)) return 0.0f;
    const float cap = 1e6f;  // error-free input reads as 60 dB
    return error * cap <= signal ? cap : signal / error;
}

}  // namespace

float qpsk_snr_dd(const complexf* sym, int n_sym)
{
    const float a = 0.70710678118654752f;
    float signal = 0.0f, error = 0.0f;
    for (int i = 0; i < n_sym; i++) {
        const float re = sym[i].real(), im = sym[i].imag();
        const float dr = re < 0 ? -a : a, di = im < 0 ? -a : a;
        signal += dr * dr + di * di;
        error += (re - dr) * (re - dr);
        error += (im - di) * (im - di);
    }
    return snr_ratio(signal, error);
}

float constellation_snr_dd(const complexf* points, int n_points, const complexf* sym, int n_sym, int stride)
{
    float signal = 0.0f, error = 0.0f;
    for (int i = 0; i < n_sym; i += stride) {
        int best = 0;
        float bd = std::norm(sym[i] - points[0]);
        for (int k = 1; k < n_points; k++) {
            const float d = std::norm(sym[i] - points[k]);
            if (d < bd) {
                bd = d;
                best = k;
            }
        }
        signal += std::norm(points[best]);
        error += bd;
    }
    return snr_ratio(signal, error);
}

std::shared_ptr<gpu_decode_budget> gpu_decode_budget::make_default()
{
    auto b = std::make_shared<gpu_decode_budget>();
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 0;
    b->cap = cus;  // one decode workgroup per compute unit
    if (const char* e = getenv("CUSAT_LDPC_INFLIGHT_CAP")) b->cap = atoi(e);
    return b;
}

rx_fec_chain::rx_fec_chain(fecframe_size_t framesize, code_rate_t rate, modulation_t constellation, int max_trials,
                           int cuda_max_batch, int n_gpu_streams, int micro_batch)
    : d_modulation(constellation), d_micro_batch(micro_batch), d_n_streams(n_gpu_streams)
{
    (void)cuda_max_batch;  // the chain decodes micro-batches; this sized the reference's spare decoder
    cusat_fec_info inf;
    if (cusat_fec_query((int)framesize, (int)rate, (int)constellation, 0, &inf) != CUSAT_OK)
        throw std::invalid_argument(std::string("rx_fec_chain: ") + cusat_last_error());
    if (micro_batch < 1 || n_gpu_streams < 1 || max_trials < 0)
        throw std::invalid_argument("rx_fec_chain: micro_batch and n_gpu_streams must be >= 1");
    d_kbch_bytes = inf.k_bch / 8;
    d_max_dfl = d_kbch_bytes - 10;
    d_kldpc_bytes = inf.k_ldpc / 8;
    d_nldpc = inf.n_ldpc;
    d_xfecframe_len = inf.xfec_len;
    d_n_mod = inf.n_mod;
    {
        std::vector<float> pts(2u << inf.n_mod);
        uint32_t cols[8];
        int nm = 0, rows = 0;
        if (cusat_modcod_constellation((int)framesize, (int)rate, (int)constellation, pts.data(), cols, &nm, &rows) !=
            CUSAT_OK)
            throw std::invalid_argument(std::string("rx_fec_chain: ") + cusat_last_error());
        for (size_t k = 0; k < pts.size() / 2; k++) d_points.emplace_back(pts[2 * k], pts[2 * k + 1]);
    }
    d_batch_size = N_FLIGHT * d_micro_batch;
    d_output.resize((size_t)d_kbch_bytes * d_batch_size);
    d_flight.resize(N_FLIGHT);
    d_streams.resize(d_n_streams);
    // frames start on 64-byte boundaries in the slots and on the device
    d_sym_stride = (d_xfecframe_len + 7) & ~7;
    const size_t mb = (size_t)d_micro_batch;
    d_res_bytes = round256(3 * mb * sizeof(int32_t));
    const size_t sym_bytes = mb * d_sym_stride * sizeof(complexf);
    const size_t blk_bytes = d_res_bytes + mb * d_max_dfl;
    try {
        abi_check(cusat_fec_ctx_create(&d_ctx, (int)framesize, (int)rate, (int)constellation, max_trials, 1, 0),
                  "context");
        for (auto& r : d_streams) {
            hipStream_t st = nullptr;
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
            r.stream = st;
            abi_check(cusat_stream_create(&r.sc, &d_ctx, 1, d_micro_batch, 0), "stream context");
            hip_check(hipMalloc(&r.sym_dev, sym_bytes), "device buffers");
            hip_check(hipMalloc(&r.blk, blk_bytes), "device buffers");
            // capture the three receive launches once; every micro-batch
            // replays them with its own plan
            abi_check(cusat_stream_rx_plan_async(r.sc, nullptr, 0, nullptr, 0, st), "plan");
            int32_t* res = reinterpret_cast<int32_t*>(r.blk);
            const float* sym = reinterpret_cast<const float*>(r.sym_dev);
            auto rx = [&] {
                return cusat_stream_rx(r.sc, sym, r.blk + d_res_bytes, res, res + mb, res + 2 * mb, st);
            };
            abi_check(rx(), "rx");
            hip_check(hipStreamSynchronize(st), "warm launch");
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture");
            const int rc = rx();
            const hipError_t ec = hipStreamEndCapture(st, &g);
            abi_check(rc, "capture");
            hip_check(ec, "capture");
            hipGraphExec_t ge = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            hip_check(ei, "graph instantiate");
            r.graph = ge;
        }
        for (auto& f : d_flight) {
            hip_check(hipHostMalloc((void**)&f.syms, sym_bytes, hipHostMallocDefault), "pinned slots");
            hip_check(hipHostMalloc((void**)&f.blk, blk_bytes, hipHostMallocDefault), "pinned slots");
            f.trials = reinterpret_cast<int32_t*>(f.blk);
            f.valid = f.trials + mb;
            f.dfl = f.trials + 2 * mb;
            f.gse = f.blk + d_res_bytes;
            // touch every page now: a first write into pinned memory costs a
            // fault per page, which would land in submit_xfecframe
            memset(f.syms, 0, sym_bytes);
            memset(f.blk, 0, blk_bytes);
            hipEvent_t ev = nullptr;
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "events");
            f.event = ev;
        }
    } catch (...) {
        free_all();
        throw;
    }
    d_budget = gpu_decode_budget::make_default();
}

void rx_fec_chain::free_all()
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
        if (r.sc) cusat_stream_destroy(r.sc);
        if (r.sym_dev) hipFree(r.sym_dev);
        if (r.blk) hipFree(r.blk);
        if (r.stream) hipStreamDestroy(r.stream);
        r = stream_res{};
    }
    if (d_ctx) cusat_fec_ctx_destroy(d_ctx);
    d_ctx = nullptr;
}

rx_fec_chain::~rx_fec_chain()
{
    // frames still in flight hold budget shared with other chains
    while (d_drain_idx < d_launch_idx) {
        flight_slot& s = d_flight[d_drain_idx % N_FLIGHT];
        if (s.in_flight) {
            hipEventSynchronize(s.event);
            release_budget(s.n_frames);
            s.in_flight = false;
        }
        d_drain_idx++;
    }
    free_all();
}

void rx_fec_chain::reset()
{
    {
        std::unique_lock<std::mutex> lk;
        if (d_mtx) lk = std::unique_lock<std::mutex>(*d_mtx);
        for (; d_drain_idx < d_submit_idx; d_drain_idx++) {
            flight_slot& s = d_flight[d_drain_idx % N_FLIGHT];
            if (s.in_flight) {
                hip_check(hipEventSynchronize(s.event), "reset");
                release_budget(s.n_frames);
            }
            s.in_flight = false;
            s.stream_idx = -1;
            s.n_frames = 0;
        }
    }
    d_submit_idx = d_launch_idx = d_drain_idx = d_next_stream = 0;
    d_batch_idx = 0;
    d_output_len = 0;
    d_snr_db = 0;
    d_noise_var = 1.0f;
    d_llr_gain = 4.0f;
    d_frame_cnt = d_frame_error_cnt = 0;
    d_gse_dfl_sizes.clear();
}

bool rx_fec_chain::submit_xfecframe(const complexf* symbols, int n_symbols)
{
    bool ready = false;
    // drain at micro-batch boundaries, when the ring fills up and while a
    // closed micro-batch waits for decode budget (a drain releases budget)
    if (d_batch_idx == 0 || d_launch_idx < d_submit_idx || d_submit_idx - d_drain_idx >= N_FLIGHT / 2) {
        d_output_len = 0;
        d_gse_dfl_sizes.clear();
        ready = drain_completed();
    }
    while (d_block_when_full && d_submit_idx - d_drain_idx >= N_FLIGHT - 2) {
        force_drain_oldest();
        ready = true;
    }
    if (d_submit_idx - d_drain_idx >= N_FLIGHT - 2 || n_symbols != d_xfecframe_len || !symbols) {
        d_gpu_drop_cnt++;
        d_frame_cnt++;
        d_frame_error_cnt++;
        return ready;
    }
    flight_slot& slot = d_flight[d_submit_idx % N_FLIGHT];
    complexf* dst = slot.syms + (size_t)d_batch_idx * d_sym_stride;
    detail::copy_frame(dst, symbols, (size_t)n_symbols * sizeof(complexf));
    if (++d_snr_skip_count >= 32) {
        d_snr_skip_count = 0;
        update_dd_snr(symbols);  // the caller's copy: pinned memory is slow to read from the host
    }
    if (d_batch_idx == 0) d_batch_start_time = std::chrono::steady_clock::now();
    if (++d_batch_idx >= d_micro_batch) {
        slot.n_frames = d_batch_idx;
        d_submit_idx++;
        d_batch_idx = 0;
        pump_launches();
    }
    return ready;
}

void rx_fec_chain::update_dd_snr(const complexf* frame)
{
    float snr_lin;
    if (d_n_mod >= 7 && d_pilot_snr_db > 5.0f)  // dense APSK: trust the pilot estimate
        snr_lin = std::pow(10.0f, d_pilot_snr_db / 10.0f);
    else if (d_modulation == MOD_QPSK)
        snr_lin = qpsk_snr_dd(frame, d_xfecframe_len);
    else
        snr_lin = constellation_snr_dd(d_points.data(), (int)d_points.size(), frame, d_xfecframe_len, 8);
    if (snr_lin > 1.0f) {
        d_snr_db = 10.0f * std::log10(snr_lin);
        d_noise_var = 1.0f / snr_lin;
        d_llr_gain = std::fmin(4.0f / d_noise_var, 1000.0f);
    }
}

void rx_fec_chain::launch_partial()
{
    if (d_batch_idx > 0) {
        d_flight[d_submit_idx % N_FLIGHT].n_frames = d_batch_idx;
        d_submit_idx++;
        d_batch_idx = 0;
    }
    pump_launches();
}

void rx_fec_chain::warmup()
{
    std::vector<complexf> zeros(d_xfecframe_len, complexf(0.0f, 0.0f));
    submit_xfecframe(zeros.data(), (int)zeros.size());
    launch_partial();
    flush();
    reset();
    d_force_drain_cnt = 0;
    d_gpu_drop_cnt = 0;
}

void rx_fec_chain::flush()
{
    d_output_len = 0;
    d_gse_dfl_sizes.clear();
    launch_partial();
    while (d_drain_idx < d_submit_idx) force_drain_oldest();
}

bool rx_fec_chain::drain_completed()
{
    d_output_len = 0;
    d_gse_dfl_sizes.clear();
    bool any = false;
    while (d_drain_idx < d_launch_idx) {
        flight_slot& s = d_flight[d_drain_idx % N_FLIGHT];
        if (!s.in_flight) {
            d_drain_idx++;
            continue;
        }
        const hipError_t q = hipEventQuery(s.event);
        if (q == hipErrorNotReady) break;  // keep the order
        hip_check(q, "event");
        post_process_slot(s);
        s.in_flight = false;
        release_budget(s.n_frames);
        d_drain_idx++;
        any = true;
    }
    pump_launches();
    return any;
}

int rx_fec_chain::get_gse_output(uint8_t* out, int max_bytes)
{
    const int n = std::max(0, std::min(d_output_len, max_bytes));
    if (n) memcpy(out, d_output.data(), (size_t)n);
    d_output_len = 0;
    return n;
}

void rx_fec_chain::pump_launches(bool bypass_cap)
{
    while (d_launch_idx < d_submit_idx) {
        flight_slot& s = d_flight[d_launch_idx % N_FLIGHT];
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

void rx_fec_chain::launch_slot(flight_slot& slot)
{
    const int si = d_next_stream++ % d_n_streams;
    stream_res& r = d_streams[si];
    const int n = slot.n_frames;
    cusat_stream_rx_frame recs[64];
    std::vector<cusat_stream_rx_frame> big;
    cusat_stream_rx_frame* rec = recs;
    if (n > 64) {
        big.resize(n);
        rec = big.data();
    }
    for (int i = 0; i < n; i++) {
        rec[i].code = 0;
        rec[i].flags = 0;
        rec[i].sym_offset = (int64_t)i * d_sym_stride;
        rec[i].out_offset = (int64_t)i * d_max_dfl;
        rec[i].noise_var = d_noise_var;  // the chain's scaling when the micro-batch launches
        rec[i].llr_gain = d_llr_gain;
    }
    {
        std::unique_lock<std::mutex> lk;
        if (d_mtx) lk = std::unique_lock<std::mutex>(*d_mtx);
        hipStream_t st = r.stream;
        const size_t sym_bytes = ((size_t)(n - 1) * d_sym_stride + d_xfecframe_len) * sizeof(complexf);
        hip_check(hipMemcpyAsync(r.sym_dev, slot.syms, sym_bytes, hipMemcpyHostToDevice, st), "copy in");
        abi_check(cusat_stream_rx_plan_async(r.sc, rec, n, nullptr, 0, st), "plan");
        hip_check(hipGraphLaunch(r.graph, st), "graph launch");
        // trials, valid, data-field lengths and data fields in one copy
        hip_check(hipMemcpyAsync(slot.blk, r.blk, d_res_bytes + (size_t)n * d_max_dfl, hipMemcpyDeviceToHost, st),
                  "copy out");
        hip_check(hipEventRecord(slot.event, st), "event");
    }
    slot.in_flight = true;
    slot.stream_idx = si;
}

void rx_fec_chain::force_drain_oldest()
{
    // the oldest slot may still wait for budget: launch it regardless
    if (d_drain_idx >= d_launch_idx && d_drain_idx < d_submit_idx) pump_launches(true);
    flight_slot& s = d_flight[d_drain_idx % N_FLIGHT];
    if (!s.in_flight) {
        d_drain_idx++;
        return;
    }
    d_force_drain_cnt++;
    hip_check(hipEventSynchronize(s.event), "drain");  // not under the launch mutex
    post_process_slot(s);
    s.in_flight = false;
    release_budget(s.n_frames);
    d_drain_idx++;
}

void rx_fec_chain::post_process_slot(flight_slot& slot)
{
    for (int i = 0; i < slot.n_frames; i++) {
        d_frame_cnt++;
        if (!frame_passes_fec_checks(slot.trials[i], slot.valid[i] == 1)) {
            d_frame_error_cnt++;
            continue;
        }
        const int n = slot.dfl[i];  // data field of the descrambled BBFRAME (0: none)
        if (n <= 0 || n > (int)d_output.size() - d_output_len) continue;
        memcpy(d_output.data() + d_output_len, slot.gse + (size_t)i * d_max_dfl, (size_t)n);
        d_gse_dfl_sizes.push_back(n);
        d_output_len += n;
    }
}

}  // namespace cusat_dsp
