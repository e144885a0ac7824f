// ACM receive throughput of the two pools of the cusat_dsp API on one GPU:
// rx_fec_pool (one rx_fec_chain per MODCOD, a flush on every MODCOD change,
// rx_fec_pool.cc:55-77) and rx_fec_stream_pool (one ring of mixed
// micro-batches, each a graph replay of cusat_stream_rx).
//
// Workload: normal FECFRAMEs of the given MODCODs, each carrying a GSE data
// field (BBHEADER + payload, BB-scrambled), encoded, mapped and noised by the
// library's host transmitter at threshold + 1.5 dB; the frames arrive in runs
// of one MODCOD whose length is uniform in 1..run.  Timed: every submit, the
// output collection and finish_all_pending, from host symbols to GSE bytes
// (host copy into pinned slots, H2D, demap, LDPC, BCH, BB, D2H included).
//
// usage: pool_bench [frames] [run] [pool: chain|chain_same|stream|stream_hostdd|both] [micro_batch]
//                   [streams] [modcods,comma-separated]
// chain: rx_fec_pool with the reference's defaults (micro-batches of 4, 8
// streams per MODCOD); chain_same: rx_fec_pool with the given micro_batch
// and streams per MODCOD.
// The stream pool waits when its ring is full (set_block_when_full), so
// its rate is what it sustains; stream_hostdd estimates the SNR on the host.
// prints one JSON line per pool.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "cusat_dsp/rx_fec_pool.h"
#include "cusat_dsp/rx_fec_stream_pool.h"
#include "cusat_fec.h"

using namespace cusat_dsp;

namespace {

struct source {
    const modcod_def_t* m = nullptr;
    int xfec = 0, kbch = 0;
    std::vector<complexf> symbols;  // frames_per_source x xfec
};

constexpr int FRAMES_PER_SOURCE = 48;

source make_source(const modcod_def_t* m, uint32_t seed)
{
    cusat_fec_info inf;
    if (cusat_fec_query(FECFRAME_NORMAL, m->code_rate, m->constellation, 0, &inf) != CUSAT_OK) {
        std::fprintf(stderr, "%s\n", cusat_last_error());
        std::exit(2);
    }
    source s;
    s.m = m;
    s.xfec = inf.xfec_len;
    s.kbch = inf.k_bch / 8;
    std::vector<uint8_t> prbs(s.kbch);
    cusat_bb_prbs(prbs.data(), s.kbch);
    std::mt19937 rng(seed);
    std::vector<uint8_t> info((size_t)FRAMES_PER_SOURCE * s.kbch);
    for (int b = 0; b < FRAMES_PER_SOURCE; b++) {
        uint8_t* f = &info[(size_t)b * s.kbch];
        for (int i = 0; i < s.kbch; i++) f[i] = (uint8_t)rng();
        const int len = (int)(rng() % (s.kbch - 10));
        f[0] = (1 << 5) | (1 << 4) | 2;  // generic packetized, SIS, CCM, roll-off 0.20
        f[1] = f[2] = f[3] = 0;
        f[4] = (uint8_t)((8 * len) >> 8);
        f[5] = (uint8_t)(8 * len);
        f[6] = 0;
        f[7] = f[8] = 0xff;
        f[9] = (uint8_t)cusat_bb_crc8(f, 9);
        for (int i = 0; i < s.kbch; i++) f[i] ^= prbs[i];
    }
    s.symbols.resize((size_t)FRAMES_PER_SOURCE * s.xfec);
    if (cusat_encode_frames(FECFRAME_NORMAL, m->code_rate, m->constellation, seed, FRAMES_PER_SOURCE,
                            m->min_snr_db + 1.5f, info.data(), reinterpret_cast<float*>(s.symbols.data()),
                            nullptr) != CUSAT_OK) {
        std::fprintf(stderr, "%s\n", cusat_last_error());
        std::exit(2);
    }
    return s;
}

struct arrival {
    int src, frame;
};

template <class P>
void run(P& pool, const char* name, const std::vector<source>& src, const std::vector<arrival>& arr, int run_len,
         int micro_batch, int streams)
{
    for (const auto& s : src) pool.update_snr(s.m->modcod_idx, s.m->min_snr_db + 1.5f);
    pool.warmup_all();
    for (const auto& s : src) pool.update_snr(s.m->modcod_idx, s.m->min_snr_db + 1.5f);
    std::vector<uint8_t> buf(pool.max_output_bytes());
    uint64_t out_bytes = 0, info_bits = 0;
    auto collect = [&] {
        for (;;) {
            const int n = pool.get_gse_output(buf.data(), (int)buf.size());
            if (n <= 0) break;
            out_bytes += (uint64_t)n;
        }
    };
    pl_layout l;
    double t_submit = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (const auto& a : arr) {
        const source& s = src[a.src];
        l.modcod = s.m->modcod_idx;
        const auto ts = std::chrono::steady_clock::now();
        const bool ready = pool.submit(&s.symbols[(size_t)a.frame * s.xfec], s.xfec, l);
        t_submit += std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
        if (ready) collect();
        info_bits += 8ull * s.kbch;
    }
    pool.finish_all_pending();
    collect();
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"pool\": \"%s\", \"frames\": %zu, \"modcods\": %zu, \"max_run\": %d, \"micro_batch\": %d, "
                "\"streams\": %d, \"seconds\": %.4f, \"seconds_in_submit\": %.4f, \"frames_per_s\": %.1f, "
                "\"gbps_info\": %.4f, \"gse_bytes\": %llu, \"frame_errors\": %llu, \"drops\": %llu}\n",
                name, arr.size(), src.size(), run_len, micro_batch, streams, sec, t_submit, arr.size() / sec,
                info_bits / sec * 1e-9, (unsigned long long)out_bytes,
                (unsigned long long)pool.total_error_count(), (unsigned long long)pool.total_gpu_drop_count());
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv)
{
    const int n_frames = argc > 1 ? std::atoi(argv[1]) : 3072;
    const int run_len = argc > 2 ? std::atoi(argv[2]) : 8;
    const std::string which = argc > 3 ? argv[3] : "both";
    const int micro_batch = argc > 4 ? std::atoi(argv[4]) : 128;
    const int streams = argc > 5 ? std::atoi(argv[5]) : 4;
    // QPSK 1/4, 1/2, 3/4, 2/3; 8PSK 3/5, 9/10, 3/4; 16APSK 2/3, 5/6, 3/4; 32APSK 3/4, 9/10
    std::vector<uint8_t> ids = {1, 4, 7, 6, 12, 17, 14, 18, 21, 19, 24, 28};
    if (argc > 6) {
        ids.clear();
        std::stringstream ss(argv[6]);
        std::string tok;
        while (std::getline(ss, tok, ',')) ids.push_back((uint8_t)std::atoi(tok.c_str()));
    }
    if (n_frames < 1 || run_len < 1 || micro_batch < 1 || streams < 1 || ids.empty()) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    std::vector<source> src;
    for (size_t k = 0; k < ids.size(); k++) {
        const modcod_def_t* m = find_modcod(ids[k]);
        if (!m) {
            std::fprintf(stderr, "unknown MODCOD %d\n", ids[k]);
            return 2;
        }
        src.push_back(make_source(m, 100 + (uint32_t)k));
    }
    std::mt19937 rng(7);
    std::vector<arrival> arr;
    int next = 0;
    while ((int)arr.size() < n_frames) {
        const int k = (int)(rng() % src.size());
        const int r = 1 + (int)(rng() % run_len);
        for (int j = 0; j < r && (int)arr.size() < n_frames; j++) arr.push_back({k, next++ % FRAMES_PER_SOURCE});
    }
    std::printf("# %zu frames generated\n", arr.size());
    std::fflush(stdout);
    {
        // floor of any pool taking host symbols: one memcpy of every frame
        // into pinned memory (a ring of 128 frame slots), on this thread
        size_t max_xfec = 0, bytes = 0;
        for (const auto& s : src) max_xfec = std::max<size_t>(max_xfec, s.xfec);
        complexf* pinned = nullptr;
        if (hipHostMalloc((void**)&pinned, 128 * max_xfec * sizeof(complexf), hipHostMallocDefault) != hipSuccess) {
            std::fprintf(stderr, "hipHostMalloc failed\n");
            return 2;
        }
        for (int pass = 0; pass < 2; pass++) {  // the first pass touches the pages
            const auto t0 = std::chrono::steady_clock::now();
            size_t k = 0;
            bytes = 0;
            for (const auto& a : arr) {
                const source& s = src[a.src];
                memcpy(pinned + (k++ % 128) * max_xfec, &s.symbols[(size_t)a.frame * s.xfec], s.xfec * sizeof(complexf));
                bytes += s.xfec * sizeof(complexf);
            }
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (pass == 1)
                std::printf("{\"host_copy_floor\": true, \"frames\": %zu, \"bytes\": %zu, \"seconds\": %.4f, "
                            "\"frames_per_s\": %.1f, \"gb_per_s\": %.2f}\n",
                            arr.size(), bytes, sec, arr.size() / sec, bytes / sec * 1e-9);
        }
        hipHostFree(pinned);
        std::fflush(stdout);
    }
    {
        // host -> device copy rate of micro-batch sized blocks (32 MiB) from
        // pinned memory of each allocation kind, on one stream
        const size_t blk = 32u << 20;
        const struct {
            unsigned flags;
            const char* name;
        } kinds[] = {{hipHostMallocDefault, "default"},
                     {hipHostMallocNonCoherent, "non_coherent"},
                     {hipHostMallocCoherent, "coherent"}};
        void* dev = nullptr;
        hipStream_t st = nullptr;
        if (hipMalloc(&dev, blk) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            std::fprintf(stderr, "device buffer failed\n");
            return 2;
        }
        for (const auto& k : kinds) {
            void* h = nullptr;
            if (hipHostMalloc(&h, blk, k.flags) != hipSuccess) continue;
            memset(h, 1, blk);
            double host_s = 0, total_s = 0;
            for (int pass = 0; pass < 2; pass++) {
                const auto t0 = std::chrono::steady_clock::now();
                for (int r = 0; r < 16; r++) hipMemcpyAsync(dev, h, blk, hipMemcpyHostToDevice, st);
                const auto t1 = std::chrono::steady_clock::now();
                hipStreamSynchronize(st);
                const auto t2 = std::chrono::steady_clock::now();
                host_s = std::chrono::duration<double>(t1 - t0).count();
                total_s = std::chrono::duration<double>(t2 - t0).count();
            }
            std::printf("{\"h2d\": \"%s\", \"bytes\": %zu, \"seconds\": %.4f, \"gb_per_s\": %.2f, "
                        "\"seconds_in_calls\": %.4f}\n",
                        k.name, 16 * blk, total_s, 16 * blk / total_s * 1e-9, host_s);
            hipHostFree(h);
        }
        hipStreamDestroy(st);
        hipFree(dev);
        std::fflush(stdout);
    }
    if (which == "chain" || which == "both") {
        // the reference's defaults: micro-batches of 4 over 8 streams per MODCOD
        rx_fec_pool pool(FECFRAME_NORMAL, ids, 25, 64, 8, 4);
        run(pool, "rx_fec_pool", src, arr, run_len, 4, 8);
    }
    if (which == "chain_same") {
        rx_fec_pool pool(FECFRAME_NORMAL, ids, 25, 64, streams, micro_batch);
        run(pool, "rx_fec_pool", src, arr, run_len, micro_batch, streams);
    }
    if (which == "stream" || which == "stream_hostdd" || which == "both") {
        rx_fec_stream_pool pool(FECFRAME_NORMAL, ids, 25, 64, streams, micro_batch);
        pool.set_block_when_full(true);
        const bool host_dd = which == "stream_hostdd";
        pool.set_gpu_dd_snr(!host_dd);
        run(pool, host_dd ? "rx_fec_stream_pool (host DD-SNR)" : "rx_fec_stream_pool", src, arr, run_len,
            micro_batch, streams);
    }
    return 0;
}
