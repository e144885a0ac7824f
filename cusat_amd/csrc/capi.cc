// C ABI (include/cusat_fec.h): contexts, device tables, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bb.h"
#include "bch.h"
#include "codes.h"
#include "ctx.h"
#include "cusat_fec.h"
#include "kernels.h"
#include "modulation.h"

namespace cusat {

static thread_local char g_err[512] = "";

void set_last_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

ldpc_args make_ldpc_args(const cusat_fec_ctx* c, const int16_t* llr, uint8_t* hard, int32_t* trials)
{
    ldpc_args a{};
    a.llr = llr;
    a.msg = c->d_msg;
    a.hard = hard;
    a.result = trials;
    a.circ = c->d_circ;
    a.lay_cnt = c->d_lay_cnt;
    a.lay_off = c->d_lay_off;
    a.lay_moff = c->d_lay_moff;
    a.steps = c->d_steps;
    a.grp_off = c->d_grp_off;
    a.step_defer = c->d_step_defer;
    a.groups = c->d_groups;
    a.N = c->geo.N;
    a.K = c->geo.K;
    a.R = c->geo.R;
    a.q = c->geo.q;
    a.n_steps = c->kern_steps;
    a.lanes = c->geo.lanes;
    a.n_slots = c->geo.n_slots;
    a.n_groups = (int)c->geo.groups.size();
    a.lt_alloc = c->geo.lt_alloc;
    a.max_trials = c->max_trials;
    a.cmax = c->cmax;
    a.exact = c->exact;
    a.split = !(c->flags & CUSAT_CTX_NO_SPLIT_KERNEL);
    return a;
}

}  // namespace cusat

using namespace cusat;


namespace {

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            set_last_error("%s: %s", #expr, hipGetErrorString(e_));                   \
            return CUSAT_E_HIP;                                                       \
        }                                                                             \
    } while (0)

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int upload_tables(cusat_fec_ctx* c)
{
    const decoder_geometry& g = c->geo;
    const int W = (c->bch.degree + 63) / 64;
    const size_t n_bytes = c->fp.n_bch / 8;
    std::vector<uint64_t> lut(256 * W), pw((size_t)64 * c->bch.degree * W, 0);
    for (int b = 0; b < 256; b++)
        for (int w = 0; w < W; w++) lut[b * W + w] = c->bch.lut[b][w];
    c->bch_chunk = (int)((n_bytes + 63) / 64);
    const std::vector<poly256> powers = bch_powers(c->bch, n_bytes * 8 + c->bch.degree + 1);
    for (int l = 0; l < 64; l++) {
        const size_t begin = (size_t)l * c->bch_chunk;
        if (begin >= n_bytes) continue;
        const size_t end = std::min(n_bytes, begin + c->bch_chunk);
        const size_t shift = (n_bytes - end) * 8;
        for (int b = 0; b < c->bch.degree; b++)
            for (int w = 0; w < W; w++) pw[((size_t)b * W + w) * 64 + l] = powers[shift + b][w];  // [bit][word][lane]
    }

    // schedule as int32 (scalar loads), padded to an even step count with an
    // idle step (the kernel double-buffers message records over step pairs)
    std::vector<int> steps32(g.steps.begin(), g.steps.end());
    std::vector<int> grp_off = g.grp_off;
    if (g.n_steps & 1) {
        steps32.push_back(-1);
        steps32.push_back(-1);
        grp_off.push_back(grp_off.back());
    }
    c->kern_steps = (int)steps32.size() / 2;
    std::vector<int> defer32(c->kern_steps, 0);
    for (int st = 0; st < g.n_steps; st++) defer32[st] = g.step_defer[st];
    // collision groups as int32 records: LDS dword base of the VN group,
    // slot count, slot dword bases (slots follow app + parity in LDS)
    std::vector<int> grp32;
    const int slot_base = (g.K + g.R) / 2;
    for (const collision_group& cg : g.groups) {
        int rec[8] = {180 * cg.row, cg.n, 0, 0, 0, 0, 0, 0};
        for (int e = 0; e < cg.n; e++) rec[2 + e] = slot_base + 180 * cg.slot[e];
        grp32.insert(grp32.end(), rec, rec + 8);
    }

    std::vector<uint8_t> prbs(c->fp.k_bch / 8);
    bb_prbs(prbs.data(), prbs.size());

    struct piece {
        const void* src;
        size_t bytes;
        void** dst;
    };
    std::vector<piece> pieces = {
        {g.circ.data(), g.circ.size() * 4, (void**)&c->d_circ},
        {g.lay_cnt.data(), g.lay_cnt.size() * 4, (void**)&c->d_lay_cnt},
        {g.lay_off.data(), g.lay_off.size() * 4, (void**)&c->d_lay_off},
        {g.lay_moff.data(), g.lay_moff.size() * 4, (void**)&c->d_lay_moff},
        {steps32.data(), steps32.size() * 4, (void**)&c->d_steps},
        {defer32.data(), defer32.size() * 4, (void**)&c->d_step_defer},
        {grp_off.data(), grp_off.size() * 4, (void**)&c->d_grp_off},
        {grp32.data(), grp32.size() * 4, (void**)&c->d_groups},
        {c->con.xy.data(), (size_t)c->con.n_points * 2 * 4, (void**)&c->d_pts},
        {lut.data(), lut.size() * 8, (void**)&c->d_bch_lut},
        {pw.data(), pw.size() * 8, (void**)&c->d_bch_pw},
        {prbs.data(), prbs.size(), (void**)&c->d_prbs},
    };
    size_t total = 0;
    for (auto& p : pieces) total += align_up(std::max<size_t>(p.bytes, 16));
    std::vector<unsigned char> host(total, 0);
    size_t off = 0;
    std::vector<size_t> offs;
    for (auto& p : pieces) {
        if (p.bytes) memcpy(host.data() + off, p.src, p.bytes);
        offs.push_back(off);
        off += align_up(std::max<size_t>(p.bytes, 16));
    }
    HIP_TRY(hipMalloc(&c->d_tables, total));
    HIP_TRY(hipMemcpy(c->d_tables, host.data(), total, hipMemcpyHostToDevice));
    for (size_t i = 0; i < pieces.size(); i++) *pieces[i].dst = (unsigned char*)c->d_tables + offs[i];
    return CUSAT_OK;
}

bch_args make_bch_args(const cusat_fec_ctx* c, const uint8_t* hard, int32_t* valid)
{
    bch_args a{};
    a.hard = hard;
    a.valid = valid;
    a.lut = c->d_bch_lut;
    a.pw = c->d_bch_pw;
    a.n_bytes = c->fp.n_bch / 8;
    a.degree = c->bch.degree;
    a.chunk = c->bch_chunk;
    return a;
}

void demap_consts(const cusat_fec_ctx* c, float noise_var, float llr_gain, float* qs, float* ps, float* g)
{
    // QPSK: 2 sqrt(2) / N0 (demapper_kernel.cu:64); others use the gain
    // capped at 2 * target / dmax (:283-291)
    *qs = 2.0f * 1.4142135623730951f / noise_var;
    float gain = llr_gain;
    if (c->llr_target > 0.0f && c->dmax > 0.0f) {
        const float cap = 2.0f * c->llr_target / c->dmax;
        gain = llr_gain < cap ? llr_gain : cap;
    }
    *g = gain;
    *ps = 2.0f * 0.38268343236508977f * gain;
}

bool check_batch(const cusat_fec_ctx* c, int batch)
{
    if (batch < 1 || batch > c->max_batch) {
        set_last_error("batch %d outside 1..%d", batch, c->max_batch);
        return false;
    }
    return true;
}

}  // namespace

extern "C" const char* cusat_last_error(void) { return g_err; }

namespace {

// Host-side set-up of a context: tables, geometry, constellation, BCH.
int prepare(cusat_fec_ctx* c, int framesize, int rate, int modulation, unsigned flags)
{
    c->framesize = framesize;
    c->flags = flags;
    c->rate = rate;
    c->modulation = modulation;
    const char* tname = ldpc_table_name(framesize, rate);
    const ldpc_table* tab = tname ? find_table(tname) : nullptr;
    if (!tab || !get_fec_params(framesize, rate, &c->fp) ||
        !make_constellation(modulation, rate, framesize, &c->con) ||
        !make_interleaver(modulation, rate, framesize, &c->il)) {
        set_last_error("unsupported (frame %d, rate %d, modulation %d)", framesize, rate, modulation);
        return CUSAT_E_INVALID;
    }
    if (!make_bch(framesize, c->fp.n_bch, c->fp.t, &c->bch) || c->bch.k_bits != c->fp.k_bch) {
        set_last_error("BCH geometry mismatch");
        return CUSAT_E_INVALID;
    }
    try {
        const bool sched = !(flags & CUSAT_CTX_SEQUENTIAL_SCHEDULE);
        c->geo = build_geometry(*tab, sched, true);
        // deferred collision groups need a second slot area: keep them only
        // where both decoders' LDS still fits in 160 KiB
        const int ks = c->geo.n_steps + (c->geo.n_steps & 1), ng = (int)c->geo.groups.size();
        const int lds_max = std::max(ldpc_lds_bytes(c->geo.K, c->geo.R, c->geo.n_slots, ks, ng),
                                     ldpc_lds_bytes_split(c->geo.K, c->geo.R, c->geo.n_slots, ks, ng, 2));
        if (c->geo.slot_parity_base > 0 && lds_max > 160 * 1024) c->geo = build_geometry(*tab, sched, false);
    } catch (const std::exception& e) {
        set_last_error("geometry: %s", e.what());
        return CUSAT_E_INVALID;
    }
    int cmin = 1 << 30, cmx = 0;
    for (int v : c->geo.lay_cnt) {
        cmin = std::min(cmin, v);
        cmx = std::max(cmx, v);
    }
    if (!ldpc_select_variant(cmin, cmx, c->geo.lanes, &c->cmax, &c->exact) || c->geo.n_slots > 63) {
        set_last_error("code exceeds kernel limits");
        return CUSAT_E_INVALID;
    }
    xfec_lengths(modulation, c->geo.N, c->con.bits_per_sym, &c->xfec, &c->dlen);
    c->dmax = llr_dmax(modulation, c->con);
    float rmax = 0.0f;
    for (int i = 0; i < c->con.n_points; i++)
        rmax = std::max(rmax, std::sqrt(c->con.xy[2 * i] * c->con.xy[2 * i] + c->con.xy[2 * i + 1] * c->con.xy[2 * i + 1]));
    c->grid_scale = demap_grid_scale(rmax);
    if (const char* e = getenv("CUSAT_DEMAP_LLR_TARGET")) c->llr_target = (float)atof(e);
    return CUSAT_OK;
}

void fill_info(const cusat_fec_ctx* c, cusat_fec_info* o)
{
    memset(o, 0, sizeof *o);
    o->n_ldpc = c->geo.N;
    o->k_ldpc = c->geo.K;
    o->k_bch = c->fp.k_bch;
    o->n_bch = c->fp.n_bch;
    o->bch_t = c->fp.t;
    o->bch_degree = c->bch.degree;
    o->q = c->geo.q;
    o->n_steps = c->geo.n_steps;
    o->lanes = c->geo.lanes;
    o->cnl = c->geo.cnl;
    o->lt_alloc = c->geo.lt_alloc;
    o->n_mod = c->con.bits_per_sym;
    o->xfec_len = c->xfec;
    o->demap_len = c->dlen;
    o->max_batch = c->max_batch;
    o->max_trials = c->max_trials;
    o->n_collision_slots = c->geo.n_slots;
    o->llr_dmax = c->dmax;
    o->llr_gain_cap = (c->llr_target > 0.0f && c->dmax > 0.0f) ? 2.0f * c->llr_target / c->dmax : 0.0f;
    const int ks = c->geo.n_steps + (c->geo.n_steps & 1), ng = (int)c->geo.groups.size();
    o->lds_bytes = (!(c->flags & CUSAT_CTX_NO_SPLIT_KERNEL) && ldpc_split_available(c->cmax, c->exact, c->geo.lanes))
                       ? ldpc_lds_bytes_split(c->geo.K, c->geo.R, c->geo.n_slots, ks, ng, c->geo.lanes)
                       : ldpc_lds_bytes(c->geo.K, c->geo.R, c->geo.n_slots, ks, ng);
}

}  // namespace

extern "C" int cusat_fec_query(int framesize, int rate, int modulation, unsigned flags, cusat_fec_info* info)
{
    if (!info) { set_last_error("null argument"); return CUSAT_E_INVALID; }
    cusat_fec_ctx c{};
    int rc = prepare(&c, framesize, rate, modulation, flags);
    if (rc == CUSAT_OK) fill_info(&c, info);
    return rc;
}

extern "C" int cusat_fec_ctx_create(cusat_fec_ctx** out, int framesize, int rate, int modulation,
                                    int max_trials, int max_batch, unsigned flags)
{
    if (!out) { set_last_error("null output pointer"); return CUSAT_E_INVALID; }
    *out = nullptr;
    if (max_trials < 0 || max_trials > 1000 || max_batch < 1 || max_batch > (1 << 20)) {
        set_last_error("max_trials %d / max_batch %d out of range", max_trials, max_batch);
        return CUSAT_E_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_last_error("no HIP device");
        return CUSAT_E_NODEVICE;
    }
    cusat_fec_ctx* c = new (std::nothrow) cusat_fec_ctx();
    if (!c) return CUSAT_E_NOMEM;
    c->max_trials = max_trials;
    c->max_batch = max_batch;
    int rc = prepare(c, framesize, rate, modulation, flags);
    if (rc != CUSAT_OK) {
        delete c;
        return rc;
    }
    hipGetDevice(&c->device);
    // set-up of cusat_ldpc_decode_noalloc on this device (the address of its
    // static scratch); a failure here shows again, as an error, in that call
    (void)ldpc_scratch_setup();
    rc = upload_tables(c);
    if (rc == CUSAT_OK) {
        const size_t rec = std::max(ldpc_msg_bytes(c->geo.q, c->cmax), ldpc_msg_bytes_split(c->geo.q, c->cmax));
        hipError_t e = hipMalloc(&c->d_msg, (size_t)max_batch * rec);
        if (e == hipSuccess) e = hipMalloc(&c->d_mask, (size_t)max_batch * sizeof(int32_t));
        if (e != hipSuccess) {
            set_last_error("message workspace: %s", hipGetErrorString(e));
            rc = CUSAT_E_NOMEM;
        }
    }
    if (rc != CUSAT_OK) {
        cusat_fec_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return CUSAT_OK;
}

extern "C" int cusat_fec_ctx_destroy(cusat_fec_ctx* c)
{
    if (!c) return CUSAT_OK;
    if (c->d_tables) hipFree(c->d_tables);
    if (c->d_msg) hipFree(c->d_msg);
    if (c->d_mask) hipFree(c->d_mask);
    if (c->d_beta) hipFree(c->d_beta);
    for (int* p : c->retired_params) hipFree(p);
    if (c->d_gen) hipFree(c->d_gen);
    delete c;
    return CUSAT_OK;
}

extern "C" int cusat_fec_ctx_info(const cusat_fec_ctx* c, cusat_fec_info* o)
{
    if (!c || !o) { set_last_error("null argument"); return CUSAT_E_INVALID; }
    fill_info(c, o);
    return CUSAT_OK;
}

extern "C" int cusat_fec_demap_scales(const cusat_fec_ctx* c, float noise_var, float llr_gain,
                                      float* qs, float* ps, float* g)
{
    if (!c || !qs || !ps || !g || !(noise_var > 0.0f)) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    demap_consts(c, noise_var, llr_gain, qs, ps, g);
    return CUSAT_OK;
}

extern "C" int cusat_fec_constellation(const cusat_fec_ctx* c, float* points, uint32_t* col)
{
    if (!c) { set_last_error("null context"); return CUSAT_E_INVALID; }
    if (points) memcpy(points, c->con.xy.data(), sizeof(float) * 2 * c->con.n_points);
    if (col) {
        if (c->il.n_cols == 0) { col[0] = 0; col[1] = 1; for (int b = 2; b < 8; b++) col[b] = 0; }
        else for (int b = 0; b < 8; b++) col[b] = c->il.col[b];
    }
    return CUSAT_OK;
}

extern "C" int cusat_fec_demap(cusat_fec_ctx* c, const float* sym, int16_t* llr, int batch,
                               float noise_var, float llr_gain, void* stream)
{
    if (!c || !sym || !llr || !(noise_var > 0.0f)) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    if (!check_batch(c, batch)) return CUSAT_E_BATCH;
    demap_args a{};
    a.sym = sym;
    a.llr = llr;
    const bool formula = c->modulation == CUSAT_MOD_QPSK || c->modulation == CUSAT_MOD_8PSK;
    a.pts = formula ? nullptr : c->d_pts;
    a.xfec = c->xfec;
    a.dlen = c->dlen;
    a.N = c->geo.N;
    a.n_mod = c->con.bits_per_sym;
    a.n_points = c->con.n_points;
    a.grid_scale = c->grid_scale;
    for (int b = 0; b < 8; b++) a.col[b] = c->il.col[b];
    demap_consts(c, noise_var, llr_gain, &a.qpsk_scale, &a.psk8_scale, &a.gain);
    HIP_TRY(launch_demap(a, batch, (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_ldpc_decode(cusat_fec_ctx* c, const int16_t* llr, uint8_t* hard, int32_t* trials,
                                 int batch, void* stream)
{
    if (!c || !llr || !hard || !trials) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    if (!check_batch(c, batch)) return CUSAT_E_BATCH;
    ldpc_args a = make_ldpc_args(c, llr, hard, trials);
    HIP_TRY(launch_ldpc_decode(a, batch, (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_bch_check(cusat_fec_ctx* c, const uint8_t* hard, int32_t* valid, int batch, void* stream)
{
    if (!c || !hard || !valid) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    if (!check_batch(c, batch)) return CUSAT_E_BATCH;
    bch_args a = make_bch_args(c, hard, valid);
    HIP_TRY(launch_bch_check(a, batch, (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_rx_fec_batch(cusat_fec_ctx* c, const float* sym, int16_t* llr, int batch, float noise_var,
                                  float llr_gain, uint8_t* hard, int32_t* trials, int32_t* valid, void* stream)
{
    int rc = cusat_fec_demap(c, sym, llr, batch, noise_var, llr_gain, stream);
    if (rc == CUSAT_OK) rc = cusat_ldpc_decode(c, llr, hard, trials, batch, stream);
    if (rc == CUSAT_OK) rc = cusat_bch_check(c, hard, valid, batch, stream);
    return rc;
}

extern "C" int cusat_ldpc_set_params(cusat_fec_ctx* c, const int16_t* beta, const int16_t* alpha, int n_iters)
{
    if (!c || n_iters < 0 || n_iters > 1000 || (n_iters > 0 && (!beta || !alpha))) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    const int q = c->geo.q;
    const size_t n = (size_t)n_iters * q;
    for (size_t k = 0; k < n; k++)
        if (beta[k] < 0 || alpha[k] < 0 || alpha[k] > 1023) {
            set_last_error("parameter %zu out of range (beta %d, alpha %d)", k, beta[k], alpha[k]);
            return CUSAT_E_INVALID;
        }
    if (n_iters > c->param_cap) {
        // the outgrown buffer stays allocated until the context is
        // destroyed: work enqueued before this call may still read it
        int* d = nullptr;
        HIP_TRY(hipMalloc(&d, 2 * n * sizeof(int)));
        if (c->d_beta) c->retired_params.push_back(c->d_beta);
        c->d_beta = d;
        c->d_alpha = d + n;
        c->param_cap = n_iters;
    }
    if (n_iters > 0) {
        // beta block then alpha block, at the current capacity's layout
        c->d_alpha = c->d_beta + (size_t)c->param_cap * q;
        std::vector<int> h((size_t)c->param_cap * q * 2, 0);
        // alpha word of the kernels: n | ceil(8096 / n) << 16 (0 = plain scale)
        for (size_t k = 0; k < n; k++) {
            h[k] = beta[k];
            const int a = alpha[k];
            h[(size_t)c->param_cap * q + k] = a ? (a | (((8096 + a - 1) / a) << 16)) : 0;
        }
        HIP_TRY(hipMemcpy(c->d_beta, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    c->n_param_iters = n_iters;
    c->param_gen++;
    return CUSAT_OK;
}

extern "C" int cusat_ldpc_decode_learned(cusat_fec_ctx* c, const int16_t* llr, uint8_t* hard, int32_t* trials,
                                         int batch, int max_trials, void* stream)
{
    if (!c || !llr || !hard || !trials || max_trials < 0 || max_trials > 1000) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    if (!check_batch(c, batch)) return CUSAT_E_BATCH;
    ldpc_args a = make_ldpc_args(c, llr, hard, trials);
    a.max_trials = max_trials;
    if (c->n_param_iters > 0) {
        a.beta = c->d_beta;
        a.alpha = c->d_alpha;
        a.n_param_iters = c->n_param_iters;
    }
    HIP_TRY(launch_ldpc_decode(a, batch, (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_rx_fec_learned(cusat_fec_ctx* c, const int16_t* llr, uint8_t* hard, int32_t* trials,
                                    int32_t* valid, int32_t* fallback, int batch, int learned_trials,
                                    void* stream)
{
    if (!c || !valid) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    int rc = cusat_ldpc_decode_learned(c, llr, hard, trials, batch, learned_trials, stream);
    if (rc == CUSAT_OK) rc = cusat_bch_check(c, hard, valid, batch, stream);
    if (rc != CUSAT_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    int32_t* mask = fallback ? fallback : c->d_mask;
    HIP_TRY(launch_fallback_mask(trials, valid, mask, batch, st));
    // plain decoder (beta = 1, no scale, max_trials of the context) on the
    // frames the learned pass did not deliver, then their BCH check
    ldpc_args a = make_ldpc_args(c, llr, hard, trials);
    a.active = mask;
    HIP_TRY(launch_ldpc_decode(a, batch, st));
    bch_args b = make_bch_args(c, hard, valid);
    b.active = mask;
    HIP_TRY(launch_bch_check(b, batch, st));
    return CUSAT_OK;
}

extern "C" int cusat_modcod_constellation(int framesize, int rate, int modulation, float* points,
                                          uint32_t* col, int* n_mod, int* rows)
{
    constellation con;
    interleaver il;
    if (!make_constellation(modulation, rate, framesize, &con) ||
        !make_interleaver(modulation, rate, framesize, &il)) {
        set_last_error("unsupported (frame %d, rate %d, modulation %d)", framesize, rate, modulation);
        return CUSAT_E_INVALID;
    }
    if (points) memcpy(points, con.xy.data(), sizeof(float) * 2 * con.n_points);
    if (col) for (int b = 0; b < 8; b++) col[b] = il.col[b];
    if (n_mod) *n_mod = con.bits_per_sym;
    if (rows) *rows = il.rows;
    return CUSAT_OK;
}

extern "C" int cusat_bch_encode(int framesize, int rate, const uint8_t* msg, uint8_t* cw, int batch)
{
    fec_params fp;
    bch_code bch;
    if (!msg || !cw || batch < 0 || !get_fec_params(framesize, rate, &fp) ||
        !make_bch(framesize, fp.n_bch, fp.t, &bch) || bch.k_bits != fp.k_bch) {
        set_last_error("cusat_bch_encode: bad arguments");
        return CUSAT_E_INVALID;
    }
    for (int b = 0; b < batch; b++)
        bch_encode(bch, msg + (size_t)b * (fp.k_bch / 8), cw + (size_t)b * (fp.n_bch / 8));
    return CUSAT_OK;
}
