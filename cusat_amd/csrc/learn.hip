// Training of the learned offset / normalised min-sum decoder: fp32 forward
// and backward passes through unrolled layered iterations, plus Adam.
//
// Model (the integer decoder of ldpc_kernel.h without rounding or the int16
// APP rails): per iteration t and schedule step, every check (i, j) of the
// step's layers reads x_e = L[v_e] - m_e at the step-start APP, computes
//   y_e = clamp(alpha[t][i] * s_e * max(mu_e - beta[t][i], 0), -127, 127),
//   mu_e = min over the check's other edges of |x|, s_e = product of the
//   other edges' signs (x < 0 is negative),
// and then L[v_e] += y_e - m_e, m_e = y_e (colliding updates add up).
// The loss is the mean over iterations, frames and bits of
// softplus(-kappa * s * L_t[v]) with s = +1 for a 0 bit (BCE of the APP
// after each iteration against the transmitted codeword).
//
// Work split: one workgroup per frame and per schedule step (768 threads:
// two lanes of 384, thread j < 360 of a lane owns check j of the lane's
// layer).  One launch per step keeps the step-synchronous rule simple: all
// reads of a step happen before the barrier, all updates (fp32 atomics for
// the APP, since two edges of a step may share a variable node) after it,
// and the next step is the next launch.  The backward pass walks the same
// launches in reverse; the forward pass stores every x_e per iteration
// ([B][T][LT_alloc] floats) and the loss gradient of every iteration's APP
// ([B][T][N]), which is all the backward pass needs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <new>
#include <vector>

#include "ctx.h"
#include "cusat_fec.h"

using namespace cusat;

struct cusat_learn_ctx {
    const cusat_fec_ctx* fec = nullptr;
    int max_batch = 0, T = 0, N = 0, K = 0, q = 0, E = 0, n_steps = 0, cm = 0;
    void* d_tab = nullptr;
    int *circ_rs = nullptr, *lay_cnt = nullptr, *lay_off = nullptr, *lay_moff = nullptr, *steps = nullptr;
    float *L = nullptr, *M = nullptr, *X = nullptr, *G = nullptr, *gL = nullptr, *gM = nullptr;
};

namespace {

constexpr int LANE_T = 384;
constexpr int BLOCK = 2 * LANE_T;

struct largs {
    const int* circ_rs;
    const int* lay_cnt;
    const int* lay_off;
    const int* lay_moff;
    const int* steps;  // [n_steps][2], -1 = idle lane
    int N, K, q, E, T;
    int quant;  // 1: the integer decoder's check node (straight-through gradients)
    float* L;   // [B][N] APP
    float* M;   // [B][E] messages
    float* X;   // [B][T][E] read-phase values
    float* G;   // [B][T][N] loss gradient of each iteration's APP
    float* gL;  // [B][N]
    float* gM;  // [B][E]
    const float* alpha;  // [T][q]
    const float* beta;
    float* ga;
    float* gb;
};

// Edges of check (i, j): info edges e < cnt at slots 0..CM-1, the two parity
// edges at slots CM, CM+1 (check (0, 0) has only the first).  v = channel
// index of the variable node, mi = index of the edge's message in a frame
// (layer offset + edge * 360 + j, the decoder's message order).
template <int CM>
__device__ __forceinline__ void load_edges(const largs& a, int i, int j, int* v, int* mi, bool* ok)
{
    const int cnt = a.lay_cnt[i], off = a.lay_off[i], moff = a.lay_moff[i];
#pragma unroll
    for (int e = 0; e < CM; e++) {
        ok[e] = e < cnt;
        int vv = 0;
        if (ok[e]) {
            const int rs = a.circ_rs[off + e];
            int jj = j - (rs & 0xffff);
            jj += jj < 0 ? 360 : 0;
            vv = 360 * (rs >> 16) + jj;
        }
        v[e] = vv;
        mi[e] = moff + e * 360 + j;
    }
    v[CM] = a.K + a.q * j + i;
    ok[CM] = true;
    mi[CM] = moff + cnt * 360 + j;
    ok[CM + 1] = (i | j) != 0;
    v[CM + 1] = i > 0 ? a.K + a.q * j + i - 1 : (j > 0 ? a.K + a.q * (j - 1) + a.q - 1 : a.K);
    mi[CM + 1] = moff + (cnt + 1) * 360 + j;
}

template <int CM>
__device__ __forceinline__ void min_sum(const float* x, const bool* ok, float& mn1, float& mn2, int& i1,
                                        int& i2, unsigned& par)
{
    mn1 = INFINITY;
    mn2 = INFINITY;
    i1 = -1;
    i2 = -1;
    par = 0;
#pragma unroll
    for (int e = 0; e < CM + 2; e++) {
        if (!ok[e]) continue;
        const float m = fabsf(x[e]);
        par ^= x[e] < 0.0f ? 1u : 0u;
        if (m < mn1) {
            mn2 = mn1;
            i2 = i1;
            mn1 = m;
            i1 = e;
        } else if (m < mn2) {
            mn2 = m;
            i2 = e;
        }
    }
}

// Check-node output y of one edge (mu = min of |x| over the other edges,
// sg = its sign) and what its derivatives need: r (the offset magnitude),
// act (the offset is active: mu above it) and pass (the +-127 clamp passes).
//  * fp32 model: y = clamp(al sg max(mu - be, 0), -127, 127).
//  * quantised model = the integer decoder: beta_q = max(rint(be), 0),
//    n = clamp(rint(64 al), 1, 1023), |y| = min(floor((r n + 32) / 64), 127),
//    r = max(mu - beta_q, 0) (ldpc_decode_kernel.cu:75-90,242-275); with
//    integer inputs every value is an exact fp32 integer (r n < 2^24).
//    Gradients go straight through the rounding: dy/dal = sg r,
//    dy/dbe = -al sg, dy/dmu = al sg, gated by act and pass.
__device__ __forceinline__ void cn_edge(float mu, float sg, float al, float be, int quant, float& y,
                                        float& r, bool& act, bool& pass)
{
    if (quant) {
        const float bq = fmaxf(rintf(be), 0.0f);
        const float n = fminf(fmaxf(rintf(64.0f * al), 1.0f), 1023.0f);
        r = fmaxf(mu - bq, 0.0f);
        const float m = floorf((r * n + 32.0f) * (1.0f / 64.0f));
        pass = m <= 127.0f;
        y = sg * fminf(m, 127.0f);
        act = mu > bq;
    } else {
        r = fmaxf(mu - be, 0.0f);
        const float v = al * sg * r;
        pass = fabsf(v) <= 127.0f;
        y = fminf(fmaxf(v, -127.0f), 127.0f);
        act = mu > be;
    }
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

template <int CM>
__global__ void __launch_bounds__(BLOCK) fwd_step(largs a, int t, int s)
{
    const int b = blockIdx.x;
    const int lane = threadIdx.x >= LANE_T ? 1 : 0;
    const int j = threadIdx.x - lane * LANE_T;
    const int i = a.steps[2 * s + lane];
    const bool act = i >= 0 && j < 360;
    float* L = a.L + (size_t)b * a.N;
    float* M = a.M + (size_t)b * a.E;
    float* X = a.X + ((size_t)b * a.T + t) * a.E;
    int v[CM + 2], mi[CM + 2];
    bool ok[CM + 2];
    float x[CM + 2], mo[CM + 2], y[CM + 2];
    if (act) {
        load_edges<CM>(a, i, j, v, mi, ok);
#pragma unroll
        for (int e = 0; e < CM + 2; e++) {
            mo[e] = 0.0f;
            x[e] = 0.0f;
            if (ok[e]) {
                mo[e] = M[mi[e]];
                x[e] = L[v[e]] - mo[e];
                X[mi[e]] = x[e];
            }
        }
        const float al = a.alpha[t * a.q + i], be = a.beta[t * a.q + i];
        float mn1, mn2;
        int i1, i2;
        unsigned par;
        min_sum<CM>(x, ok, mn1, mn2, i1, i2, par);
#pragma unroll
        for (int e = 0; e < CM + 2; e++) {
            const float sg = (par ^ (x[e] < 0.0f ? 1u : 0u)) ? -1.0f : 1.0f;
            const float mu = e == i1 ? mn2 : mn1;
            float r;
            bool act_e, pass;
            cn_edge(mu, sg, al, be, a.quant, y[e], r, act_e, pass);
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int e = 0; e < CM + 2; e++) {
            if (!ok[e]) continue;
            atomicAdd(&L[v[e]], y[e] - mo[e]);
            M[mi[e]] = y[e];
        }
    }
}

template <int CM>
__global__ void __launch_bounds__(BLOCK) bwd_step(largs a, int t, int s)
{
    __shared__ float red[BLOCK / 64][2];
    const int b = blockIdx.x;
    const int lane = threadIdx.x >= LANE_T ? 1 : 0;
    const int j = threadIdx.x - lane * LANE_T;
    const int i = a.steps[2 * s + lane];
    const bool act = i >= 0 && j < 360;
    float* gL = a.gL + (size_t)b * a.N;
    float* gM = a.gM + (size_t)b * a.E;
    const float* X = a.X + ((size_t)b * a.T + t) * a.E;
    int v[CM + 2], mi[CM + 2];
    bool ok[CM + 2];
    float gx[CM + 2], gmo[CM + 2];
    float gal = 0.0f, gbe = 0.0f;
    if (act) {
        load_edges<CM>(a, i, j, v, mi, ok);
        float x[CM + 2], gl[CM + 2], dmu[CM + 2];
#pragma unroll
        for (int e = 0; e < CM + 2; e++) x[e] = ok[e] ? X[mi[e]] : 0.0f;
        const float al = a.alpha[t * a.q + i], be = a.beta[t * a.q + i];
        float mn1, mn2;
        int i1, i2;
        unsigned par;
        min_sum<CM>(x, ok, mn1, mn2, i1, i2, par);
#pragma unroll
        for (int e = 0; e < CM + 2; e++) {
            gl[e] = 0.0f;
            dmu[e] = 0.0f;
            if (!ok[e]) continue;
            gl[e] = gL[v[e]];
            // y_e feeds the APP update and becomes the edge's next message
            const float gy = gl[e] + gM[mi[e]];
            const float sg = (par ^ (x[e] < 0.0f ? 1u : 0u)) ? -1.0f : 1.0f;
            const float mu = e == i1 ? mn2 : mn1;
            float y, r;
            bool act_e, pass;
            cn_edge(mu, sg, al, be, a.quant, y, r, act_e, pass);
            if (pass) {  // the clamp passes the gradient on [-127, 127]
                gal += gy * sg * r;
                if (act_e) {
                    gbe -= gy * al * sg;
                    dmu[e] = gy * al * sg;
                }
            }
        }
        // mu_e = |x_i1| for e != i1 and |x_i2| for e == i1
        float d_rest = 0.0f, d_i1 = 0.0f;
#pragma unroll
        for (int e = 0; e < CM + 2; e++) {
            d_rest += e == i1 ? 0.0f : dmu[e];
            d_i1 += e == i1 ? dmu[e] : 0.0f;
        }
#pragma unroll
        for (int e = 0; e < CM + 2; e++) {
            // d|x|/dx with the subgradient 0 at x = 0
            const float sx = x[e] < 0.0f ? -1.0f : (x[e] > 0.0f ? 1.0f : 0.0f);
            gx[e] = (e == i1 ? sx * d_rest : 0.0f) + (e == i2 ? sx * d_i1 : 0.0f);
            // x_e = L_before - m_old and L_after = L_before + y - m_old
            gmo[e] = -gx[e] - gl[e];
        }
    }
    gal = wave_sum(gal);
    gbe = wave_sum(gbe);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][0] = gal;
        red[threadIdx.x >> 6][1] = gbe;
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int e = 0; e < CM + 2; e++) {
            if (!ok[e]) continue;
            atomicAdd(&gL[v[e]], gx[e]);
            gM[mi[e]] = gmo[e];
        }
    }
    if (j == 0 && i >= 0) {
        float sa = 0.0f, sb = 0.0f;
        const int w0 = lane * (LANE_T / 64);
        for (int w = w0; w < w0 + LANE_T / 64; w++) {
            sa += red[w][0];
            sb += red[w][1];
        }
        atomicAdd(&a.ga[t * a.q + i], sa);
        atomicAdd(&a.gb[t * a.q + i], sb);
    }
}

__global__ void __launch_bounds__(256) loss_kernel(const float* L, const uint8_t* cw, float* G, float* loss, int N,
                                                   int T, int t, float kappa, float gscale, float lscale)
{
    __shared__ float red[4];
    const int b = blockIdx.y;
    const int v = blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (v < N) {
        const int bit = (cw[(size_t)b * (N / 8) + (v >> 3)] >> (7 - (v & 7))) & 1;
        const float s = bit ? -1.0f : 1.0f;
        const float z = kappa * s * L[(size_t)b * N + v];
        const float ez = expf(-fabsf(z));
        l = fmaxf(-z, 0.0f) + log1pf(ez);               // softplus(-z)
        const float sig = z >= 0.0f ? ez / (1.0f + ez) : 1.0f / (1.0f + ez);  // sigmoid(-z)
        G[((size_t)b * T + t) * N + v] = -kappa * s * sig * gscale;
    }
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&loss[t], (red[0] + red[1] + red[2] + red[3]) * lscale);
}

__global__ void __launch_bounds__(256) fill_kernel(float* p, size_t n, float val)
{
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) p[k] = val;
}

__global__ void __launch_bounds__(256) copy_kernel(float* dst, const float* src, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) dst[k] = src[k];
}

// gL[b][v] += G[b][t][v]
__global__ void __launch_bounds__(256) add_grad_kernel(float* gL, const float* G, int N, int T, int t)
{
    const int b = blockIdx.y;
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < N) gL[(size_t)b * N + v] += G[((size_t)b * T + t) * N + v];
}

__global__ void __launch_bounds__(256) adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2, float lo,
                                                   float hi)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float upd = lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    p[k] = fminf(fmaxf(p[k] - upd, lo), hi);
}

dim3 fill_grid(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)); }

#define HIP_TRY(expr)                                                       \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            set_last_error("%s: %s", #expr, hipGetErrorString(e_));         \
            return CUSAT_E_HIP;                                             \
        }                                                                   \
    } while (0)

template <int CM>
hipError_t run_steps(const largs& a, int B, int n_steps, bool backward, int t, hipStream_t st)
{
    if (!backward) {
        for (int s = 0; s < n_steps; s++) hipLaunchKernelGGL(fwd_step<CM>, dim3(B), dim3(BLOCK), 0, st, a, t, s);
    } else {
        for (int s = n_steps - 1; s >= 0; s--) hipLaunchKernelGGL(bwd_step<CM>, dim3(B), dim3(BLOCK), 0, st, a, t, s);
    }
    return hipGetLastError();
}

hipError_t steps_for(int cm, const largs& a, int B, int n_steps, bool backward, int t, hipStream_t st)
{
    switch (cm) {
    case 4: return run_steps<4>(a, B, n_steps, backward, t, st);
    case 8: return run_steps<8>(a, B, n_steps, backward, t, st);
    case 16: return run_steps<16>(a, B, n_steps, backward, t, st);
    default: return run_steps<32>(a, B, n_steps, backward, t, st);
    }
}

}  // namespace

extern "C" int cusat_learn_ctx_create(cusat_learn_ctx** out, const cusat_fec_ctx* fec, int max_batch, int n_iters)
{
    if (!out || !fec || max_batch < 1 || n_iters < 1 || n_iters > 100) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    *out = nullptr;
    const decoder_geometry& g = fec->geo;
    int cmx = 0;
    for (int c : g.lay_cnt) cmx = std::max(cmx, c);
    const int cm = cmx <= 4 ? 4 : cmx <= 8 ? 8 : cmx <= 16 ? 16 : cmx <= 32 ? 32 : 0;
    if (!cm) {
        set_last_error("code has %d info circulants per layer (training supports 32)", cmx);
        return CUSAT_E_INVALID;
    }
    cusat_learn_ctx* lc = new (std::nothrow) cusat_learn_ctx();
    if (!lc) return CUSAT_E_NOMEM;
    lc->fec = fec;
    lc->max_batch = max_batch;
    lc->T = n_iters;
    lc->N = g.N;
    lc->K = g.K;
    lc->q = g.q;
    lc->E = g.lt_alloc;
    lc->n_steps = g.n_steps;
    lc->cm = cm;
    std::vector<int> tab;
    auto put = [&](const std::vector<int>& v) {
        size_t at = tab.size();
        tab.insert(tab.end(), v.begin(), v.end());
        while (tab.size() & 63) tab.push_back(0);
        return at;
    };
    const size_t o_rs = put(g.circ_rs), o_cnt = put(g.lay_cnt), o_off = put(g.lay_off),
                 o_moff = put(g.lay_moff), o_st = put(std::vector<int>(g.steps.begin(), g.steps.end()));
    const size_t B = max_batch;
    const size_t nL = B * g.N, nM = B * lc->E, nX = B * n_iters * (size_t)lc->E, nG = B * n_iters * (size_t)g.N;
    hipError_t e = hipMalloc(&lc->d_tab, tab.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(lc->d_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&lc->L, (2 * nL + 2 * nM) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&lc->X, nX * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&lc->G, nG * sizeof(float));
    if (e != hipSuccess) {
        set_last_error("training buffers (%zu frames x %d iterations): %s", B, n_iters, hipGetErrorString(e));
        cusat_learn_ctx_destroy(lc);
        return CUSAT_E_NOMEM;
    }
    int* t = (int*)lc->d_tab;
    lc->circ_rs = t + o_rs;
    lc->lay_cnt = t + o_cnt;
    lc->lay_off = t + o_off;
    lc->lay_moff = t + o_moff;
    lc->steps = t + o_st;
    lc->gL = lc->L + nL;
    lc->M = lc->gL + nL;
    lc->gM = lc->M + nM;
    *out = lc;
    return CUSAT_OK;
}

extern "C" int cusat_learn_ctx_destroy(cusat_learn_ctx* lc)
{
    if (!lc) return CUSAT_OK;
    if (lc->d_tab) hipFree(lc->d_tab);
    if (lc->L) hipFree(lc->L);
    if (lc->X) hipFree(lc->X);
    if (lc->G) hipFree(lc->G);
    delete lc;
    return CUSAT_OK;
}

extern "C" int cusat_learn_fwd_bwd(cusat_learn_ctx* lc, const float* llr, const uint8_t* cw, int batch,
                                   const float* alpha, const float* beta, float kappa, unsigned flags,
                                   float* grad_alpha,
                                   float* grad_beta, float* loss, float* app_out, float* grad_llr, void* stream)
{
    if (!lc || !llr || !cw || !alpha || !beta || !grad_alpha || !grad_beta || !loss) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    if (batch < 1 || batch > lc->max_batch) {
        set_last_error("batch %d outside 1..%d", batch, lc->max_batch);
        return CUSAT_E_BATCH;
    }
    hipStream_t st = (hipStream_t)stream;
    const size_t nL = (size_t)batch * lc->N, nM = (size_t)batch * lc->E;
    largs a{};
    a.circ_rs = lc->circ_rs;
    a.lay_cnt = lc->lay_cnt;
    a.lay_off = lc->lay_off;
    a.lay_moff = lc->lay_moff;
    a.steps = lc->steps;
    a.N = lc->N;
    a.K = lc->K;
    a.q = lc->q;
    a.E = lc->E;
    a.T = lc->T;
    a.quant = (flags & CUSAT_LEARN_QUANTIZED) ? 1 : 0;
    a.L = lc->L;
    a.M = lc->M;
    a.X = lc->X;
    a.G = lc->G;
    a.gL = lc->gL;
    a.gM = lc->gM;
    a.alpha = alpha;
    a.beta = beta;
    a.ga = grad_alpha;
    a.gb = grad_beta;

    // forward
    hipLaunchKernelGGL(copy_kernel, fill_grid(nL), dim3(256), 0, st, lc->L, llr, nL);
    hipLaunchKernelGGL(fill_kernel, fill_grid(nM), dim3(256), 0, st, lc->M, nM, 0.0f);
    const dim3 vgrid((lc->N + 255) / 256, batch);
    const float gscale = 1.0f / ((float)lc->T * (float)nL), lscale = 1.0f / (float)nL;
    for (int t = 0; t < lc->T; t++) {
        HIP_TRY(steps_for(lc->cm, a, batch, lc->n_steps, false, t, st));
        hipLaunchKernelGGL(loss_kernel, vgrid, dim3(256), 0, st, lc->L, cw, lc->G, loss, lc->N, lc->T, t, kappa,
                           gscale, lscale);
    }
    if (app_out) hipLaunchKernelGGL(copy_kernel, fill_grid(nL), dim3(256), 0, st, app_out, lc->L, nL);

    // backward
    hipLaunchKernelGGL(fill_kernel, fill_grid(nL), dim3(256), 0, st, lc->gL, nL, 0.0f);
    hipLaunchKernelGGL(fill_kernel, fill_grid(nM), dim3(256), 0, st, lc->gM, nM, 0.0f);
    for (int t = lc->T - 1; t >= 0; t--) {
        hipLaunchKernelGGL(add_grad_kernel, vgrid, dim3(256), 0, st, lc->gL, lc->G, lc->N, lc->T, t);
        HIP_TRY(steps_for(lc->cm, a, batch, lc->n_steps, true, t, st));
    }
    if (grad_llr) hipLaunchKernelGGL(copy_kernel, fill_grid(nL), dim3(256), 0, st, grad_llr, lc->gL, nL);
    HIP_TRY(hipGetLastError());
    return CUSAT_OK;
}

extern "C" int cusat_learn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                                float eps, int step, float lo, float hi, void* stream)
{
    if (!p || !g || !m || !v || n < 1 || step < 1) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1,
                       b2, eps, c1, c2, lo, hi);
    HIP_TRY(hipGetLastError());
    return CUSAT_OK;
}
