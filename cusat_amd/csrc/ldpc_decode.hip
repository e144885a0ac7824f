// Layered offset/normalised min-sum LDPC decoder for the DVB-S2/S2X IRA
// codes on gfx950.
//
// Semantics follow lib/cusat_dsp/cuda/ldpc_decode_kernel.cu (wineslab/cusat):
// int16 APP with saturation, int8 check-to-variable messages clamped to
// +-127, offset beta (plain = 1) inside the check node (:242-275), optional
// scale n/64 (:75-90), data writes that re-apply the delta to the APP
// (:793-801), parity writes from the read-phase value (:803-833), syndrome
// early exit before every iteration (:689-749), result = T - t or -1.
//
// Layout (one workgroup = one frame):
//  * LDS: app_d[K] (information VNs), app_p[q][360] (parity VN of check
//    (i, j) = channel bit K + q j + i, stored layer-major so the 360 checks
//    of a layer touch consecutive words), then the delta slots
//    [n_slots][360] int16 for colliding circulants.
//  * Thread j (0..359) of lane l handles check j of the layer scheduled on
//    lane l in the current step; six wave64s per lane, lanes in different
//    waves so every branch on the layer id is wave-uniform.
//  * Check j of layer i reads, for each circulant (r, s) of the layer, VN
//    360 r + (j - s) mod 360: a cyclic shift, so a wave reads 64 consecutive
//    int16 words of LDS (at most one wrap) -- no bank conflicts.
//  * Messages live in HBM as int8 [layer][edge][360]: a wave's loads and
//    stores of one edge are 64 consecutive bytes.
//  * A step touches each VN once except on colliding circulants (a
//    weight-2/3 circulant inside a layer, or a VN group shared by the two
//    layers of a step).  Those edges park their delta (y - old) in an LDS
//    slot; after the step's barrier one thread per VN adds the deltas in
//    (lane, circulant) order with int16 saturation per add.  All reads of a
//    step therefore see the step-start APP, and every VN that is touched
//    once gets exactly the reference's update.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace cusat {

namespace {

constexpr int LANE_THREADS = 384;  // 6 x wave64, 360 checks

__device__ __forceinline__ int sat16(int v) { return min(max(v, -32768), 32767); }
// |x| with |-32768| = 32767 (qabs, ldpc_decode_kernel.cu:33-37)
__device__ __forceinline__ int qabs16(int x) { return min(x < 0 ? -x : x, 32767); }
__device__ __forceinline__ int nms(int v, int n)
{
    if (n == 0) return v;
    const int m = ((v < 0 ? -v : v) * n + 32) >> 6;
    return v < 0 ? -m : m;
}

template <int CMAX, int LANES>
__global__ void __launch_bounds__(LANES* LANE_THREADS)
    ldpc_layered_kernel(ldpc_args a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    int16_t* app_d = reinterpret_cast<int16_t*>(lds_raw);
    int16_t* app_p = app_d + a.K;
    int16_t* dslot = app_p + a.R;

    constexpr int NTHR = LANES * LANE_THREADS;
    const int cw = blockIdx.x;
    const int t = threadIdx.x;
    const int lane = LANES == 2 ? (t >= LANE_THREADS ? 1 : 0) : 0;
    const int j = t - lane * LANE_THREADS;
    const bool act = j < CIRC;
    const int q = a.q;
    const int16_t* llr = a.llr + (size_t)cw * a.N;
    int8_t* msg = a.msg + (size_t)cw * a.lt_alloc;

    // Load with the int16-APP input clamp (ldpc_decode_kernel.cu:655-670):
    // -32768 -> -32767.
    for (int v = t; v < a.K; v += NTHR) app_d[v] = (int16_t)max((int)llr[v], -32767);
    for (int p = t; p < a.R; p += NTHR) {
        const int i = p % q, jj = p / q;
        app_p[i * CIRC + jj] = (int16_t)max((int)llr[a.K + p], -32767);
    }
    __syncthreads();

    int res = -1;
    for (int it = 0;; it++) {
        // ---- syndrome over all checks, step by step, early exit ----
        int ok = 1;
        for (int s = 0; s < a.n_steps; s++) {
            const int i = a.steps[2 * s + lane];
            int bad = 0;
            if (act && i >= 0) {
                int syn = app_p[i * CIRC + j];
                if (i > 0) syn ^= app_p[(i - 1) * CIRC + j];
                else if (j > 0) syn ^= app_p[(q - 1) * CIRC + j - 1];
                const int cnt = a.lay_cnt[i];
                const uint32_t* cc = a.circ + a.lay_off[i];
                for (int e = 0; e < cnt; e++) {
                    const uint32_t c = cc[e];
                    int jj = j - (int)(c & 511u);
                    jj += jj < 0 ? CIRC : 0;
                    syn ^= app_d[(int)((c >> 9) & 511u) * CIRC + jj];
                }
                bad = syn < 0;
            }
            if (__syncthreads_or(bad)) { ok = 0; break; }
        }
        if (ok) { res = a.max_trials - it; break; }
        if (it == a.max_trials) break;

        const bool first = it == 0;
        // ---- one layered iteration ----
        for (int s = 0; s < a.n_steps; s++) {
            const int i = a.steps[2 * s + lane];
            if (act && i >= 0) {
                int bt = 1, al = 0;
                if (a.beta && it < a.n_param_iters) bt = a.beta[it * q + i];
                if (a.alpha && it < a.n_param_iters) al = a.alpha[it * q + i];
                const int cnt = a.lay_cnt[i];
                const uint32_t* cc = a.circ + a.lay_off[i];
                int8_t* m = msg + a.lay_moff[i] + j;

                int xv[CMAX], av[CMAX], ov[CMAX];
                int mn1 = 32767, mn2 = 32767, sg = 0;
#pragma unroll
                for (int e = 0; e < CMAX; e++) {
                    if (e < cnt) {
                        const uint32_t c = cc[e];
                        int jj = j - (int)(c & 511u);
                        jj += jj < 0 ? CIRC : 0;
                        const int ap = app_d[(int)((c >> 9) & 511u) * CIRC + jj];
                        const int o = first ? 0 : (int)m[e * CIRC];
                        const int x = sat16(ap - o);
                        av[e] = ap;
                        ov[e] = o;
                        xv[e] = x;
                        const int mm = max(qabs16(x) - bt, 0);
                        mn2 = min(mn2, max(mn1, mm));
                        mn1 = min(mn1, mm);
                        sg ^= x;
                    }
                }
                // parity edges: own P(i, j), previous P(i-1, j) or P(q-1, j-1)
                const int p0 = i * CIRC + j;
                const int o0 = first ? 0 : (int)m[cnt * CIRC];
                const int x0 = sat16((int)app_p[p0] - o0);
                {
                    const int mm = max(qabs16(x0) - bt, 0);
                    mn2 = min(mn2, max(mn1, mm));
                    mn1 = min(mn1, mm);
                    sg ^= x0;
                }
                const int p1 = i > 0 ? (i - 1) * CIRC + j : (j > 0 ? (q - 1) * CIRC + j - 1 : -1);
                int x1 = 0;
                if (p1 >= 0) {
                    const int o1 = first ? 0 : (int)m[(cnt + 1) * CIRC];
                    x1 = sat16((int)app_p[p1] - o1);
                    const int mm = max(qabs16(x1) - bt, 0);
                    mn2 = min(mn2, max(mn1, mm));
                    mn1 = min(mn1, mm);
                    sg ^= x1;
                }
                auto out = [&](int x) {
                    const int mm = max(qabs16(x) - bt, 0);
                    const int oth = mm == mn1 ? mn2 : mn1;
                    const int v = (sg ^ x) < 0 ? -oth : oth;
                    return min(max(nms(v, al), -127), 127);
                };
#pragma unroll
                for (int e = 0; e < CMAX; e++) {
                    if (e < cnt) {
                        const int y = out(xv[e]);
                        const uint32_t c = cc[e];
                        int jj = j - (int)(c & 511u);
                        jj += jj < 0 ? CIRC : 0;
                        const int slot = (int)(c >> 18);
                        if (slot == 0)
                            app_d[(int)((c >> 9) & 511u) * CIRC + jj] = (int16_t)sat16(av[e] + y - ov[e]);
                        else
                            dslot[(slot - 1) * CIRC + jj] = (int16_t)(y - ov[e]);
                        m[e * CIRC] = (int8_t)y;
                    }
                }
                const int y0 = out(x0);
                app_p[p0] = (int16_t)sat16(x0 + y0);
                m[cnt * CIRC] = (int8_t)y0;
                if (p1 >= 0) {
                    const int y1 = out(x1);
                    app_p[p1] = (int16_t)sat16(x1 + y1);
                    m[(cnt + 1) * CIRC] = (int8_t)y1;
                }
            }
            __syncthreads();
            const int g0 = a.grp_off[s], g1 = a.grp_off[s + 1];
            if (g1 > g0) {
                const int total = (g1 - g0) * CIRC;
                for (int idx = t; idx < total; idx += NTHR) {
                    const int gi = idx / CIRC, jj = idx - gi * CIRC;
                    const collision_group cg = a.groups[g0 + gi];
                    const int v = cg.row * CIRC + jj;
                    int val = app_d[v];
                    for (int e = 0; e < cg.n; e++) val = sat16(val + dslot[cg.slot[e] * CIRC + jj]);
                    app_d[v] = (int16_t)val;
                }
                __syncthreads();
            }
        }
    }

    // hard decisions, K bits MSB first (ldpc_decode_kernel.cu:468-493)
    uint8_t* h = a.hard + (size_t)cw * (a.K >> 3);
    for (int b = t; b < (a.K >> 3); b += NTHR) {
        int byte = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) byte |= (app_d[8 * b + k] < 0 ? 1 : 0) << (7 - k);
        h[b] = (uint8_t)byte;
    }
    if (t == 0) a.result[cw] = res;
}

template <int CMAX, int LANES>
hipError_t launch_t(const ldpc_args& a, int batch, hipStream_t stream)
{
    const int lds = ldpc_lds_bytes(a.K, a.R, a.n_slots);
    static int configured = 0;
    if (lds > configured) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&ldpc_layered_kernel<CMAX, LANES>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        configured = lds;
    }
    hipLaunchKernelGGL((ldpc_layered_kernel<CMAX, LANES>), dim3(batch), dim3(LANES * LANE_THREADS), lds,
                       stream, a);
    return hipGetLastError();
}

template <int LANES>
hipError_t launch_lanes(const ldpc_args& a, int batch, hipStream_t stream)
{
    if (a.cmax <= 4) return launch_t<4, LANES>(a, batch, stream);
    if (a.cmax <= 8) return launch_t<8, LANES>(a, batch, stream);
    if (a.cmax <= 12) return launch_t<12, LANES>(a, batch, stream);
    if (a.cmax <= 16) return launch_t<16, LANES>(a, batch, stream);
    if (a.cmax <= 20) return launch_t<20, LANES>(a, batch, stream);
    if (a.cmax <= 28) return launch_t<28, LANES>(a, batch, stream);
    return hipErrorInvalidValue;
}

}  // namespace

int ldpc_lds_bytes(int K, int R, int n_slots)
{
    return ((K + R) * 2 + n_slots * CIRC * 2 + 15) & ~15;
}

hipError_t launch_ldpc_decode(const ldpc_args& a, int batch, hipStream_t stream)
{
    if (a.lanes == 2) return launch_lanes<2>(a, batch, stream);
    return launch_lanes<1>(a, batch, stream);
}

}  // namespace cusat
