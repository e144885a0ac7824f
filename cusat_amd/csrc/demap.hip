// Fused soft demap + bit deinterleave to int16 LLRs on gfx950
// (lib/cusat_dsp/cuda/demapper_kernel.cu:31-184 of wineslab/cusat).
//
// One thread per symbol, grid (ceil(L/256), B).  LLR b of symbol j lands at
// col[b] + j of the frame (the column-row deinterleave of EN 302 307
// clause 5.3.3); QPSK is not interleaved and writes the pair (2j, 2j+1).
// Every float operation is written out (fmaf where a fused multiply-add is
// meant, contraction off otherwise) so the result is a fixed function of the
// inputs; the APSK log-MAP correction uses this file's own exp/log:
// exp by Cody-Waite reduction + Taylor polynomial, log by frexp reduction +
// a least-squares polynomial for log1p.
#include <hip/hip_runtime.h>

#include "kernels.h"

#pragma clang fp contract(off)

namespace cusat {

namespace {

__device__ __forceinline__ float exp_neg(float x)
{
    const float n = rintf(x * 1.44269502e+00f);
    float r = fmaf(-n, 6.93359375e-01f, x);
    r = fmaf(-n, -2.12194442e-04f, r);
    float p = 1.98412701e-04f;
    p = fmaf(p, r, 1.38888892e-03f);
    p = fmaf(p, r, 8.33333377e-03f);
    p = fmaf(p, r, 4.16666679e-02f);
    p = fmaf(p, r, 1.66666672e-01f);
    p = fmaf(p, r, 5.00000000e-01f);
    p = fmaf(p, r, 1.00000000e+00f);
    p = fmaf(p, r, 1.00000000e+00f);
    return ldexpf(p, (int)n);
}

__device__ __forceinline__ float log_pos(float x)
{
    int e;
    float m = frexpf(x, &e);
    if (m < 7.07106769e-01f) {
        m = m + m;
        e -= 1;
    }
    const float f = m - 1.0f;
    float p = -7.638592273e-02f;
    p = fmaf(p, f, 1.291417032e-01f);
    p = fmaf(p, f, -1.324084252e-01f);
    p = fmaf(p, f, 1.418007910e-01f);
    p = fmaf(p, f, -1.660912633e-01f);
    p = fmaf(p, f, 2.000207156e-01f);
    p = fmaf(p, f, -2.500158548e-01f);
    p = fmaf(p, f, 3.333332539e-01f);
    p = fmaf(p, f, -4.999998808e-01f);
    p = fmaf(p, f, 1.0f);
    return fmaf((float)e, 6.93147182e-01f, f * p);
}

__device__ __forceinline__ int16_t to_llr(float v)
{
    v = rintf(v);
    v = fminf(fmaxf(v, -32768.0f), 32767.0f);
    return (int16_t)v;
}

// LLRs of symbol j of one frame (sym = the frame's symbols, llr = its N
// LLRs); pts = the constellation in the LDS (APSK).
__device__ __forceinline__ void demap_symbol(const demap_args& a, const float* pts, int j, const float2* sym,
                                             int16_t* llr)
{
    const float2 y = sym[j];
    const float re = y.x, im = y.y;

    if (a.n_mod == 2) {
        if (2 * j + 1 < a.N) {
            llr[2 * j] = to_llr(re * a.qpsk_scale);
            llr[2 * j + 1] = to_llr(im * a.qpsk_scale);
        }
    } else if (a.n_mod == 3 && !a.pts) {
        const float C = 9.23879504e-01f, S = 3.82683426e-01f;
        const float cr = fmaf(re, C, im * S);
        const float ci = fmaf(im, C, -(re * S));
        const float l0 = (7.07106769e-01f * (fabsf(cr) - fabsf(ci))) * a.psk8_scale;
        llr[a.col[0] + j] = to_llr(l0);
        llr[a.col[1] + j] = to_llr(cr * a.psk8_scale);
        llr[a.col[2] + j] = to_llr(ci * a.psk8_scale);
    } else {
        const int nm = a.n_mod;
        float mn0[8], mn1[8], s0[8], s1[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            mn0[b] = 1e30f;
            mn1[b] = 1e30f;
            s0[b] = 0.0f;
            s1[b] = 0.0f;
        }
        for (int i = 0; i < a.n_points; i++) {
            const float dr = re - pts[2 * i], di = im - pts[2 * i + 1];
            const float d = fmaf(dr, dr, di * di);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                if (b < nm) {
                    if (i & (1 << (nm - 1 - b))) mn1[b] = d < mn1[b] ? d : mn1[b];
                    else                         mn0[b] = d < mn0[b] ? d : mn0[b];
                }
            }
        }
        const float quarter = a.gain * 0.25f, half = a.gain * 0.5f;
        for (int i = 0; i < a.n_points; i++) {
            const float dr = re - pts[2 * i], di = im - pts[2 * i + 1];
            const float d = fmaf(dr, dr, di * di);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                if (b < nm) {
                    if (i & (1 << (nm - 1 - b))) {
                        const float dl = (d - mn1[b]) * quarter;
                        if (dl < 16.0f) s1[b] += exp_neg(-dl);
                    } else {
                        const float dl = (d - mn0[b]) * quarter;
                        if (dl < 16.0f) s0[b] += exp_neg(-dl);
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < 8; b++) {
            if (b < nm) {
                const float corr = log_pos(fmaxf(s0[b], 1e-30f)) - log_pos(fmaxf(s1[b], 1e-30f));
                const float v = fmaf(2.0f, corr, (mn1[b] - mn0[b]) * half);
                const uint32_t pos = a.col[b] + (uint32_t)j;
                if (pos < (uint32_t)a.N) llr[pos] = to_llr(v);
            }
        }
    }
}

__global__ void __launch_bounds__(256) demap_kernel(demap_args a)
{
    __shared__ float pts[512];
    if (a.pts)
        for (int i = threadIdx.x; i < 2 * a.n_points; i += blockDim.x) pts[i] = a.pts[i];
    __syncthreads();

    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int cw = blockIdx.y;
    if (j >= a.dlen) return;
    demap_symbol(a, pts, j, reinterpret_cast<const float2*>(a.sym) + (size_t)cw * a.xfec, a.llr + (size_t)cw * a.N);
}

// Mixed-MODCOD batch: a fixed grid walks the batch's 256-symbol blocks
// (frame f owns blocks recs[f].blk0 .. recs[f + 1].blk0 - 1; hdr[0] =
// frames, hdr[1] = blocks, both read from device memory so a captured launch
// serves any batch).  Every symbol gets exactly the arithmetic of demap_kernel.
__global__ void __launch_bounds__(256)
    demap_mixed_kernel(const rx_code* __restrict__ codes, const rx_rec* __restrict__ recs,
                       const rx_scale* __restrict__ snap, const int* __restrict__ hdr,
                       const float* __restrict__ sym, int16_t* __restrict__ llr)
{
    __shared__ float pts[512];
    const int n = hdr[0], total = hdr[1];
    int cur = -1;  // code whose constellation the LDS holds
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = n - 1;  // last frame with blk0 <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (recs[mid].blk0 <= item) lo = mid;
            else hi = mid - 1;
        }
        const rx_rec& r = recs[lo];
        const rx_code& c = codes[r.code];
        demap_args a;
        a.pts = c.pts;
        a.xfec = c.xfec;
        a.dlen = c.dlen;
        a.N = c.N;
        a.n_mod = c.n_mod;
        a.n_points = c.n_points;
#pragma unroll
        for (int b = 0; b < 8; b++) a.col[b] = c.col[b];
        if (r.flags & RX_SCALE_STATE) {
            const rx_scale& sc = snap[r.code];
            a.qpsk_scale = sc.qpsk_scale;
            a.psk8_scale = sc.psk8_scale;
            a.gain = sc.gain;
        } else {
            a.qpsk_scale = r.qpsk_scale;
            a.psk8_scale = r.psk8_scale;
            a.gain = r.gain;
        }
        if (c.pts && r.code != cur) {
            __syncthreads();  // every lane is done with the previous table
            for (int i = threadIdx.x; i < 2 * c.n_points; i += blockDim.x) pts[i] = c.pts[i];
            __syncthreads();
            cur = r.code;
        }
        const int j = (item - r.blk0) * 256 + threadIdx.x;
        if (j < a.dlen)
            demap_symbol(a, pts, j, reinterpret_cast<const float2*>(sym) + r.sym_off, llr + r.llr_off);
    }
}

// Decision-directed SNR of one frame, the host estimate's arithmetic
// exactly (rx_fec_chain.cc qpsk_snr_dd / constellation_snr_dd, after
// modulation/dd_snr.cc:23-51): QPSK over every symbol with sign decisions,
// other constellations over every 8th symbol with the nearest point; the
// sums run in symbol order in float.  The 256 lanes compute the per-symbol
// terms of a 256-symbol chunk into the LDS; lanes 0 and 1 then add the
// chunk to the signal and error sums, one term after the other, as the
// host loop does (lane 0 adds +0.0f after each signal term, an identity
// for sums that are never -0, so both lanes run one instruction stream).
__device__ __forceinline__ float dd_ratio(float signal, float error)
{
    if (!isfinite(signal) || !isfinite(error)) return 0.0f;
    const float cap = 1e6f;  // error-free input reads as 60 dB
    return error * cap <= signal ? cap : signal / error;
}

__global__ void __launch_bounds__(256)
    dd_estimate_kernel(const rx_code* __restrict__ codes, const rx_rec* __restrict__ recs,
                       const rx_event* __restrict__ ev, const int* __restrict__ hdr,
                       const float* __restrict__ sym, float* __restrict__ est)
{
    __shared__ float pts[512];
    __shared__ __align__(16) float term_a[2][256];  // [0]: signal, [1]: error
    __shared__ __align__(16) float term_b[2][256];
    const int n_ev = hdr[2];
    const int t = threadIdx.x;
    for (int e = blockIdx.x; e < n_ev; e += gridDim.x) {
        const int f = ev[e].frame;
        if (f < 0) continue;
        const rx_rec& r = recs[f];
        const rx_code& c = codes[r.code];
        const bool qpsk = c.n_mod == 2;
        const int np = c.n_points;
        __syncthreads();  // the previous frame's LDS reads are done
        if (!qpsk)
            for (int i = t; i < 2 * np; i += 256) pts[i] = c.dd_pts[i];
        __syncthreads();
        const float2* y = reinterpret_cast<const float2*>(sym) + r.sym_off;
        const int stride = qpsk ? 1 : 8;
        const int m = (c.xfec + stride - 1) / stride;
        float acc = 0.0f;  // lane 0: signal, lane 1: error
        for (int c0 = 0; c0 < m; c0 += 256) {
            const int j = c0 + t;
            if (j < m) {
                const float2 v = y[(size_t)j * stride];
                float sig, e1, e2;
                if (qpsk) {
                    const float a = 0.70710678118654752f;
                    const float dr = v.x < 0 ? -a : a, di = v.y < 0 ? -a : a;
                    sig = dr * dr + di * di;
                    e1 = (v.x - dr) * (v.x - dr);
                    e2 = (v.y - di) * (v.y - di);
                } else {
                    int best = 0;
                    float xr = v.x - pts[0], xi = v.y - pts[1];
                    float bd = xr * xr + xi * xi;
                    for (int k = 1; k < np; k++) {
                        xr = v.x - pts[2 * k];
                        xi = v.y - pts[2 * k + 1];
                        const float d = xr * xr + xi * xi;
                        if (d < bd) {
                            bd = d;
                            best = k;
                        }
                    }
                    const float pr = pts[2 * best], pi = pts[2 * best + 1];
                    sig = pr * pr + pi * pi;
                    e1 = bd;
                    e2 = 0.0f;
                }
                term_a[0][t] = sig;
                term_b[0][t] = 0.0f;
                term_a[1][t] = e1;
                term_b[1][t] = e2;
            }
            __syncthreads();
            if (t < 2) {
                const int n = min(256, m - c0);
                const float* A = term_a[t];
                const float* B = term_b[t];
                int k = 0;
                // eight terms per 16-byte LDS reads, the next eight loaded
                // before these are added: the adds stay in symbol order
                if (n >= 8) {
                    float4 a0 = *reinterpret_cast<const float4*>(A), a1 = *reinterpret_cast<const float4*>(A + 4);
                    float4 b0 = *reinterpret_cast<const float4*>(B), b1 = *reinterpret_cast<const float4*>(B + 4);
                    for (; k + 8 <= n; k += 8) {
                        float4 na0 = a0, na1 = a1, nb0 = b0, nb1 = b1;
                        if (k + 16 <= n) {
                            na0 = *reinterpret_cast<const float4*>(A + k + 8);
                            na1 = *reinterpret_cast<const float4*>(A + k + 12);
                            nb0 = *reinterpret_cast<const float4*>(B + k + 8);
                            nb1 = *reinterpret_cast<const float4*>(B + k + 12);
                        }
                        acc = (acc + a0.x) + b0.x;
                        acc = (acc + a0.y) + b0.y;
                        acc = (acc + a0.z) + b0.z;
                        acc = (acc + a0.w) + b0.w;
                        acc = (acc + a1.x) + b1.x;
                        acc = (acc + a1.y) + b1.y;
                        acc = (acc + a1.z) + b1.z;
                        acc = (acc + a1.w) + b1.w;
                        a0 = na0;
                        a1 = na1;
                        b0 = nb0;
                        b1 = nb1;
                    }
                }
                for (; k < n; k++) acc = (acc + A[k]) + B[k];
            }
            __syncthreads();
        }
        const float error = __shfl(acc, 1);
        if (t == 0) est[e] = dd_ratio(acc, error);
    }
}

// One thread applies the batch's events in order (a host value, or an
// estimate above 1: noise_var = 1 / snr, llr_gain = min(4 / noise_var,
// 1000), as rx_fec_chain::update_snr / update_dd_snr), then every code's
// demapper scales are computed from its state as cusat_fec_demap does.
__global__ void __launch_bounds__(64)
    dd_resolve_kernel(const rx_code* __restrict__ codes, const rx_event* __restrict__ ev,
                      const float* __restrict__ est, const int* __restrict__ hdr, int n_codes, float4* state,
                      rx_scale* __restrict__ snap, float4* __restrict__ state_out)
{
    if (threadIdx.x == 0) {
        const int n_ev = hdr[2];
        for (int e = 0; e < n_ev; e++) {
            const rx_event x = ev[e];
            float s, db;
            if (x.frame < 0) {
                s = x.snr_lin;
                db = x.snr_db;
            } else {
                s = est[e];
                if (!(s > 1.0f)) continue;
                db = __builtin_nanf("");  // the host takes 10 log10(snr_lin)
            }
            const float nv = 1.0f / s;
            state[x.code] = make_float4(nv, fminf(4.0f / nv, 1000.0f), s, db);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n_codes; k += blockDim.x) {
        const float4 st = state[k];
        const rx_code& c = codes[k];
        rx_scale o;
        o.qpsk_scale = 2.0f * 1.4142135623730951f / st.x;
        float gain = st.y;
        if (c.has_cap) gain = st.y < c.gain_cap ? st.y : c.gain_cap;
        o.gain = gain;
        o.psk8_scale = 2.0f * 0.38268343236508977f * gain;
        o.snr_lin = st.z;
        snap[k] = o;
        if (state_out) state_out[k] = st;
    }
}

}  // namespace

hipError_t launch_demap(const demap_args& a, int batch, hipStream_t stream)
{
    dim3 grid((a.dlen + 255) / 256, batch);
    hipLaunchKernelGGL(demap_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_demap_mixed(const rx_code* codes, const rx_rec* recs, const rx_scale* snap, const int* hdr,
                              const float* sym, int16_t* llr, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(demap_mixed_kernel, dim3(grid), dim3(256), 0, stream, codes, recs, snap, hdr, sym, llr);
    return hipGetLastError();
}

hipError_t launch_dd_estimate(const rx_code* codes, const rx_rec* recs, const rx_event* ev, const int* hdr,
                              const float* sym, float* est, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(dd_estimate_kernel, dim3(grid), dim3(256), 0, stream, codes, recs, ev, hdr, sym, est);
    return hipGetLastError();
}

hipError_t launch_dd_resolve(const rx_code* codes, const rx_event* ev, const float* est, const int* hdr,
                             int n_codes, float4* state, rx_scale* snap, float4* state_out, hipStream_t stream)
{
    hipLaunchKernelGGL(dd_resolve_kernel, dim3(1), dim3(64), 0, stream, codes, ev, est, hdr, n_codes, state, snap,
                       state_out);
    return hipGetLastError();
}

}  // namespace cusat
