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
                       const int* __restrict__ hdr, const float* __restrict__ sym, int16_t* __restrict__ llr)
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
        a.qpsk_scale = r.qpsk_scale;
        a.psk8_scale = r.psk8_scale;
        a.gain = r.gain;
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

}  // namespace

hipError_t launch_demap(const demap_args& a, int batch, hipStream_t stream)
{
    dim3 grid((a.dlen + 255) / 256, batch);
    hipLaunchKernelGGL(demap_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_demap_mixed(const rx_code* codes, const rx_rec* recs, const int* hdr, const float* sym,
                              int16_t* llr, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(demap_mixed_kernel, dim3(grid), dim3(256), 0, stream, codes, recs, hdr, sym, llr);
    return hipGetLastError();
}

}  // namespace cusat
