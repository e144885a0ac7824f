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

__global__ void __launch_bounds__(256) demap_kernel(demap_args a)
{
    __shared__ float pts[512];
    if (a.pts)
        for (int i = threadIdx.x; i < 2 * a.n_points; i += blockDim.x) pts[i] = a.pts[i];
    __syncthreads();

    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int cw = blockIdx.y;
    if (j >= a.dlen) return;
    const float2 y = reinterpret_cast<const float2*>(a.sym)[(size_t)cw * a.xfec + j];
    const float re = y.x, im = y.y;
    int16_t* llr = a.llr + (size_t)cw * a.N;

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

}  // namespace

hipError_t launch_demap(const demap_args& a, int batch, hipStream_t stream)
{
    dim3 grid((a.dlen + 255) / 256, batch);
    hipLaunchKernelGGL(demap_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace cusat
