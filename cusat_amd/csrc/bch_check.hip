// BCH syndrome check on gfx950: valid = (c(x) mod g(x) == 0) for the LDPC
// hard output (lib/cusat_dsp/cuda/bch_common.cuh:91-158 of wineslab/cusat).
//
// One wave64 per frame.  Lane l folds its byte chunk into
// r_l = chunk_l(x) x^P mod g with the byte table lut[b] = b(x) x^P mod g
// (kept in LDS), then moves it to its place in the frame with
// r_l x^(8 (n - end_l)) mod g = XOR of pw[bit][l] over the set bits of r_l
// (pw[b][l] = x^(8 (n - end_l) + b) mod g, precomputed per geometry), and the
// wave XOR-reduces the 64 partial remainders with cross-lane shuffles.
// Because g(0) = 1, c(x) x^P mod g == 0 exactly when c(x) mod g == 0.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace cusat {

namespace {

template <int W>
__global__ void __launch_bounds__(256) bch_check_kernel(bch_args a, int batch)
{
    __shared__ uint64_t tab[256 * W];
    for (int i = threadIdx.x; i < 256 * W; i += blockDim.x) tab[i] = a.lut[i];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int frame = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (frame >= batch || (a.active && a.active[frame] == 0)) return;
    const int P = a.degree;
    const int top_w = (P - 8) >> 6, top_b = (P - 8) & 63;
    const uint64_t top_mask = (P & 63) ? ((1ull << (P & 63)) - 1) : ~0ull;

    uint64_t r[W];
#pragma unroll
    for (int w = 0; w < W; w++) r[w] = 0;
    const uint8_t* src = a.hard + (size_t)frame * a.n_bytes;
    const int begin = lane * a.chunk;
    const int end = min(a.n_bytes, begin + a.chunk);
    for (int i = begin; i < end; i++) {
        const unsigned idx = (unsigned)src[i] ^ (unsigned)((r[top_w] >> top_b) & 0xffu);
#pragma unroll
        for (int w = W - 1; w >= 0; w--) r[w] = (r[w] << 8) | (w ? r[w - 1] >> 56 : 0);
        r[W - 1] &= top_mask;
#pragma unroll
        for (int w = 0; w < W; w++) r[w] ^= tab[idx * W + w];
    }

    uint64_t acc[W];
#pragma unroll
    for (int w = 0; w < W; w++) acc[w] = 0;
    // pw rows are [bit][word][lane]: one row's load is 512 contiguous bytes
    // for the wave.  Branch-free (a lane without bytes has r = 0, and its
    // table entries are 0); the waves of the launch hide the load latency.
    const uint64_t* pw = a.pw + lane;
#pragma unroll
    for (int wi = 0; wi < W; wi++) {
        const int nb = min(64, P - 64 * wi);
#pragma unroll 8
        for (int bb = 0; bb < nb; bb++) {
            const uint64_t m = 0ull - ((r[wi] >> bb) & 1ull);
            const size_t row = (size_t)(64 * wi + bb) * W;
#pragma unroll
            for (int w = 0; w < W; w++) acc[w] ^= pw[(row + w) * 64] & m;
        }
    }
#pragma unroll
    for (int w = 0; w < W; w++) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) acc[w] ^= __shfl_xor(acc[w], d, 64);
    }
    if (lane == 0) {
        uint64_t any = 0;
#pragma unroll
        for (int w = 0; w < W; w++) any |= acc[w];
        a.valid[frame] = any == 0 ? 1 : 0;
    }
}

__global__ void __launch_bounds__(256) fallback_mask_kernel(const int32_t* trials, const int32_t* valid,
                                                              int32_t* mask, int batch)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch) mask[b] = (trials[b] >= 0 && valid[b] == 1) ? 0 : 1;
}

}  // namespace

hipError_t launch_fallback_mask(const int32_t* trials, const int32_t* valid, int32_t* mask, int batch,
                                hipStream_t stream)
{
    hipLaunchKernelGGL(fallback_mask_kernel, dim3((batch + 255) / 256), dim3(256), 0, stream, trials, valid,
                       mask, batch);
    return hipGetLastError();
}

hipError_t launch_bch_check(const bch_args& a, int batch, hipStream_t stream)
{
    const int W = (a.degree + 63) / 64;
    dim3 grid((batch + 3) / 4);
    switch (W) {
    case 1: hipLaunchKernelGGL(bch_check_kernel<1>, grid, dim3(256), 0, stream, a, batch); break;
    case 2: hipLaunchKernelGGL(bch_check_kernel<2>, grid, dim3(256), 0, stream, a, batch); break;
    case 3: hipLaunchKernelGGL(bch_check_kernel<3>, grid, dim3(256), 0, stream, a, batch); break;
    case 4: hipLaunchKernelGGL(bch_check_kernel<4>, grid, dim3(256), 0, stream, a, batch); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cusat
