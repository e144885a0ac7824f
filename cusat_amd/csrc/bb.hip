// Post-BCH baseband step on the GPU: BB descrambling, BBHEADER CRC-8 check
// and GSE data-field extraction of every frame that passed the FEC checks --
// the per-frame host work of rx_fec_chain::post_process_slot
// (rx_fec_chain.cc:565-627), bb_scramble (bb_scrambler.cc) and
// extract_gse_datafield (bbframe_parser.cc:13-56).
//
// * PRBS of EN 302 307-1 clause 5.2.2 (figure 5): 15-stage register, initial
//   content 100101010000000 (stage 1 first), output = stage 14 XOR stage 15,
//   fed back into stage 1; restarted for every BBFRAME; applied MSB first.
// * CRC-8 of clause 5.1.6: generator x^8 + x^7 + x^6 + x^4 + x^2 + 1,
//   register cleared, MSB first; over the 10 header bytes (data + CRC) it is 0.
// * Table 3 fields: MATYPE-1 TS/GS = byte 0 >> 6 (generic packetized = 0),
//   DFL = bytes 4..5 (bits); DFL must be a byte multiple no longer than
//   8 k_bch - 80.
//
// One workgroup per frame.  dfl_out[b] = -1: the frame failed the FEC checks
// (a frame error); 0: no data field (bad header, not GSE, or empty); n > 0:
// n descrambled data-field bytes at out + b * (k_bch/8 - 10).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "bb.h"
#include "ctx.h"
#include "cusat_fec.h"
#include "kernels.h"

namespace cusat {

void bb_prbs(uint8_t* out, size_t n_bytes)
{
    unsigned reg = 0;  // stage s (1..15) in bit s - 1
    const char* init = "100101010000000";
    for (int s = 0; s < 15; s++)
        if (init[s] == '1') reg |= 1u << s;
    for (size_t i = 0; i < n_bytes; i++) {
        unsigned byte = 0;
        for (int b = 0; b < 8; b++) {
            const unsigned o = ((reg >> 13) ^ (reg >> 14)) & 1u;
            reg = ((reg << 1) | o) & 0x7fffu;
            byte = (byte << 1) | o;
        }
        out[i] = (uint8_t)byte;
    }
}

__host__ __device__ uint8_t bb_crc8(const uint8_t* d, int n)
{
    unsigned r = 0;
    for (int i = 0; i < n; i++) {
        r ^= d[i];
        for (int b = 0; b < 8; b++) r = (r & 0x80u) ? ((r << 1) ^ 0xd5u) & 0xffu : (r << 1) & 0xffu;
    }
    return (uint8_t)r;
}

namespace {

// One frame by one 256-thread workgroup: f = its hard bytes, o = its
// data-field slot (kbch_bytes - 10 bytes); the result goes to *dfl_out.
__device__ __forceinline__ void bb_frame(const uint8_t* f, int32_t trials, int32_t valid, const uint8_t* prbs,
                                         int kbch_bytes, uint8_t* o, int32_t* dfl_out, int* n_s)
{
    const int max_dfl = kbch_bytes - 10;
    if (threadIdx.x == 0) {
        int n = -1;
        if (trials >= 0 && valid == 1) {
            uint8_t h[10];
            for (int i = 0; i < 10; i++) h[i] = f[i] ^ prbs[i];
            const int dfl = (h[4] << 8) | h[5];
            n = 0;
            if (bb_crc8(h, 10) == 0 && (h[0] >> 6) == 0 && dfl % 8 == 0 && dfl <= 8 * max_dfl) n = dfl / 8;
        }
        *n_s = n;
        *dfl_out = n;
    }
    __syncthreads();
    const int n = *n_s;
    for (int i = threadIdx.x; i < n; i += 256) o[i] = f[10 + i] ^ prbs[10 + i];
}

__global__ void __launch_bounds__(256)
    bb_extract_kernel(const uint8_t* __restrict__ hard, int hard_stride, const int32_t* __restrict__ trials,
                      const int32_t* __restrict__ valid, const uint8_t* __restrict__ prbs, int kbch_bytes,
                      uint8_t* __restrict__ out, int32_t* __restrict__ dfl_out)
{
    __shared__ int n_s;
    const int b = blockIdx.x;
    bb_frame(hard + (size_t)b * hard_stride, trials[b], valid[b], prbs, kbch_bytes,
             out + (size_t)b * (kbch_bytes - 10), dfl_out + b, &n_s);
}

// Mixed-MODCOD batch: a fixed grid walks the hdr[0] frames (read from
// device memory, so a captured launch serves any batch).
__global__ void __launch_bounds__(256)
    bb_extract_mixed_kernel(const rx_code* __restrict__ codes, const rx_rec* __restrict__ recs,
                            const int* __restrict__ hdr, const uint8_t* __restrict__ hard,
                            const int32_t* __restrict__ trials, const int32_t* __restrict__ valid,
                            uint8_t* __restrict__ out, int32_t* __restrict__ dfl_out)
{
    __shared__ int n_s;
    const int n = hdr[0];
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const rx_rec& r = recs[b];
        const rx_code& c = codes[r.code];
        bb_frame(hard + r.hard_off, trials[b], valid[b], c.prbs, c.kbch_bytes, out + r.out_off, dfl_out + b, &n_s);
        __syncthreads();  // n_s is rewritten by the next frame
    }
}

}  // namespace

hipError_t launch_bb_extract(const uint8_t* hard, int hard_stride, const int32_t* trials, const int32_t* valid,
                             const uint8_t* prbs, int kbch_bytes, uint8_t* out, int32_t* dfl_out, int batch,
                             hipStream_t stream)
{
    hipLaunchKernelGGL(bb_extract_kernel, dim3(batch), dim3(256), 0, stream, hard, hard_stride, trials, valid,
                       prbs, kbch_bytes, out, dfl_out);
    return hipGetLastError();
}

hipError_t launch_bb_extract_mixed(const rx_code* codes, const rx_rec* recs, const int* hdr, const uint8_t* hard,
                                   const int32_t* trials, const int32_t* valid, uint8_t* out, int32_t* dfl_out,
                                   int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(bb_extract_mixed_kernel, dim3(grid), dim3(256), 0, stream, codes, recs, hdr, hard, trials,
                       valid, out, dfl_out);
    return hipGetLastError();
}

}  // namespace cusat

using namespace cusat;

extern "C" int cusat_bb_extract(cusat_fec_ctx* c, const uint8_t* hard, const int32_t* trials,
                                const int32_t* valid, uint8_t* out, int32_t* dfl_out, int batch, void* stream)
{
    if (!c || !hard || !trials || !valid || !out || !dfl_out) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    if (batch < 1 || batch > c->max_batch) {
        set_last_error("batch %d outside 1..%d", batch, c->max_batch);
        return CUSAT_E_BATCH;
    }
    hipError_t e = launch_bb_extract(hard, c->geo.K / 8, trials, valid, c->d_prbs, c->fp.k_bch / 8, out, dfl_out,
                                     batch, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_last_error("bb_extract: %s", hipGetErrorString(e));
        return CUSAT_E_HIP;
    }
    return CUSAT_OK;
}

extern "C" int cusat_bb_prbs(uint8_t* out, int n_bytes)
{
    if (!out || n_bytes < 0) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    bb_prbs(out, (size_t)n_bytes);
    return CUSAT_OK;
}

extern "C" int cusat_bb_crc8(const uint8_t* data, int n_bytes)
{
    if (!data || n_bytes < 0) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    return bb_crc8(data, n_bytes);
}
