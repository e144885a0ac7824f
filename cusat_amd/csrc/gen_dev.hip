// Test-traffic generator on the GPU: the seeded DVB-S2 transmit chain of
// cusat_gen_frames (info bytes -> BCH -> IRA LDPC -> interleave + map ->
// AWGN) with frames produced where the decoder reads them, so that FER
// sweeps at 1e-3 and learned-parameter training are not held up by a host
// generator.
//
// One workgroup (256 threads) per frame, the frame's codeword staged in LDS:
//  1. info bytes: given (info_in), or counter-based random bytes of
//     (seed, frame, byte);
//  2. BCH parity = m(x) x^P mod g(x) by wave 0: lane l folds its chunk of
//     message bytes with the byte table lut[b] = b(x) x^P mod g, moves the
//     partial remainder to its place with the per-lane powers
//     x^(8 (k - end_l) + bit) mod g, and the lanes XOR-reduce (the scheme of
//     the BCH check kernel, over the message instead of the codeword);
//     parity byte j = remainder bits P - 8 j - 1 .. P - 8 j - 8
//     (fec/bch_code.cc:124-143);
//  3. IRA parity (ldpc_encode_kernel.cu:1-17): u[(a + (i mod 360) q) mod R]
//     ^= info_i as LDS atomic XORs, then p_c = u_0 ^ ... ^ u_c as a
//     word-level prefix XOR and a workgroup scan of the word parities;
//  4. symbol j: label bits k at interleaver positions col[k] + j stride
//     (0 past N; the 128APSK padding symbols past the data carry the
//     all-ones label), point + AWGN of sigma = sqrt(N0/2) from a
//     counter-based Gaussian of (seed, frame, symbol).
// The codeword bits and noiseless symbols equal cusat_encode_frames' for
// the same info bytes; the random bytes and the noise are this generator's
// own (a counter-based hash, so independent of the launch shape), not the
// host generator's std::mt19937 streams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.h"
#include "cusat_fec.h"

using namespace cusat;

namespace {

constexpr int GEN_THREADS = 256;

struct gen_args {
    const uint8_t* info_in;
    float* sym_out;
    uint8_t* info_out;
    uint8_t* cw_out;
    const float* pts;
    const uint64_t* lut;
    const uint64_t* pw;   // [64][P][W] lane powers over the message
    const int* row_off;   // [K/360 + 1]
    const int* row_addr;  // accumulator addresses, table order
    uint64_t seed;
    float sigma;
    int noisy;
    int N, K, R, q, kb, P, chunk, xfec, dlen, n_mod, stride, n_points;
    uint32_t cols[8];
};

// splitmix64 finaliser
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t domain, uint64_t frame, uint64_t idx)
{
    return mix64(mix64(seed ^ domain) + 0x9e3779b97f4a7c15ull * (frame + 1) + idx);
}

template <int W>
__global__ void __launch_bounds__(GEN_THREADS) gen_kernel(gen_args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int nw = (a.R + 31) >> 5;  // parity words
    uint8_t* cw = lds;                                             // N/8 bytes
    uint32_t* u = reinterpret_cast<uint32_t*>(lds + ((a.N / 8 + 15) & ~15));
    uint64_t* tab = reinterpret_cast<uint64_t*>(u + ((nw + 3) & ~3));
    uint32_t* carry = reinterpret_cast<uint32_t*>(tab + 256 * W);  // [GEN_THREADS]

    for (int i = t; i < 256 * W; i += GEN_THREADS) tab[i] = a.lut[i];
    for (int i = t; i < nw; i += GEN_THREADS) u[i] = 0;
    for (int i = t; i < a.kb; i += GEN_THREADS) {
        uint8_t v;
        if (a.info_in) {
            v = a.info_in[(size_t)b * a.kb + i];
        } else {
            const uint64_t h = hash3(a.seed, 0x696e666f, (uint64_t)b, (uint64_t)(i >> 3));
            v = (uint8_t)(h >> (8 * (i & 7)));
        }
        cw[i] = v;
    }
    __syncthreads();

    // ---- BCH parity (wave 0) ----
    if (t < 64) {
        const int P = a.P;
        const int top_w = (P - 8) >> 6, top_b = (P - 8) & 63;
        const uint64_t top_mask = (P & 63) ? ((1ull << (P & 63)) - 1) : ~0ull;
        uint64_t r[W];
#pragma unroll
        for (int w = 0; w < W; w++) r[w] = 0;
        const int begin = t * a.chunk, end = min(a.kb, begin + a.chunk);
        for (int i = begin; i < end; i++) {
            const unsigned idx = (unsigned)cw[i] ^ (unsigned)((r[top_w] >> top_b) & 0xffu);
#pragma unroll
            for (int w = W - 1; w >= 0; w--) r[w] = (r[w] << 8) | (w ? r[w - 1] >> 56 : 0);
            r[W - 1] &= top_mask;
#pragma unroll
            for (int w = 0; w < W; w++) r[w] ^= tab[idx * W + w];
        }
        uint64_t acc[W];
#pragma unroll
        for (int w = 0; w < W; w++) acc[w] = 0;
        if (begin < end) {
            const uint64_t* pl = a.pw + (size_t)t * P * W;
            for (int bit = 0; bit < P; bit++)
                if ((r[bit >> 6] >> (bit & 63)) & 1) {
#pragma unroll
                    for (int w = 0; w < W; w++) acc[w] ^= pl[bit * W + w];
                }
        }
#pragma unroll
        for (int w = 0; w < W; w++) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) acc[w] ^= __shfl_xor(acc[w], d, 64);
        }
        // parity byte j = remainder bits lo .. lo + 7, lo = P - 8 (j + 1)
        // (P is a multiple of 8, so a byte never straddles two words)
        for (int j = t; j < P / 8; j += 64) {
            const int lo = P - 8 * (j + 1);
            uint64_t word = 0;
#pragma unroll
            for (int w = 0; w < W; w++)
                if (w == (lo >> 6)) word = acc[w];
            cw[a.kb + j] = (uint8_t)(word >> (lo & 63));
        }
    }
    __syncthreads();

    // ---- IRA accumulators: u[(addr + (i mod 360) q) mod R] ^= bit i ----
    for (int i = t; i < a.K; i += GEN_THREADS) {
        if (!((cw[i >> 3] >> (7 - (i & 7))) & 1)) continue;
        const int row = i / 360, jj = i - row * 360;
        const int e1 = a.row_off[row + 1];
        for (int e = a.row_off[row]; e < e1; e++) {
            int c = a.row_addr[e] + jj * a.q;  // < 2 R: addresses < R, jj q < R
            if (c >= a.R) c -= a.R;
            atomicXor(&u[c >> 5], 1u << (c & 31));
        }
    }
    __syncthreads();

    // ---- p_c = u_0 ^ ... ^ u_c (bit c at word c / 32, bit c % 32) ----
    const int per = (nw + GEN_THREADS - 1) / GEN_THREADS;
    const int w0 = min(nw, t * per), w1 = min(nw, w0 + per);
    uint32_t run = 0;  // parity of this thread's words so far
    for (int w = w0; w < w1; w++) {
        uint32_t x = u[w];
        x ^= x << 1;
        x ^= x << 2;
        x ^= x << 4;
        x ^= x << 8;
        x ^= x << 16;
        x ^= run ? 0xffffffffu : 0u;
        u[w] = x;
        run = x >> 31;
    }
    carry[t] = run;
    __syncthreads();
    if (t == 0) {
        uint32_t c = 0;  // exclusive XOR scan of the threads' parities
        for (int k = 0; k < GEN_THREADS; k++) {
            const uint32_t v = carry[k];
            carry[k] = c;
            c ^= v;
        }
    }
    __syncthreads();
    if (carry[t])
        for (int w = w0; w < w1; w++) u[w] = ~u[w];
    __syncthreads();
    // parity byte k = bits 8k .. 8k + 7, the first one the byte's MSB
    for (int k = t; k < a.R / 8; k += GEN_THREADS) {
        const uint32_t v = (u[k >> 2] >> (8 * (k & 3))) & 0xffu;
        cw[a.K / 8 + k] = (uint8_t)(__brev(v) >> 24);
    }
    __syncthreads();

    // ---- interleave + map + AWGN ----
    const float2* pts = reinterpret_cast<const float2*>(a.pts);
    float2* out = reinterpret_cast<float2*>(a.sym_out) + (size_t)b * a.xfec;
    for (int j = t; j < a.xfec; j += GEN_THREADS) {
        int label;
        if (j < a.dlen) {
            label = 0;
            for (int k = 0; k < a.n_mod; k++) {
                const unsigned pos = a.cols[k] + (unsigned)j * (unsigned)a.stride;
                const int bit = pos < (unsigned)a.N ? (cw[pos >> 3] >> (7 - (pos & 7))) & 1 : 0;
                label |= bit << (a.n_mod - 1 - k);
            }
        } else {
            label = a.n_points - 1;
        }
        float2 s = pts[label];
        if (a.noisy) {
            const uint64_t h = hash3(a.seed, 0x6e6f697365, (uint64_t)b, (uint64_t)j);
            // u1 in (0, 1], u2 in [0, 1): 24-bit grids
            const float u1 = (float)((uint32_t)(h >> 40) + 1u) * (1.0f / 16777216.0f);
            const float u2 = (float)((uint32_t)(h >> 8) & 0xffffffu) * (1.0f / 16777216.0f);
            const float rad = a.sigma * sqrtf(-2.0f * logf(u1));
            float sn, cs;
            sincospif(2.0f * u2, &sn, &cs);
            s.x += rad * cs;
            s.y += rad * sn;
        }
        out[j] = s;
    }
    if (a.info_out)
        for (int i = t; i < a.kb; i += GEN_THREADS) a.info_out[(size_t)b * a.kb + i] = cw[i];
    if (a.cw_out)
        for (int i = t; i < a.N / 8; i += GEN_THREADS) a.cw_out[(size_t)b * (a.N / 8) + i] = cw[i];
}

#define HIP_TRY(expr)                                               \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) {                                     \
            set_last_error("%s: %s", #expr, hipGetErrorString(e_)); \
            return CUSAT_E_HIP;                                     \
        }                                                           \
    } while (0)

// Device tables of the generator, made on first use: accumulator address
// rows and the lane powers of the message-length BCH fold.
int gen_tables(cusat_fec_ctx* c)
{
    if (c->d_gen) return CUSAT_OK;
    const char* tname = ldpc_table_name(c->framesize, c->rate);
    const ldpc_table* tab = tname ? find_table(tname) : nullptr;
    if (!tab) {
        set_last_error("generator: no LDPC table");
        return CUSAT_E_INVALID;
    }
    std::vector<int> off(1, 0), addr;
    for (const auto& row : tab->rows) {
        addr.insert(addr.end(), row.begin(), row.end());
        off.push_back((int)addr.size());
    }
    const int P = c->bch.degree, W = (P + 63) / 64;
    const int kb = c->fp.k_bch / 8;
    const int chunk = (kb + 63) / 64;
    const std::vector<poly256> powers = bch_powers(c->bch, (size_t)kb * 8 + P + 1);
    std::vector<uint64_t> pw((size_t)64 * P * W, 0);
    for (int l = 0; l < 64; l++) {
        const int begin = l * chunk;
        if (begin >= kb) continue;
        const int end = std::min(kb, begin + chunk);
        const size_t shift = (size_t)(kb - end) * 8;
        for (int bit = 0; bit < P; bit++)
            for (int w = 0; w < W; w++) pw[((size_t)l * P + bit) * W + w] = powers[shift + bit][w];
    }
    const size_t b_pw = pw.size() * 8, b_off = off.size() * 4, b_addr = std::max<size_t>(addr.size(), 1) * 4;
    const size_t o_off = (b_pw + 255) & ~(size_t)255, o_addr = o_off + ((b_off + 255) & ~(size_t)255);
    char* d = nullptr;
    HIP_TRY(hipMalloc(&d, o_addr + b_addr));
    hipError_t e = hipMemcpy(d, pw.data(), b_pw, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + o_off, off.data(), b_off, hipMemcpyHostToDevice);
    if (e == hipSuccess && !addr.empty()) e = hipMemcpy(d + o_addr, addr.data(), addr.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(d);
        set_last_error("generator tables: %s", hipGetErrorString(e));
        return CUSAT_E_HIP;
    }
    c->d_gen = d;
    c->d_gen_pw = reinterpret_cast<uint64_t*>(d);
    c->d_gen_off = reinterpret_cast<int*>(d + o_off);
    c->d_gen_addr = reinterpret_cast<int*>(d + o_addr);
    c->gen_chunk = chunk;
    return CUSAT_OK;
}

}  // namespace

extern "C" int cusat_gen_frames_dev(cusat_fec_ctx* c, uint64_t seed, int batch, float es_n0_db,
                                    const uint8_t* info_in, float* sym_out, uint8_t* info_out,
                                    uint8_t* cw_out, void* stream)
{
    if (!c || !sym_out || batch < 0) {
        set_last_error("cusat_gen_frames_dev: bad argument");
        return CUSAT_E_INVALID;
    }
    if (batch == 0) return CUSAT_OK;
    int rc = gen_tables(c);
    if (rc != CUSAT_OK) return rc;
    gen_args a{};
    a.info_in = info_in;
    a.sym_out = sym_out;
    a.info_out = info_out;
    a.cw_out = cw_out;
    a.pts = c->d_pts;
    a.lut = c->d_bch_lut;
    a.pw = c->d_gen_pw;
    a.row_off = c->d_gen_off;
    a.row_addr = c->d_gen_addr;
    a.seed = seed;
    a.noisy = es_n0_db < 1000.0f ? 1 : 0;
    const float n0 = std::pow(10.0f, -es_n0_db / 10.0f);
    a.sigma = std::sqrt(n0 / 2.0f);
    a.N = c->geo.N;
    a.K = c->geo.K;
    a.R = c->geo.R;
    a.q = c->geo.q;
    a.kb = c->fp.k_bch / 8;
    a.P = c->bch.degree;
    a.chunk = c->gen_chunk;
    a.xfec = c->xfec;
    a.dlen = c->dlen;
    a.n_mod = c->con.bits_per_sym;
    a.n_points = c->con.n_points;
    if (c->il.n_cols == 0) {
        a.cols[0] = 0;
        a.cols[1] = 1;
        a.stride = 2;
    } else {
        for (int k = 0; k < 8; k++) a.cols[k] = c->il.col[k];
        a.stride = 1;
    }
    if (c->fp.n_bch != a.K || (a.P & 7) || a.P > 256) {
        set_last_error("cusat_gen_frames_dev: code geometry");
        return CUSAT_E_INVALID;
    }
    const int W = (a.P + 63) / 64;
    const int nw = (a.R + 31) >> 5;
    const size_t lds = ((a.N / 8 + 15) & ~15) + (size_t)((nw + 3) & ~3) * 4 + 256 * W * 8 + GEN_THREADS * 4;
    hipStream_t st = (hipStream_t)stream;
    switch (W) {
    case 1: hipLaunchKernelGGL(gen_kernel<1>, dim3(batch), dim3(GEN_THREADS), lds, st, a); break;
    case 2: hipLaunchKernelGGL(gen_kernel<2>, dim3(batch), dim3(GEN_THREADS), lds, st, a); break;
    case 3: hipLaunchKernelGGL(gen_kernel<3>, dim3(batch), dim3(GEN_THREADS), lds, st, a); break;
    default: hipLaunchKernelGGL(gen_kernel<4>, dim3(batch), dim3(GEN_THREADS), lds, st, a); break;
    }
    HIP_TRY(hipGetLastError());
    return CUSAT_OK;
}
