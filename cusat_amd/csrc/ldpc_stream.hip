// Persistent decoder of a mixed ACM stream: frames of several codes (normal
// and short FECFRAMEs, any MODCOD of the contexts given) in ONE launch.
//
// A fixed grid of workgroups (one frame at a time each) takes jobs from an
// atomic queue: a job is a run of frames (same code, consecutive in the
// plan's order).  Two builds of the kernel: 6-wave workgroups running
// decode_frame, and 12-wave workgroups running decode_frame_split (the edges
// of a check over two wave groups); the tuning picks one.
// Per frame the workgroup decodes with the code's decoder variant (a switch
// over the compiled instantiations; one-lane short-frame schedules run with
// the second lane idle, exactly as the per-code decoder); in learned mode
// (codes with parameters) a frame whose learned pass does not meet its
// syndrome is decoded again by the plain decoder in the same workgroup.
// The workgroup checks the BCH syndromes of its decoded frames in rounds,
// one wave per frame (bch_check.hip's algorithm, each wave with its code's
// byte table in its own LDS slice), once it holds one frame per wave and
// when the queue is empty.  In learned mode a frame that met its syndrome
// in the learned pass but fails the check is decoded again by the plain
// decoder and checked again in the same round: every frame failing
// frame_passes_fec_checks (rx_fec_chain.h:193-196) after its learned pass
// gets the plain decoder, as cusat_rx_fec_learned does.
// Results equal the per-code launches frame for frame (cusat_ldpc_decode +
// cusat_bch_check, or cusat_rx_fec_learned in learned mode).
//
// Message records live in a per-workgroup workspace (one frame in flight per
// workgroup), not per frame.  The queue counter resets itself when the last
// workgroup leaves, so the launch allocates, clears and copies nothing and
// can be captured in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "ctx.h"
#include "cusat_fec.h"
#include "kernels.h"
#include "ldpc_kernel.h"
#include "ldpc_kernel_split.h"

using namespace cusat;

namespace {

// workgroup size of the 6-wave (WIDE = false) and 12-wave (WIDE = true) kernel
template <bool WIDE>
constexpr int stream_threads()
{
    return WIDE ? 2 * ldpc_dev::SPLIT_LANE_THREADS : 2 * ldpc_dev::LANE_THREADS;
}

struct code_desc {
    const ldpc_dev::circ_rec* circ;
    const int* lay_cnt;
    const int* lay_off;
    const int* steps;
    const int* grp_off;
    const int* step_defer;
    const int* groups;
    const int* beta;  // learned parameters, or null
    const int* alpha;
    const uint64_t* bch_lut;
    const uint64_t* bch_pw;
    int N, K, R, q, n_steps, n_slots, n_groups;
    int max_trials, learned_trials, n_param_iters;
    int variant;        // decode_frame variant (CUSAT_STREAM_VARIANTS id)
    int variant_split;  // decode_frame_split variant (CUSAT_STREAM_SPLIT_VARIANTS id), or -1
    int bch_bytes, bch_degree, bch_chunk, bch_words;
    int pair;  // 1: the 12-wave kernel decodes two frames of this code at once (one-lane schedule)
};

// two-lane decoder variants in the stream kernel: (id, CMAX, exact)
#define CUSAT_STREAM_VARIANTS(X)                                                                    \
    X(0, 2, true) X(1, 3, true) X(2, 4, true) X(3, 5, true) X(4, 7, true) X(5, 8, true)             \
    X(6, 9, true) X(7, 11, true) X(8, 12, true) X(9, 16, true) X(10, 20, true) X(11, 25, true)     \
    X(12, 28, true) X(13, 6, false) X(14, 12, false) X(15, 18, false) X(16, 24, false)              \
    X(17, 30, false)

// two-lane split-edge variants of the 12-wave kernel: (id, CMAX, exact), the
// same set (every code the stream takes has one)
#define CUSAT_STREAM_SPLIT_VARIANTS(X) CUSAT_STREAM_VARIANTS(X)

// split-edge variants with a two-frame form: those of the codes with a
// one-lane schedule (the short FECFRAMEs: at most 25 info circulants per
// layer, the guarded ones at most 17)
#define CUSAT_STREAM_PAIR_VARIANTS(X)                                                                \
    X(0, 2, true) X(1, 3, true) X(2, 4, true) X(3, 5, true) X(4, 7, true) X(5, 8, true)             \
    X(6, 9, true) X(7, 11, true) X(11, 25, true) X(13, 6, false) X(14, 12, false) X(15, 18, false)

bool pair_variant(int id)
{
#define X(ID, C, E) \
    if (id == ID) return true;
    CUSAT_STREAM_PAIR_VARIANTS(X)
#undef X
    return false;
}

int variant_id(int cmax, bool exact)
{
#define X(ID, C, E) \
    if (cmax == C && exact == E) return ID;
    CUSAT_STREAM_VARIANTS(X)
#undef X
    return -1;
}

int split_variant_id(int cmax, bool exact)
{
    if (!ldpc_split_available(cmax, exact, 2)) return -1;
#define X(ID, C, E) \
    if (cmax == C && exact == E) return ID;
    CUSAT_STREAM_SPLIT_VARIANTS(X)
#undef X
    return -1;
}

// The LDS of the stream kernels, at namespace scope so that the decoder
// functions below address it directly (LDS instructions, not flat ones):
// the whole LDS as a static array (FIXED kernels: address known at compile
// time, no base add on every access) or the dynamic allocation.
__shared__ __align__(16) unsigned char stream_lds_fixed[ldpc_dev::SPLIT_LDS_STATIC];
extern __shared__ __align__(16) unsigned char stream_lds_dyn[];

template <bool FIXED>
__device__ __forceinline__ unsigned char* stream_lds()
{
    if constexpr (FIXED)
        return stream_lds_fixed;
    else
        return stream_lds_dyn;
}

// Address spaces for pointers that cross into the decoder functions.  A
// kernel knows its pointer arguments (and pointers loaded through them) are
// global; a device function sees generic pointers and would use flat
// instructions, and no scalar loads.  The read-only tables are constant
// memory (uniform loads become scalar loads), outputs global memory.
#define CUSAT_AS_GLOBAL __attribute__((address_space(1)))
#define CUSAT_AS_CONST __attribute__((address_space(4)))

// code_desc as the decoder functions read it: the same layout, the tables
// in constant memory
struct code_desc_dev {
    CUSAT_AS_CONST const ldpc_dev::circ_rec* circ;
    CUSAT_AS_CONST const int* lay_cnt;
    CUSAT_AS_CONST const int* lay_off;
    CUSAT_AS_CONST const int* steps;
    CUSAT_AS_CONST const int* grp_off;
    CUSAT_AS_CONST const int* step_defer;
    CUSAT_AS_CONST const int* groups;
    CUSAT_AS_CONST const int* beta;
    CUSAT_AS_CONST const int* alpha;
    CUSAT_AS_CONST const uint64_t* bch_lut;
    CUSAT_AS_CONST const uint64_t* bch_pw;
    int N, K, R, q, n_steps, n_slots, n_groups;
    int max_trials, learned_trials, n_param_iters;
    int variant;
    int variant_split;
    int bch_bytes, bch_degree, bch_chunk, bch_words;
    int pair;
};
static_assert(sizeof(code_desc_dev) == sizeof(code_desc), "code_desc_dev mirrors code_desc");

// Function arguments arrive in VGPRs; wave-uniform values are made uniform
// again so that loads through them stay scalar and branches on them stay
// scalar branches.
template <class T>
__device__ __forceinline__ T* uniform(T* ptr)
{
    const uint64_t v = reinterpret_cast<uint64_t>(ptr);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// One decoder variant of one frame, as a function of its own: every variant
// gets its own register allocation, as in the per-variant batch kernels.
// With all variants inlined into the stream kernel the allocator spilled
// (over 100 VGPRs at the 12-wave register budget) and every code paid for
// it.  SPLIT: 12-wave split-edge decoder, else the 6-wave decode_frame.
template <bool SPLIT, int C, bool E, bool FIXED>
__device__ __attribute__((noinline)) int decode_one(CUSAT_AS_CONST const code_desc_dev* codes_v, int ci_v,
                                                    CUSAT_AS_CONST const int16_t* llr_v, CUSAT_AS_GLOBAL uint4* rec_v,
                                                    CUSAT_AS_GLOBAL uint8_t* h_v, int lp_v)
{
    const auto& c = uniform(codes_v)[__builtin_amdgcn_readfirstlane(ci_v)];
    const bool lp = __builtin_amdgcn_readfirstlane(lp_v) != 0;
    const ldpc_dev::kparams p{c.N,      c.K, c.R, c.q, c.n_steps, lp ? c.learned_trials : c.max_trials,
                              lp ? c.n_param_iters : 0, c.n_slots, c.n_groups};
    // generic pointers for the decoder; once it is inlined here, the
    // compiler carries the address spaces through to every access
    const int16_t* llr = (const int16_t*)uniform(llr_v);
    uint4* rec = (uint4*)uniform(rec_v);
    uint8_t* h = (uint8_t*)uniform(h_v);
    const auto* circ = (const ldpc_dev::circ_rec*)c.circ;
    const int* beta = lp ? (const int*)c.beta : nullptr;
    const int* alpha = lp ? (const int*)c.alpha : nullptr;
    if constexpr (SPLIT)
        return ldpc_dev::decode_frame_split<C, E, 2>(stream_lds<FIXED>(), llr, rec, h, circ, (const int*)c.lay_cnt,
                                                     (const int*)c.lay_off, (const int*)c.steps,
                                                     (const int*)c.grp_off, (const int*)c.step_defer,
                                                     (const int*)c.groups, beta, alpha, p);
    else
        return ldpc_dev::decode_frame<C, E, 2>(stream_lds<FIXED>(), llr, rec, h, circ, (const int*)c.lay_cnt,
                                               (const int*)c.lay_off, (const int*)c.steps, (const int*)c.grp_off,
                                               (const int*)c.groups, beta, alpha, p);
}

// 12-wave workgroups: the split-edge decoder; 6-wave: decode_frame
template <bool WIDE, bool FIXED>
__device__ __forceinline__ int decode_variant(const code_desc* codes_g, int ci, const code_desc& c,
                                              const int16_t* llr_g, uint4* rec_g, uint8_t* h_g, int lp)
{
    const auto* codes = (CUSAT_AS_CONST const code_desc_dev*)codes_g;
    const auto* llr = (CUSAT_AS_CONST const int16_t*)llr_g;
    auto* rec = (CUSAT_AS_GLOBAL uint4*)rec_g;
    auto* h = (CUSAT_AS_GLOBAL uint8_t*)h_g;
    if (WIDE) {
        switch (c.variant_split) {
#define X(ID, C, E) \
    case ID:        \
        return decode_one<true, C, E, FIXED>(codes, ci, llr, rec, h, lp);
            CUSAT_STREAM_SPLIT_VARIANTS(X)
#undef X
        }
        return -1;
    }
    switch (c.variant) {
#define X(ID, C, E) \
    case ID:        \
        return decode_one<false, C, E, FIXED>(codes, ci, llr, rec, h, lp);
        CUSAT_STREAM_VARIANTS(X)
#undef X
    }
    return -1;
}

// Two frames of a code with a one-lane schedule at once, one per lane
// (decode_frame_split's PAIR form), as a function of its own as decode_one.
// Returns frame 0's trial count in the low 16 bits, frame 1's in the high
// 16 bits.
template <int C, bool E, bool FIXED>
__device__ __attribute__((noinline)) int decode_two(CUSAT_AS_CONST const code_desc_dev* codes_v, int ci_v,
                                                    CUSAT_AS_CONST const int16_t* llr0_v, CUSAT_AS_GLOBAL uint4* rec0_v,
                                                    CUSAT_AS_GLOBAL uint8_t* h0_v,
                                                    CUSAT_AS_CONST const int16_t* llr1_v, CUSAT_AS_GLOBAL uint4* rec1_v,
                                                    CUSAT_AS_GLOBAL uint8_t* h1_v, int lp_v)
{
    const auto& c = uniform(codes_v)[__builtin_amdgcn_readfirstlane(ci_v)];
    const bool lp = __builtin_amdgcn_readfirstlane(lp_v) != 0;
    const ldpc_dev::kparams p{c.N,      c.K, c.R, c.q, c.n_steps, lp ? c.learned_trials : c.max_trials,
                              lp ? c.n_param_iters : 0, c.n_slots, c.n_groups};
    const auto* circ = (const ldpc_dev::circ_rec*)c.circ;
    const int* beta = lp ? (const int*)c.beta : nullptr;
    const int* alpha = lp ? (const int*)c.alpha : nullptr;
    return ldpc_dev::decode_frame_split<C, E, 2, true>(
        stream_lds<FIXED>(), (const int16_t*)uniform(llr0_v), (uint4*)uniform(rec0_v), (uint8_t*)uniform(h0_v), circ,
        (const int*)c.lay_cnt, (const int*)c.lay_off, (const int*)c.steps, (const int*)c.grp_off,
        (const int*)c.step_defer, (const int*)c.groups, beta, alpha, p, (const int16_t*)uniform(llr1_v),
        (uint4*)uniform(rec1_v), (uint8_t*)uniform(h1_v));
}

template <bool FIXED>
__device__ __forceinline__ int decode_variant_two(const code_desc* codes_g, int ci, const code_desc& c,
                                                  const int16_t* llr0, uint4* rec0, uint8_t* h0,
                                                  const int16_t* llr1, uint4* rec1, uint8_t* h1, int lp)
{
    const auto* codes = (CUSAT_AS_CONST const code_desc_dev*)codes_g;
    switch (c.variant_split) {
#define X(ID, C, E)                                                                                                \
    case ID:                                                                                                       \
        return decode_two<C, E, FIXED>(codes, ci, (CUSAT_AS_CONST const int16_t*)llr0, (CUSAT_AS_GLOBAL uint4*)rec0, \
                                       (CUSAT_AS_GLOBAL uint8_t*)h0, (CUSAT_AS_CONST const int16_t*)llr1,           \
                                       (CUSAT_AS_GLOBAL uint4*)rec1, (CUSAT_AS_GLOBAL uint8_t*)h1, lp);
        CUSAT_STREAM_PAIR_VARIANTS(X)
#undef X
    }
    return -1;  // both halves -1 (the host never marks another variant as a pair)
}

// c(x) mod g(x) == 0 over the frame's n_bytes hard bytes, by wave 0 (the
// algorithm of bch_check.hip: lane chunks folded with the byte table, moved
// to their place with x^(8 (n - end)) mod g, XOR-reduced over the wave).
template <int W>
__device__ __forceinline__ int bch_wave(const uint8_t* __restrict__ src, const uint64_t* tab,
                                        const uint64_t* __restrict__ pw, int n_bytes, int P, int chunk,
                                        int lane)
{
    constexpr int TOPW = W - 1;  // (P - 8) >> 6 == W - 1 for every DVB-S2 BCH degree
    const int top_b = (P - 8) & 63;
    const uint64_t top_mask = (P & 63) ? ((1ull << (P & 63)) - 1) : ~0ull;
    uint64_t r[W];
#pragma unroll
    for (int w = 0; w < W; w++) r[w] = 0;
    const int begin = lane * chunk;
    const int end = min(n_bytes, begin + chunk);
    for (int i = begin; i < end; i++) {
        const unsigned idx = (unsigned)src[i] ^ (unsigned)((r[TOPW] >> top_b) & 0xffu);
#pragma unroll
        for (int w = W - 1; w >= 0; w--) r[w] = (r[w] << 8) | (w ? r[w - 1] >> 56 : 0);
        r[W - 1] &= top_mask;
#pragma unroll
        for (int w = 0; w < W; w++) r[w] ^= tab[idx * W + w];
    }
    uint64_t acc[W];
#pragma unroll
    for (int w = 0; w < W; w++) acc[w] = 0;
    // pw rows are [bit][word][lane] (one row's load is 512 contiguous bytes
    // for the wave).  Branch-free, 8 rows' loads issued before their uses.
    const uint64_t* pl = pw + lane;
#pragma unroll
    for (int wi = 0; wi < W; wi++) {
        const int nb = min(64, P - 64 * wi);
        for (int b0 = 0; b0 < nb; b0 += 8) {
            uint64_t v[8][W];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const size_t row = (size_t)(64 * wi + b0 + j) * W;
#pragma unroll
                for (int w = 0; w < W; w++) v[j][w] = b0 + j < nb ? pl[(row + w) * 64] : 0ull;
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint64_t m = 0ull - ((r[wi] >> (b0 + j)) & 1ull);
#pragma unroll
                for (int w = 0; w < W; w++) acc[w] ^= v[j][w] & m;
            }
        }
    }
    uint64_t any = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) acc[w] ^= __shfl_xor(acc[w], d, 64);
        any |= acc[w];
    }
    return any == 0 ? 1 : 0;
}

// Whole workgroup: BCH check of the hard bytes at h (written by this
// workgroup before the call); returns 1 = valid, the same value in every thread.
template <int NT>
__device__ __forceinline__ int bch_frame(unsigned char* lds, int* flag, const uint8_t* __restrict__ h,
                                         const uint64_t* __restrict__ lut, const uint64_t* __restrict__ pw,
                                         int n_bytes, int degree, int chunk, int words)
{
    uint64_t* tab = reinterpret_cast<uint64_t*>(lds);
    __syncthreads();  // the decoder is done with the LDS; h is written
    for (int k = threadIdx.x; k < 256 * words; k += NT) tab[k] = lut[k];
    __syncthreads();
    if (threadIdx.x < 64) {
        int ok = 0;
        switch (words) {
        case 1: ok = bch_wave<1>(h, tab, pw, n_bytes, degree, chunk, threadIdx.x); break;
        case 2: ok = bch_wave<2>(h, tab, pw, n_bytes, degree, chunk, threadIdx.x); break;
        case 3: ok = bch_wave<3>(h, tab, pw, n_bytes, degree, chunk, threadIdx.x); break;
        default: ok = bch_wave<4>(h, tab, pw, n_bytes, degree, chunk, threadIdx.x); break;
        }
        if (threadIdx.x == 0) *flag = ok;
    }
    __syncthreads();
    const int ok = *flag;
    __syncthreads();  // flag read by all before it is reused
    return ok;
}

// The BCH check of one frame as a function of its own, for the same
// reasons as decode_one: inside the kernel body, at the kernel's register
// limit, the table loads were generic (flat) and each waited for the one
// before.  ctl_off: the byte offset of the control ints in the LDS.
template <int NT, bool FIXED>
__device__ __attribute__((noinline)) int bch_one(CUSAT_AS_CONST const code_desc_dev* codes_v, int ci_v,
                                                 CUSAT_AS_GLOBAL const uint8_t* h_v, int ctl_off_v)
{
    const auto& c = uniform(codes_v)[__builtin_amdgcn_readfirstlane(ci_v)];
    unsigned char* lds = stream_lds<FIXED>();
    int* flag = reinterpret_cast<int*>(lds + __builtin_amdgcn_readfirstlane(ctl_off_v)) + 1;
    return bch_frame<NT>(lds, flag, (const uint8_t*)uniform(h_v), (const uint64_t*)c.bch_lut,
                         (const uint64_t*)c.bch_pw, c.bch_bytes, c.bch_degree, c.bch_chunk, c.bch_words);
}

// LDS control area of the stream kernel, at ctl_off (after the decoders'
// own LDS): the queue slot, bch_one's flag, and the frames this workgroup
// has decoded whose BCH check is pending (frame, trials, redone, redo flag)
constexpr int CTL_BYTES = 256;
constexpr int BCH_SLICE = 256 * 4 * 8;  // a code's byte table (at most 4 words per entry)
template <int PEND>
struct stream_ctl {
    int job, bch_flag;
    int frame[PEND], res[PEND], fb[PEND], redo[PEND];
};
static_assert(sizeof(stream_ctl<12>) <= CTL_BYTES, "control area");

// ctl_off: byte offset in the LDS of the control area.  FIXED: the whole
// LDS as a static array (address known at compile time, no base add on
// every LDS access), for streams whose LDS allows one workgroup per CU
// anyway.
//
// A workgroup decodes its frames one after the other and checks their BCH
// syndromes in rounds, one wave per frame, each wave with its code's byte
// table in its own LDS slice: when it holds one decoded frame per wave, and
// when the queue is empty.  A frame of a learned code that met its
// syndrome in the learned pass but fails the check is decoded again by the
// plain decoder and checked again (bch_one) in the same round.
template <bool WIDE, bool FIXED = false>
__global__ void __launch_bounds__(stream_threads<WIDE>())
    ldpc_stream_kernel(const code_desc* __restrict__ codes, const cusat_stream_frame* __restrict__ frames,
                       const int2* __restrict__ jobs, const int* __restrict__ order,
                       unsigned* counters, const int16_t* __restrict__ llr, uint8_t* __restrict__ hard,
                       int32_t* __restrict__ trials, int32_t* __restrict__ valid,
                       int32_t* __restrict__ fallback, uint4* __restrict__ ws, size_t ws_v4, int learned,
                       int ctl_off)
{
    constexpr int NT = stream_threads<WIDE>();
    constexpr int PEND = NT / 64;  // one pending frame per wave
    unsigned char* lds_raw = stream_lds<FIXED>();
    auto& ctl = *reinterpret_cast<stream_ctl<PEND>*>(lds_raw + ctl_off);
    const int slices = min(PEND, ctl_off / BCH_SLICE);  // byte tables that fit below the control area
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint4* rec = ws + (size_t)blockIdx.x * ws_v4;
    int npend = 0;
    auto bch_round = [&]() {
        __syncthreads();  // the pending list is written; the decoder's LDS is free
        for (int g = 0; g < npend; g += slices) {
            const int k = g + wv;
            const bool live = wv < slices && k < npend;
            const int f = live ? __builtin_amdgcn_readfirstlane(ctl.frame[k]) : 0;
            const int ci = live ? __builtin_amdgcn_readfirstlane(frames[f].code) : 0;
            const code_desc& c = codes[ci];
            uint64_t* tab = reinterpret_cast<uint64_t*>(lds_raw + wv * BCH_SLICE);
            if (live)
                for (int x = lane; x < 256 * c.bch_words; x += 64) tab[x] = c.bch_lut[x];
            __syncthreads();
            if (live) {
                const uint8_t* h = hard + frames[f].hard_offset;
                int ok;
                switch (c.bch_words) {
                case 1: ok = bch_wave<1>(h, tab, c.bch_pw, c.bch_bytes, c.bch_degree, c.bch_chunk, lane); break;
                case 2: ok = bch_wave<2>(h, tab, c.bch_pw, c.bch_bytes, c.bch_degree, c.bch_chunk, lane); break;
                case 3: ok = bch_wave<3>(h, tab, c.bch_pw, c.bch_bytes, c.bch_degree, c.bch_chunk, lane); break;
                default: ok = bch_wave<4>(h, tab, c.bch_pw, c.bch_bytes, c.bch_degree, c.bch_chunk, lane); break;
                }
                if (lane == 0) {
                    valid[f] = ok;
                    ctl.redo[k] = learned && c.learned_trials > 0 && !ok && ctl.res[k] >= 0 && !ctl.fb[k];
                }
            }
            __syncthreads();  // the slices are reused; the redo flags are written
        }
        for (int k = 0; k < npend; k++) {
            if (!__builtin_amdgcn_readfirstlane(ctl.redo[k])) continue;
            const int f = __builtin_amdgcn_readfirstlane(ctl.frame[k]);
            const int ci = __builtin_amdgcn_readfirstlane(frames[f].code);
            const code_desc& c = codes[ci];
            uint8_t* fh = hard + frames[f].hard_offset;
            const int res = decode_variant<WIDE, FIXED>(codes, ci, c, llr + frames[f].llr_offset, rec, fh, 0);
            const int ok = bch_one<NT, FIXED>((CUSAT_AS_CONST const code_desc_dev*)codes, ci,
                                              (CUSAT_AS_GLOBAL const uint8_t*)fh, ctl_off);
            if (threadIdx.x == 0) {
                trials[f] = res;
                valid[f] = ok;
                if (fallback) fallback[f] = 1;
            }
        }
        __syncthreads();  // the pending list is reused
        npend = 0;
    };
    for (;;) {
        if (threadIdx.x == 0) ctl.job = (int)atomicAdd(&counters[0], 1u);
        __syncthreads();
        const int jb = __builtin_amdgcn_readfirstlane(ctl.job);
        __syncthreads();
        if (jb >= (int)counters[2]) break;  // the plan's job count (device: graph replays see new plans)
        const int2 J = jobs[jb];
        for (int k = 0; k < J.y; k++) {
            const int f = __builtin_amdgcn_readfirstlane(order[J.x + k]);
            const int ci = __builtin_amdgcn_readfirstlane(frames[f].code);
            const code_desc& c = codes[ci];
            const int16_t* fl = llr + frames[f].llr_offset;
            uint8_t* fh = hard + frames[f].hard_offset;
            const bool use_learned = learned && c.learned_trials > 0;
            if constexpr (WIDE && FIXED) {
                if (c.pair && k + 1 < J.y) {
                    // two frames of the job (one code) at once, one per lane;
                    // frame 1's message records in the workspace's second half
                    const int f1 = __builtin_amdgcn_readfirstlane(order[J.x + k + 1]);
                    const int two = decode_variant_two<FIXED>(codes, ci, c, fl, rec, fh, llr + frames[f1].llr_offset,
                                                              rec + ws_v4 / 2, hard + frames[f1].hard_offset,
                                                              use_learned ? 1 : 0);
                    __syncthreads();
                    for (int m = 0; m < 2; m++) {
                        const int fm = m ? f1 : f;
                        int res = m ? two >> 16 : (int)(int16_t)(two & 0xffff);
                        const int fb = use_learned && res < 0 ? 1 : 0;
                        if (fb) {
                            res = decode_variant<WIDE, FIXED>(codes, ci, c, llr + frames[fm].llr_offset, rec,
                                                              hard + frames[fm].hard_offset, 0);
                            __syncthreads();
                        }
                        if (threadIdx.x == 0) {
                            trials[fm] = res;
                            if (fallback) fallback[fm] = fb;
                            ctl.frame[npend] = fm;
                            ctl.res[npend] = res;
                            ctl.fb[npend] = fb;
                        }
                        if (++npend == PEND) bch_round();
                    }
                    k++;
                    continue;
                }
            }
            // learned pass, then the plain decoder at once when the
            // syndrome is not met (a barrier after each decode: its last
            // step packs the hard decisions from the LDS that the next
            // decode starts writing)
            int res = decode_variant<WIDE, FIXED>(codes, ci, c, fl, rec, fh, use_learned ? 1 : 0);
            __syncthreads();
            const int fb = use_learned && res < 0 ? 1 : 0;
            if (fb) {
                res = decode_variant<WIDE, FIXED>(codes, ci, c, fl, rec, fh, 0);
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                trials[f] = res;
                if (fallback) fallback[f] = fb;
                ctl.frame[npend] = f;
                ctl.res[npend] = res;
                ctl.fb[npend] = fb;
            }
            if (++npend == PEND) bch_round();
        }
    }
    if (npend) bch_round();
    // the last workgroup to leave resets the queue for the next launch
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&counters[1], 1u) == gridDim.x - 1) {
            atomicExch(&counters[0], 0u);
            atomicExch(&counters[1], 0u);
        }
    }
}

#define HIP_TRY(expr)                                               \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) {                                     \
            set_last_error("%s: %s", #expr, hipGetErrorString(e_)); \
            return CUSAT_E_HIP;                                     \
        }                                                           \
    } while (0)

}  // namespace

struct cusat_stream_ctx {
    std::vector<cusat_fec_ctx*> codes;
    std::vector<int> variant;        // stream decoder variant of each code
    std::vector<int> variant_split;  // its split-edge variant, or -1
    std::vector<int> pair;           // 1: two frames per workgroup (code_desc::pair)
    int max_frames = 0;
    unsigned flags = 0;
    int n_cu = 0;
    int max_wg[2] = {0, 0};  // workgroups per CU of the 6-wave / 12-wave kernel (0: does not fit)
    int lds[2] = {0, 0};     // dynamic LDS bytes of each, control ints included
    size_t ws_v4 = 0;  // uint4 per workgroup
    // device
    code_desc* d_codes = nullptr;
    cusat_stream_frame* d_frames = nullptr;
    int2* d_jobs = nullptr;
    int* d_order = nullptr;
    unsigned* d_counters = nullptr;
    uint4* d_ws = nullptr;
    // plan region (one allocation, this layout on both sides): 256 bytes of
    // counters ([2] = job count), frames, jobs, order
    size_t off_jobs = 0, off_order = 0, plan_bytes = 0;
    static constexpr int N_STAGE = 4;
    // pinned images for the async plan calls: the plan region, then the
    // code descriptors (stage_bytes = plan_bytes + descriptors)
    unsigned char* stage[N_STAGE] = {};
    hipEvent_t stage_done[N_STAGE] = {};
    bool stage_used[N_STAGE] = {};
    int next_stage = 0;
    size_t stage_bytes = 0;
    // param_gen of each context when its descriptor was last uploaded: a
    // plan call re-uploads the descriptors when a context's learned
    // parameters changed since
    std::vector<unsigned> code_gen;
    // receive batches (cusat_stream_rx): per-code records after d_codes, the
    // rx header (256 bytes: frames, demap blocks) and records at off_rx in the
    // plan region, LLR and hard-decision work buffers (all allocated by
    // cusat_stream_create)
    rx_code* d_rx_codes = nullptr;
    size_t off_rx = 0, off_dl = 0, off_ev = 0;  // rx header + records; demap order; SNR events
    int max_events = 0;
    rx_scale* d_snap = nullptr;     // per-code demapper scales of the current batch
    float* d_est = nullptr;         // per-event decision-directed estimates
    int64_t llr_max = 0, hard_max = 0;  // per frame
    int16_t* d_llr_work = nullptr;
    uint8_t* d_hard_work = nullptr;
    bool rx_planned = false;  // the current plan is an rx plan
    // plan
    int n_frames = 0, n_jobs = 0;
    std::vector<cusat_stream_frame> frames;
    std::vector<int> demap_order;  // host scratch of the rx plan
    cusat_stream_tuning tuning{};
};

namespace {

int frame_cost(const cusat_fec_ctx* c)
{
    // work of one iteration: steps x (info circulants + parity) of the variant
    return c->kern_steps * (c->cmax + 2);
}

// Host image of the plan region for the frames sc->frames under tuning tu:
// counters (queue and exit count 0, [2] the job count), the frames, the jobs
// (first position in order, frame count) and the decode order.  Returns the
// job count.
int stage_plan(const cusat_stream_ctx* sc, const cusat_stream_tuning& tu, unsigned char* img)
{
    const int n = sc->n_frames;
    int* order = reinterpret_cast<int*>(img + sc->off_order);
    int2* jobs = reinterpret_cast<int2*>(img + sc->off_jobs);
    std::iota(order, order + n, 0);
    // codes of equal cost kept apart, so that a code's frames are neighbours
    if (tu.order == CUSAT_STREAM_ORDER_COST)
        std::stable_sort(order, order + n, [&](int a, int b) {
            const int ca = sc->frames[a].code, cb = sc->frames[b].code;
            const int xa = frame_cost(sc->codes[ca]), xb = frame_cost(sc->codes[cb]);
            return xa != xb ? xa > xb : ca < cb;
        });
    // a code decoded two frames at a time takes twice the frames per job,
    // when the plan has more frames than the grid has workgroups (else every
    // frame has a workgroup of its own, and a pair would only be slower)
    const bool pairs = tu.split_edges && n > sc->n_cu * tu.wg_per_cu;
    int nj = 0;
    for (int k = 0; k < n;) {
        const int code = sc->frames[order[k]].code;
        const int per = pairs && sc->pair[code] ? 2 * tu.frames_per_job : tu.frames_per_job;
        int m = 1;
        while (m < per && k + m < n && sc->frames[order[k + m]].code == code) m++;
        jobs[nj++] = make_int2(k, m);
        k += m;
    }
    unsigned* ctr = reinterpret_cast<unsigned*>(img);
    memset(ctr, 0, 256);
    ctr[2] = (unsigned)nj;
    if (n) memcpy(img + 256, sc->frames.data(), n * sizeof(cusat_stream_frame));
    return nj;
}

// Copy the used parts of a plan image (n frames, nj jobs) to the device.
hipError_t copy_plan(const cusat_stream_ctx* sc, const unsigned char* img, int n, int nj, hipStream_t st, bool async)
{
    unsigned char* d = reinterpret_cast<unsigned char*>(sc->d_counters);
    const size_t part[3][2] = {{0, 256 + n * sizeof(cusat_stream_frame)},
                               {sc->off_jobs, nj * sizeof(int2)},
                               {sc->off_order, n * sizeof(int)}};
    for (const auto& q : part) {
        if (!q[1]) continue;
        const hipError_t e = async ? hipMemcpyAsync(d + q[0], img + q[0], q[1], hipMemcpyHostToDevice, st)
                                   : hipMemcpy(d + q[0], img + q[0], q[1], hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int build_plan(cusat_stream_ctx* sc, const cusat_stream_tuning& tu)
{
    std::vector<unsigned char> img(sc->plan_bytes);
    const int nj = stage_plan(sc, tu, img.data());
    HIP_TRY(copy_plan(sc, img.data(), sc->n_frames, nj, nullptr, false));
    sc->n_jobs = nj;
    sc->tuning = tu;
    return CUSAT_OK;
}

// The code descriptors as the kernel reads them, from the contexts' current
// tables and learned parameters.
void fill_code_descs(const cusat_stream_ctx* sc, code_desc* d)
{
    for (size_t k = 0; k < sc->codes.size(); k++) {
        const cusat_fec_ctx* c = sc->codes[k];
        code_desc& o = d[k];
        memset(&o, 0, sizeof o);
        o.circ = reinterpret_cast<const ldpc_dev::circ_rec*>(c->d_circ);
        o.lay_cnt = c->d_lay_cnt;
        o.lay_off = c->d_lay_off;
        o.steps = c->d_steps;
        o.grp_off = c->d_grp_off;
        o.step_defer = c->d_step_defer;
        o.groups = c->d_groups;
        o.beta = c->n_param_iters > 0 ? c->d_beta : nullptr;
        o.alpha = c->n_param_iters > 0 ? c->d_alpha : nullptr;
        o.bch_lut = c->d_bch_lut;
        o.bch_pw = c->d_bch_pw;
        o.N = c->geo.N;
        o.K = c->geo.K;
        o.R = c->geo.R;
        o.q = c->geo.q;
        o.n_steps = c->kern_steps;
        o.n_slots = c->geo.n_slots;
        o.n_groups = (int)c->geo.groups.size();
        o.max_trials = c->max_trials;
        o.n_param_iters = c->n_param_iters;
        o.learned_trials = c->n_param_iters;
        o.variant = sc->variant[k];
        o.variant_split = sc->variant_split[k];
        o.bch_bytes = c->fp.n_bch / 8;
        o.bch_degree = c->bch.degree;
        o.bch_chunk = c->bch_chunk;
        o.bch_words = (c->bch.degree + 63) / 64;
        o.pair = sc->pair[k];
    }
}

bool params_changed(const cusat_stream_ctx* sc)
{
    for (size_t k = 0; k < sc->codes.size(); k++)
        if (sc->codes[k]->param_gen != sc->code_gen[k]) return true;
    return false;
}

void note_params(cusat_stream_ctx* sc)
{
    for (size_t k = 0; k < sc->codes.size(); k++) sc->code_gen[k] = sc->codes[k]->param_gen;
}

// Synchronous upload of the code descriptors (and, at creation, of the
// receive-side per-code records).
int upload_codes(cusat_stream_ctx* sc, bool rx_codes)
{
    std::vector<code_desc> d(sc->codes.size());
    fill_code_descs(sc, d.data());
    HIP_TRY(hipMemcpy(sc->d_codes, d.data(), d.size() * sizeof(code_desc), hipMemcpyHostToDevice));
    note_params(sc);
    if (!rx_codes) return CUSAT_OK;
    std::vector<rx_code> r(sc->codes.size());
    for (size_t k = 0; k < sc->codes.size(); k++) {
        const cusat_fec_ctx* c = sc->codes[k];
        rx_code& o = r[k];
        memset(&o, 0, sizeof o);
        const bool formula = c->modulation == CUSAT_MOD_QPSK || c->modulation == CUSAT_MOD_8PSK;
        o.pts = formula ? nullptr : c->d_pts;
        o.prbs = c->d_prbs;
        o.dd_pts = c->d_pts;
        o.grid_scale = c->grid_scale;
        o.span = demap_span(!formula, c->con.bits_per_sym);
        o.has_cap = c->llr_target > 0.0f && c->dmax > 0.0f;
        o.gain_cap = o.has_cap ? 2.0f * c->llr_target / c->dmax : 0.0f;  // as cusat_fec_demap
        o.xfec = c->xfec;
        o.dlen = c->dlen;
        o.N = c->geo.N;
        o.n_mod = c->con.bits_per_sym;
        o.n_points = c->con.n_points;
        o.kbch_bytes = c->fp.k_bch / 8;
        o.hard_bytes = c->geo.K / 8;
        for (int b = 0; b < 8; b++) o.col[b] = c->il.col[b];
    }
    HIP_TRY(hipMemcpy(sc->d_rx_codes, r.data(), r.size() * sizeof(rx_code), hipMemcpyHostToDevice));
    return CUSAT_OK;
}

// Take the next pinned staging image (waiting only if its previous copy has
// not run yet) and, when a context's learned parameters changed since the
// last upload, stage the new descriptors behind the plan and copy them on
// `st`, ahead of the plan that will use them.
int next_stage_image(cusat_stream_ctx* sc, hipStream_t st, unsigned char** img)
{
    const int k = sc->next_stage;
    if (sc->stage_used[k]) HIP_TRY(hipEventSynchronize(sc->stage_done[k]));  // its last copy has run
    *img = sc->stage[k];
    if (params_changed(sc)) {
        code_desc* d = reinterpret_cast<code_desc*>(sc->stage[k] + sc->plan_bytes);
        fill_code_descs(sc, d);
        HIP_TRY(hipMemcpyAsync(sc->d_codes, d, sc->codes.size() * sizeof(code_desc), hipMemcpyHostToDevice, st));
        note_params(sc);
    }
    return CUSAT_OK;
}

// The staging image taken by next_stage_image has been copied on `st`.
int release_stage_image(cusat_stream_ctx* sc, hipStream_t st)
{
    const int k = sc->next_stage;
    HIP_TRY(hipEventRecord(sc->stage_done[k], st));
    sc->stage_used[k] = true;
    sc->next_stage = (k + 1) % cusat_stream_ctx::N_STAGE;
    return CUSAT_OK;
}

}  // namespace

// raise the kernel's dynamic-LDS limit to `lds` (the attribute is per
// function: only ever raised, another stream may need more) and report how
// many workgroups fit on a CU (0 if none)
template <bool WIDE>
hipError_t fit_kernel(int lds, int* occ)
{
    static int configured = 0;
    auto kern = &ldpc_stream_kernel<WIDE>;
    *occ = 0;
    if (WIDE && lds > ldpc_dev::SPLIT_LDS_STATIC / 2 && lds <= ldpc_dev::SPLIT_LDS_STATIC) {
        // one workgroup per CU either way: the static-LDS build
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, &ldpc_stream_kernel<true, true>,
                                                            stream_threads<true>(), 0);
    }
    if (lds > __atomic_load_n(&configured, __ATOMIC_ACQUIRE)) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return hipSuccess;  // does not fit: occ stays 0
        }
        __atomic_store_n(&configured, lds, __ATOMIC_RELEASE);
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, kern, stream_threads<WIDE>(), lds);
}

extern "C" int cusat_stream_create(cusat_stream_ctx** out, cusat_fec_ctx* const* codes, int n_codes,
                                   int max_frames, unsigned flags)
{
    if (!out || !codes || n_codes < 1 || n_codes > 1024 || max_frames < 1) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    *out = nullptr;
    cusat_stream_ctx* sc = new (std::nothrow) cusat_stream_ctx();
    if (!sc) return CUSAT_E_NOMEM;
    sc->max_frames = max_frames;
    sc->flags = flags;
    size_t rec_max = 0;
    for (int k = 0; k < n_codes; k++) {
        cusat_fec_ctx* c = codes[k];
        if (!c) {
            delete sc;
            set_last_error("null context %d", k);
            return CUSAT_E_INVALID;
        }
        int cmax = 0;
        bool exact = false;
        int cmin = 1 << 30, cmx = 0;
        for (int v : c->geo.lay_cnt) {
            cmin = std::min(cmin, v);
            cmx = std::max(cmx, v);
        }
        // the stream runs every code with a two-lane variant (a one-lane
        // schedule leaves lane 1 idle), so the variant may differ from the
        // one the context's own launches use
        if (!ldpc_select_variant(cmin, cmx, 2, &cmax, &exact) || variant_id(cmax, exact) < 0 ||
            ((c->bch.degree - 8) >> 6) != (c->bch.degree + 63) / 64 - 1) {
            delete sc;
            set_last_error("code %d has no stream decoder variant", k);
            return CUSAT_E_INVALID;
        }
        sc->codes.push_back(c);
        sc->variant.push_back(variant_id(cmax, exact));
        const int vs = split_variant_id(cmax, exact);
        sc->variant_split.push_back(vs);
        sc->pair.push_back(0);
        rec_max = std::max(rec_max, ldpc_msg_bytes(c->geo.q, cmax));
        const int ng = (int)c->geo.groups.size();
        const int lds = ldpc_lds_bytes(c->geo.K, c->geo.R, c->geo.n_slots, c->kern_steps, ng);
        int lds_w = lds;
        if (vs >= 0) {
            rec_max = std::max(rec_max, ldpc_msg_bytes_split(c->geo.q, cmax));
            lds_w = ldpc_lds_bytes_split(c->geo.K, c->geo.R, c->geo.n_slots, c->kern_steps, ng, 2);
        }
        // + the BCH byte table (up to 256 x 4 words) staged in the LDS
        sc->lds[0] = std::max(sc->lds[0], std::max(lds, 256 * 4 * 8));
        sc->lds[1] = std::max(sc->lds[1], std::max(lds_w, 256 * 4 * 8));
    }
    // Two frames per workgroup for the codes with a one-lane schedule, when
    // the 12-wave kernel holds one workgroup per CU anyway (its LDS is more
    // than half a CU's) and two such frames fit in that LDS: lane 1's six
    // waves would otherwise idle through those frames.  The message
    // workspace then holds two frames' records.
    if (!(flags & CUSAT_STREAM_NO_PAIRS) && sc->lds[1] > ldpc_dev::SPLIT_LDS_STATIC / 2)
        for (int k = 0; k < n_codes; k++) {
            const cusat_fec_ctx* c = sc->codes[k];
            if (c->geo.lanes != 1 || !pair_variant(sc->variant_split[k])) continue;
            const int lp = ldpc_lds_bytes_pair(c->geo.K, c->geo.R, c->geo.n_slots, c->kern_steps,
                                               (int)c->geo.groups.size());
            if (lp > sc->lds[1]) continue;
            int cmax = 0;
            bool exact = false;
            int cmin = 1 << 30, cmx = 0;
            for (int v : c->geo.lay_cnt) {
                cmin = std::min(cmin, v);
                cmx = std::max(cmx, v);
            }
            ldpc_select_variant(cmin, cmx, 2, &cmax, &exact);  // as above: it succeeded there
            sc->pair[k] = 1;
            rec_max = std::max(rec_max, 2 * ldpc_msg_bytes_split(c->geo.q, cmax));
        }
    for (int w = 0; w < 2; w++) sc->lds[w] += CTL_BYTES;  // control area
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete sc;
        set_last_error("no device properties");
        return CUSAT_E_HIP;
    }
    sc->n_cu = prop.multiProcessorCount;
    hipError_t e = fit_kernel<false>(sc->lds[0], &sc->max_wg[0]);
    if (e == hipSuccess) e = fit_kernel<true>(sc->lds[1], &sc->max_wg[1]);
    if (e != hipSuccess || sc->max_wg[0] < 1) {
        delete sc;
        set_last_error("stream kernel does not fit (%d bytes LDS): %s", sc->lds[0], hipGetErrorString(e));
        return CUSAT_E_HIP;
    }
    sc->ws_v4 = rec_max / 16;
    // one message workspace per workgroup the launch can have (the grid is
    // capped at max_frames)
    const size_t n_wg = std::min((size_t)sc->n_cu * std::max(sc->max_wg[0], sc->max_wg[1]),
                                 (size_t)std::max(1, max_frames));
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    sc->off_jobs = 256 + al(sizeof(cusat_stream_frame) * max_frames);
    sc->off_order = sc->off_jobs + al(sizeof(int2) * max_frames);
    sc->off_rx = sc->off_order + al(sizeof(int) * max_frames);
    sc->max_events = 2 * max_frames + n_codes;
    sc->off_dl = sc->off_rx + 256 + al(sizeof(rx_rec) * max_frames);
    sc->off_ev = sc->off_dl + al(sizeof(int2) * max_frames * DEMAP_MAX_PARTS);
    sc->plan_bytes = sc->off_ev + al(sizeof(rx_event) * sc->max_events);
    const size_t bytes = sc->plan_bytes + al(sizeof(code_desc) * n_codes) + al(sizeof(rx_code) * n_codes);
    sc->stage_bytes = sc->plan_bytes + sizeof(code_desc) * n_codes;
    for (const cusat_fec_ctx* c : sc->codes) {
        sc->llr_max = std::max<int64_t>(sc->llr_max, c->geo.N);
        sc->hard_max = std::max<int64_t>(sc->hard_max, c->geo.K / 8);
    }
    sc->code_gen.assign(n_codes, 0);
    // every device and pinned buffer of the context is allocated here, so
    // that no later call allocates (the async plan calls are capturable and
    // cheap); destroy frees whatever part was allocated
    char* base = nullptr;
    e = hipMalloc(&base, bytes);
    if (e == hipSuccess) sc->d_counters = reinterpret_cast<unsigned*>(base);
    if (e == hipSuccess) e = hipMalloc(&sc->d_ws, n_wg * rec_max);
    if (e == hipSuccess) e = hipMalloc(&sc->d_llr_work, (size_t)max_frames * sc->llr_max * sizeof(int16_t));
    if (e == hipSuccess) e = hipMalloc(&sc->d_hard_work, (size_t)max_frames * sc->hard_max);
    if (e == hipSuccess) e = hipMalloc(&sc->d_snap, n_codes * sizeof(rx_scale));
    if (e == hipSuccess) e = hipMalloc(&sc->d_est, sc->max_events * sizeof(float));
    for (int k = 0; k < cusat_stream_ctx::N_STAGE && e == hipSuccess; k++) {
        e = hipHostMalloc((void**)&sc->stage[k], sc->stage_bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sc->stage_done[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        set_last_error("stream buffers: %s", hipGetErrorString(e));
        cusat_stream_destroy(sc);
        return CUSAT_E_NOMEM;
    }
    sc->d_frames = reinterpret_cast<cusat_stream_frame*>(base + 256);
    sc->d_jobs = reinterpret_cast<int2*>(base + sc->off_jobs);
    sc->d_order = reinterpret_cast<int*>(base + sc->off_order);
    sc->d_codes = reinterpret_cast<code_desc*>(base + sc->plan_bytes);
    sc->d_rx_codes = reinterpret_cast<rx_code*>(base + sc->plan_bytes + al(sizeof(code_desc) * n_codes));
    if (hipMemset(sc->d_counters, 0, 256) != hipSuccess ||
        hipMemset(sc->d_snap, 0, n_codes * sizeof(rx_scale)) != hipSuccess) {
        cusat_stream_destroy(sc);
        set_last_error("counter init failed");
        return CUSAT_E_HIP;
    }
    const int urc = upload_codes(sc, true);
    if (urc != CUSAT_OK || hipStreamSynchronize(nullptr) != hipSuccess) {
        cusat_stream_destroy(sc);
        if (urc == CUSAT_OK) set_last_error("stream set-up: %s", hipGetErrorString(hipGetLastError()));
        return urc != CUSAT_OK ? urc : CUSAT_E_HIP;
    }
    // the 12-wave kernel needs a split variant of every code
    if (std::any_of(sc->variant_split.begin(), sc->variant_split.end(), [](int v) { return v < 0; }))
        sc->max_wg[1] = 0;
    sc->tuning.split_edges = sc->max_wg[1] >= 1 ? 1 : 0;
    sc->tuning.wg_per_cu = sc->max_wg[sc->tuning.split_edges];
    sc->tuning.frames_per_job = 1;
    sc->tuning.order = CUSAT_STREAM_ORDER_COST;
    *out = sc;
    return CUSAT_OK;
}

extern "C" int cusat_stream_destroy(cusat_stream_ctx* sc)
{
    if (!sc) return CUSAT_OK;
    for (int k = 0; k < cusat_stream_ctx::N_STAGE; k++) {
        if (sc->stage_done[k]) {
            hipEventSynchronize(sc->stage_done[k]);
            hipEventDestroy(sc->stage_done[k]);
        }
        if (sc->stage[k]) hipHostFree(sc->stage[k]);
    }
    if (sc->d_counters) hipFree(sc->d_counters);
    if (sc->d_ws) hipFree(sc->d_ws);
    if (sc->d_llr_work) hipFree(sc->d_llr_work);
    if (sc->d_snap) hipFree(sc->d_snap);
    if (sc->d_est) hipFree(sc->d_est);
    if (sc->d_hard_work) hipFree(sc->d_hard_work);
    delete sc;
    return CUSAT_OK;
}

extern "C" int cusat_stream_plan(cusat_stream_ctx* sc, const cusat_stream_frame* frames, int n_frames,
                                 const cusat_stream_tuning* tuning)
{
    if (!sc || (n_frames > 0 && !frames) || n_frames < 0 || n_frames > sc->max_frames) {
        set_last_error("bad argument (n_frames %d, max %d)", n_frames, sc ? sc->max_frames : 0);
        return CUSAT_E_INVALID;
    }
    cusat_stream_tuning tu = tuning ? *tuning : sc->tuning;
    const int wmax = (tu.split_edges == 0 || tu.split_edges == 1) ? sc->max_wg[tu.split_edges] : 0;
    if (tu.wg_per_cu < 1 || tu.wg_per_cu > wmax || tu.frames_per_job < 1 || tu.frames_per_job > 64 ||
        (tu.order != CUSAT_STREAM_ORDER_GIVEN && tu.order != CUSAT_STREAM_ORDER_COST)) {
        set_last_error("tuning out of range (split_edges 0 / 1, wg_per_cu 1..%d / 1..%d, frames_per_job 1..64)",
                       sc->max_wg[0], sc->max_wg[1]);
        return CUSAT_E_INVALID;
    }
    for (int k = 0; k < n_frames; k++)
        if (frames[k].code < 0 || frames[k].code >= (int)sc->codes.size() || frames[k].llr_offset < 0 ||
            frames[k].hard_offset < 0) {
            set_last_error("frame %d: code %d / offsets out of range", k, frames[k].code);
            return CUSAT_E_INVALID;
        }
    sc->frames.assign(frames, frames + n_frames);
    sc->n_frames = n_frames;
    int rc = upload_codes(sc, false);
    if (rc != CUSAT_OK) return rc;
    sc->rx_planned = false;
    return build_plan(sc, tu);
}

extern "C" int cusat_stream_plan_async(cusat_stream_ctx* sc, const cusat_stream_frame* frames, int n_frames,
                                       void* stream)
{
    if (!sc || (n_frames > 0 && !frames) || n_frames < 0 || n_frames > sc->max_frames) {
        set_last_error("bad argument (n_frames %d, max %d)", n_frames, sc ? sc->max_frames : 0);
        return CUSAT_E_INVALID;
    }
    for (int k = 0; k < n_frames; k++)
        if (frames[k].code < 0 || frames[k].code >= (int)sc->codes.size() || frames[k].llr_offset < 0 ||
            frames[k].hard_offset < 0) {
            set_last_error("frame %d: code %d / offsets out of range", k, frames[k].code);
            return CUSAT_E_INVALID;
        }
    hipStream_t st = (hipStream_t)stream;
    unsigned char* img = nullptr;
    int rc = next_stage_image(sc, st, &img);
    if (rc != CUSAT_OK) return rc;
    sc->frames.assign(frames, frames + n_frames);
    sc->n_frames = n_frames;
    const int nj = stage_plan(sc, sc->tuning, img);
    HIP_TRY(copy_plan(sc, img, n_frames, nj, st, true));
    rc = release_stage_image(sc, st);
    if (rc != CUSAT_OK) return rc;
    sc->n_jobs = nj;
    sc->rx_planned = false;
    return CUSAT_OK;
}

extern "C" int cusat_stream_get_tuning(const cusat_stream_ctx* sc, cusat_stream_tuning* out, int* max_wg_per_cu)
{
    if (!sc || !out) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    *out = sc->tuning;
    if (max_wg_per_cu) *max_wg_per_cu = sc->max_wg[sc->tuning.split_edges];
    return CUSAT_OK;
}

namespace {

// The decode launch: persistent LDPC (learned pass and plain decoder on a
// syndrome failure), the BCH checks in rounds, the plain redo of learned
// codewords the check rejects.
hipError_t launch_stream(cusat_stream_ctx* sc, const int16_t* llr, uint8_t* hard, int32_t* trials, int32_t* valid,
                         int32_t* fallback, void* stream)
{
    // never more workgroups than a plan can have jobs: a context for small
    // batches (an ACM receiver's micro-batches) leaves the other CUs to the
    // next batch's launches instead of parking idle workgroups on them
    const int grid = std::min(sc->n_cu * sc->tuning.wg_per_cu, std::max(1, sc->max_frames));
    const int w = sc->tuning.split_edges;
    const int learned = (sc->flags & CUSAT_STREAM_LEARNED) ? 1 : 0;
    hipStream_t st = (hipStream_t)stream;
    if (w && sc->lds[1] > ldpc_dev::SPLIT_LDS_STATIC / 2)  // as fit_kernel chose
        hipLaunchKernelGGL((ldpc_stream_kernel<true, true>), dim3(grid), dim3(stream_threads<true>()), 0, st,
                           sc->d_codes, sc->d_frames, sc->d_jobs, sc->d_order, sc->d_counters, llr, hard, trials,
                           valid, fallback, sc->d_ws, sc->ws_v4, learned, sc->lds[1] - CTL_BYTES);
    else if (w)
        hipLaunchKernelGGL(ldpc_stream_kernel<true>, dim3(grid), dim3(stream_threads<true>()), sc->lds[1], st,
                           sc->d_codes, sc->d_frames, sc->d_jobs, sc->d_order, sc->d_counters, llr, hard, trials,
                           valid, fallback, sc->d_ws, sc->ws_v4, learned, sc->lds[1] - CTL_BYTES);
    else
        hipLaunchKernelGGL(ldpc_stream_kernel<false>, dim3(grid), dim3(stream_threads<false>()), sc->lds[0], st,
                           sc->d_codes, sc->d_frames, sc->d_jobs, sc->d_order, sc->d_counters, llr, hard, trials,
                           valid, fallback, sc->d_ws, sc->ws_v4, learned, sc->lds[0] - CTL_BYTES);
    return hipGetLastError();
}

}  // namespace

// Launches even for an empty plan (every workgroup leaves at once), so a
// captured launch exists whatever plan it later replays.
extern "C" int cusat_stream_decode(cusat_stream_ctx* sc, const int16_t* llr, uint8_t* hard, int32_t* trials,
                                   int32_t* valid, int32_t* fallback, void* stream)
{
    if (!sc || !llr || !hard || !trials || !valid) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    HIP_TRY(launch_stream(sc, llr, hard, trials, valid, fallback, stream));
    return CUSAT_OK;
}

extern "C" int cusat_stream_rx_plan_async(cusat_stream_ctx* sc, const cusat_stream_rx_frame* frames, int n_frames,
                                          const cusat_stream_rx_event* events, int n_events, void* stream)
{
    if (!sc || (n_frames > 0 && !frames) || n_frames < 0 || n_frames > sc->max_frames || (n_events > 0 && !events) ||
        n_events < 0 || n_events > sc->max_events) {
        set_last_error("bad argument (n_frames %d, max %d; n_events %d, max %d)", n_frames, sc ? sc->max_frames : 0,
                       n_events, sc ? sc->max_events : 0);
        return CUSAT_E_INVALID;
    }
    for (int k = 0; k < n_frames; k++) {
        const bool state = frames[k].flags & CUSAT_RX_SCALE_STATE;
        if (frames[k].code < 0 || frames[k].code >= (int)sc->codes.size() || frames[k].sym_offset < 0 ||
            frames[k].out_offset < 0 || (frames[k].flags & ~CUSAT_RX_SCALE_STATE) ||
            (!state && !(frames[k].noise_var > 0.0f))) {
            set_last_error("rx frame %d: code %d / offsets / flags / noise_var out of range", k, frames[k].code);
            return CUSAT_E_INVALID;
        }
    }
    for (int k = 0; k < n_events; k++) {
        const cusat_stream_rx_event& e = events[k];
        if (e.code < 0 || e.code >= (int)sc->codes.size() || e.frame < -1 || e.frame >= n_frames ||
            (e.frame < 0 && !(e.snr_lin > 0.0f && e.snr_lin < 1e30f))) {
            set_last_error("rx event %d: code %d / frame %d / snr out of range", k, e.code, e.frame);
            return CUSAT_E_INVALID;
        }
    }
    hipStream_t st = (hipStream_t)stream;
    unsigned char* img = nullptr;
    int rc = next_stage_image(sc, st, &img);
    if (rc != CUSAT_OK) return rc;
    int* hdr = reinterpret_cast<int*>(img + sc->off_rx);
    rx_rec* rec = reinterpret_cast<rx_rec*>(img + sc->off_rx + 256);
    // demap order: the frames of the sorted path, then the others, each part
    // costliest first (arrival order among equals); blocks of the others
    // numbered in that order
    int2* dl = reinterpret_cast<int2*>(img + sc->off_dl);
    int n_sorted = 0, n_dl = 0, blk = 0;
    {
        std::vector<int>& ord = sc->demap_order;
        ord.resize(n_frames);
        std::iota(ord.begin(), ord.end(), 0);
        auto sorted_path = [&](int f) {
            const cusat_fec_ctx* c = sc->codes[frames[f].code];
            const bool table = !(c->modulation == CUSAT_MOD_QPSK || c->modulation == CUSAT_MOD_8PSK);
            return demap_sorted(table, c->con.n_points, c->dlen);
        };
        auto cost = [&](int f) {
            const cusat_fec_ctx* c = sc->codes[frames[f].code];
            const bool table = !(c->modulation == CUSAT_MOD_QPSK || c->modulation == CUSAT_MOD_8PSK);
            return (int64_t)c->dlen * (table ? c->con.n_points * c->con.bits_per_sym : 1);
        };
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
            const bool sa = sorted_path(a), sb = sorted_path(b);
            if (sa != sb) return sa;
            return cost(a) > cost(b);
        });
        for (int k = 0; k < n_frames; k++) {
            const int f = ord[k];
            const cusat_fec_ctx* c = sc->codes[frames[f].code];
            if (sorted_path(f)) {
                const int parts = demap_parts(c->dlen, c->con.n_points);
                for (int p = 0; p < parts; p++) dl[n_dl++] = make_int2(f, p | (parts << 8));
                n_sorted += parts;
            } else {
                const bool table = !(c->modulation == CUSAT_MOD_QPSK || c->modulation == CUSAT_MOD_8PSK);
                const int span = demap_span(table, c->con.bits_per_sym);
                dl[n_dl++] = make_int2(f, blk);
                blk += (c->dlen + span - 1) / span;
            }
        }
    }
    // the decoder's frames: LLRs and hard decisions packed in the work buffers
    sc->frames.resize(n_frames);
    int64_t lo = 0, ho = 0;
    for (int f = 0; f < n_frames; f++) {
        const cusat_fec_ctx* c = sc->codes[frames[f].code];
        sc->frames[f] = cusat_stream_frame{frames[f].code, 0, lo, ho};
        rx_rec& r = rec[f];
        memset(&r, 0, sizeof r);
        r.code = frames[f].code;
        r.blk0 = -1;
        r.sym_off = frames[f].sym_offset;
        r.llr_off = lo;
        r.hard_off = ho;
        r.out_off = frames[f].out_offset;
        if (frames[f].flags & CUSAT_RX_SCALE_STATE)
            r.flags = RX_SCALE_STATE;
        else
            cusat_fec_demap_scales(c, frames[f].noise_var, frames[f].llr_gain, &r.qpsk_scale, &r.psk8_scale,
                                   &r.gain);
        lo += c->geo.N;
        ho += c->geo.K / 8;
    }
    for (int k = n_sorted; k < n_dl; k++) rec[dl[k].x].blk0 = dl[k].y;
    memset(hdr, 0, 256);
    hdr[0] = n_frames;
    hdr[1] = blk;
    hdr[2] = n_events;
    hdr[3] = n_sorted;
    hdr[4] = n_dl;
    rx_event* ev = reinterpret_cast<rx_event*>(img + sc->off_ev);
    for (int e = 0; e < n_events; e++)
        ev[e] = rx_event{events[e].code, events[e].frame, events[e].snr_lin, events[e].snr_db};
    sc->n_frames = n_frames;
    const int nj = stage_plan(sc, sc->tuning, img);
    unsigned char* d = reinterpret_cast<unsigned char*>(sc->d_counters);
    const size_t span = sc->off_ev + n_events * sizeof(rx_event);
    if (span <= (256u << 10)) {
        // one copy of the whole image up to the last event (the unused
        // tails of its arrays ride along; no launch reads past a count)
        HIP_TRY(hipMemcpyAsync(d, img, span, hipMemcpyHostToDevice, st));
    } else {
        HIP_TRY(copy_plan(sc, img, n_frames, nj, st, true));
        HIP_TRY(hipMemcpyAsync(d + sc->off_rx, img + sc->off_rx, 256 + n_frames * sizeof(rx_rec),
                               hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d + sc->off_dl, img + sc->off_dl, n_dl * sizeof(int2), hipMemcpyHostToDevice, st));
        if (n_events)
            HIP_TRY(hipMemcpyAsync(d + sc->off_ev, img + sc->off_ev, n_events * sizeof(rx_event),
                                   hipMemcpyHostToDevice, st));
    }
    rc = release_stage_image(sc, st);
    if (rc != CUSAT_OK) return rc;
    sc->n_jobs = nj;
    sc->rx_planned = true;
    return CUSAT_OK;
}

namespace {

int need_rx_plan(const cusat_stream_ctx* sc, const char* what)
{
    if (!sc->rx_planned) {
        set_last_error("%s needs an rx plan (cusat_stream_rx_plan_async)", what);
        return CUSAT_E_INVALID;
    }
    return CUSAT_OK;
}

const int* rx_header(const cusat_stream_ctx* sc)
{
    return reinterpret_cast<const int*>(reinterpret_cast<const unsigned char*>(sc->d_counters) + sc->off_rx);
}

}  // namespace

extern "C" int cusat_stream_rx_demap(cusat_stream_ctx* sc, const float* sym, void* stream)
{
    if (!sc || !sym) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    const int rc = need_rx_plan(sc, "cusat_stream_rx_demap");
    if (rc != CUSAT_OK) return rc;
    int* hdr = const_cast<int*>(rx_header(sc));
    const rx_rec* rec = reinterpret_cast<const rx_rec*>(hdr + 64);
    const int2* dlist =
        reinterpret_cast<const int2*>(reinterpret_cast<const unsigned char*>(sc->d_counters) + sc->off_dl);
    HIP_TRY(launch_demap_mixed(sc->d_rx_codes, rec, dlist, sc->d_snap, hdr, sym, sc->d_llr_work, sc->n_cu,
                               (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_stream_rx_decode(cusat_stream_ctx* sc, int32_t* trials, int32_t* valid, int32_t* fallback,
                                      void* stream)
{
    if (!sc || !trials || !valid) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    const int rc = need_rx_plan(sc, "cusat_stream_rx_decode");
    if (rc != CUSAT_OK) return rc;
    HIP_TRY(launch_stream(sc, sc->d_llr_work, sc->d_hard_work, trials, valid, fallback, stream));
    return CUSAT_OK;
}

extern "C" int cusat_stream_rx_bb(cusat_stream_ctx* sc, uint8_t* out, const int32_t* trials, const int32_t* valid,
                                  int32_t* dfl, void* stream)
{
    if (!sc || !out || !trials || !valid || !dfl) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    const int rc = need_rx_plan(sc, "cusat_stream_rx_bb");
    if (rc != CUSAT_OK) return rc;
    const int* hdr = rx_header(sc);
    const rx_rec* rec = reinterpret_cast<const rx_rec*>(hdr + 64);
    HIP_TRY(launch_bb_extract_mixed(sc->d_rx_codes, rec, hdr, sc->d_hard_work, trials, valid, out, dfl, sc->n_cu * 2,
                                    (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_stream_rx(cusat_stream_ctx* sc, const float* sym, uint8_t* out, int32_t* trials,
                               int32_t* valid, int32_t* dfl, void* stream)
{
    if (!sc || !sym || !out || !trials || !valid || !dfl) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    int rc = cusat_stream_rx_demap(sc, sym, stream);
    if (rc == CUSAT_OK) rc = cusat_stream_rx_decode(sc, trials, valid, nullptr, stream);
    if (rc == CUSAT_OK) rc = cusat_stream_rx_bb(sc, out, trials, valid, dfl, stream);
    return rc;
}

extern "C" int cusat_stream_rx_work(const cusat_stream_ctx* sc, const int16_t** llr, const uint8_t** hard)
{
    if (!sc) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    if (llr) *llr = sc->d_llr_work;
    if (hard) *hard = sc->d_hard_work;
    return CUSAT_OK;
}

extern "C" int cusat_stream_rx_dd_estimate(cusat_stream_ctx* sc, const float* sym, void* stream)
{
    if (!sc || !sym) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    const int rc = need_rx_plan(sc, "cusat_stream_rx_dd_estimate");
    if (rc != CUSAT_OK) return rc;
    const unsigned char* base = reinterpret_cast<const unsigned char*>(sc->d_counters);
    const int* hdr = reinterpret_cast<const int*>(base + sc->off_rx);
    const rx_rec* rec = reinterpret_cast<const rx_rec*>(hdr + 64);
    const rx_event* ev = reinterpret_cast<const rx_event*>(base + sc->off_ev);
    HIP_TRY(launch_dd_estimate(sc->d_rx_codes, rec, ev, hdr, sym, sc->d_est, 64, (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_stream_rx_dd_resolve(cusat_stream_ctx* sc, float* state, float* state_out, void* stream)
{
    if (!sc || !state) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    const int rc = need_rx_plan(sc, "cusat_stream_rx_dd_resolve");
    if (rc != CUSAT_OK) return rc;
    const unsigned char* base = reinterpret_cast<const unsigned char*>(sc->d_counters);
    const int* hdr = reinterpret_cast<const int*>(base + sc->off_rx);
    const rx_event* ev = reinterpret_cast<const rx_event*>(base + sc->off_ev);
    HIP_TRY(launch_dd_resolve(sc->d_rx_codes, ev, sc->d_est, hdr, (int)sc->codes.size(),
                              reinterpret_cast<float4*>(state), sc->d_snap, reinterpret_cast<float4*>(state_out),
                              (hipStream_t)stream));
    return CUSAT_OK;
}

extern "C" int cusat_stream_rx_dd(cusat_stream_ctx* sc, const float* sym, float* state, float* state_out,
                                  void* stream)
{
    if (!sc || !sym || !state) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    const int rc = cusat_stream_rx_dd_estimate(sc, sym, stream);
    return rc != CUSAT_OK ? rc : cusat_stream_rx_dd_resolve(sc, state, state_out, stream);
}

extern "C" int cusat_stream_autotune(cusat_stream_ctx* sc, const int16_t* llr, uint8_t* hard, int32_t* trials,
                                     int32_t* valid, int32_t* fallback, int reps, void* stream,
                                     cusat_stream_tuning* chosen, float* best_ms)
{
    if (!sc || reps < 1) { set_last_error("bad argument"); return CUSAT_E_INVALID; }
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    cusat_stream_tuning best = sc->tuning;
    float best_t = 1e30f;
    int rc = CUSAT_OK;
    // the 12-wave kernel has a form for every code and has been the faster
    // one on every measured mix: the 6-wave kernel is timed only as the
    // fallback when the 12-wave one does not fit
    for (int split = sc->max_wg[1] >= 1 ? 1 : 0; split < 2 && rc == CUSAT_OK; split++)
    for (int order = 0; order < 2 && rc == CUSAT_OK; order++)
        for (int fpj : {1, 2, 4})
            for (int wg = 1; wg <= sc->max_wg[split] && rc == CUSAT_OK; wg++) {
                cusat_stream_tuning tu{wg, fpj, order, split};
                rc = build_plan(sc, tu);
                if (rc != CUSAT_OK) break;
                rc = cusat_stream_decode(sc, llr, hard, trials, valid, fallback, stream);  // warm
                if (rc != CUSAT_OK) break;
                hipEventRecord(e0, st);
                for (int r = 0; r < reps && rc == CUSAT_OK; r++)
                    rc = cusat_stream_decode(sc, llr, hard, trials, valid, fallback, stream);
                hipEventRecord(e1, st);
                if (hipEventSynchronize(e1) != hipSuccess) {
                    set_last_error("autotune: event sync failed");
                    rc = CUSAT_E_HIP;
                    break;
                }
                float ms = 0;
                hipEventElapsedTime(&ms, e0, e1);
                ms /= reps;
                if (ms < best_t) {
                    best_t = ms;
                    best = tu;
                }
            }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != CUSAT_OK) return rc;
    rc = build_plan(sc, best);
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return rc;
}
