// Layered offset/normalised min-sum LDPC decode kernel for the DVB-S2/S2X
// IRA codes on gfx950 (device code; instantiated in ldpc_inst_*.hip).
//
// Semantics follow lib/cusat_dsp/cuda/ldpc_decode_kernel.cu (wineslab/cusat):
// int16 APP with saturation, int8 check-to-variable messages clamped to
// +-127, offset beta (plain = 1) inside the check node (:242-275), optional
// scale n/64 (:75-90), data writes that re-apply the delta to the APP
// (:793-801), parity writes from the read-phase value (:803-833), syndrome
// early exit before every iteration (:689-749), result = T - t or -1.
//
// Work split (one workgroup = one frame):
//  * A lane (one of the two layers of a schedule step) is 3 wave64s; thread
//    t < 180 of a lane owns checks j = t and j = t + 180 of the lane's layer
//    and evaluates both in the two halves of packed int16 registers
//    (v_pk_sub_i16 clamp = qsat16, v_pk_max_i16(x, sat(0 - x)) = qabs,
//    v_pk_sub_u16 clamp = the offset qsubu, v_pk_min/max for min1/min2).
//    Lanes are whole waves, so branches on the layer id are wave-uniform and
//    the schedule / circulant tables are read with scalar loads.
//
// LDS (int16 APP, ~130 KB for a normal frame):
//  * VN group r (VNs 360 r .. 360 r + 359) is 180 dwords; dword x holds VN
//    360 r + x (low half) and 360 r + x + 180 (high half).  Check j of layer
//    i reads VN 360 r + (j - s) mod 360 for each circulant (r, s) of the
//    layer, so the pair (t, t + 180) reads ONE dword, (t - s) mod 180, with
//    its halves swapped when (t - s) mod 360 >= 180.  A wave reads 64
//    consecutive dwords (one wrap at most): no bank conflicts.
//  * Parity VNs of layer i are 180 dwords: dword x = (P(i, x), P(i, x+180)),
//    P(i, j) = channel bit K + q j + i.  No swap is ever needed.
//  * Delta slots [n_slots][180] dwords in the same interleaved layout.
//
// Messages (HBM): one record per (frame, layer, thread): both checks' int8
// messages, edge-interleaved (A0 B0 A1 B1 ...), data edges first, the two
// parity edges at slots CMAX and CMAX + 1, padded to 16-byte multiples and
// moved with dwordx4; the record of the next step is prefetched while the
// current step computes.
//
// Collisions: a step touches each VN once except on colliding circulants (a
// weight-2/3 circulant inside a layer, or a VN group shared by the two
// layers of a step).  Those edges park their delta (y - old) in an LDS slot;
// after the step's barrier one thread per dword adds the deltas in
// (lane, circulant) order with int16 saturation per add.  All reads of a
// step therefore see the step-start APP, and a VN touched once gets exactly
// the reference's update.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace cusat {
namespace ldpc_dev {

constexpr int LANE_THREADS = 192;  // 3 x wave64 per lane
constexpr int HALF = 180;
constexpr uint32_t SEL_ID = 0x03020100u;   // v_perm: keep halves
constexpr uint32_t SEL_ROT = 0x01000302u;  // v_perm: swap halves

typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s2 as_s2(uint32_t w) { return __builtin_bit_cast(s2, w); }
__device__ __forceinline__ uint32_t bits(s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u2 as_u2(s2 v) { return __builtin_bit_cast(u2, v); }
__device__ __forceinline__ s2 as_s2u(u2 v) { return __builtin_bit_cast(s2, v); }
__device__ __forceinline__ s2 splat(int a)
{
    s2 v;
    v.x = (short)a;
    v.y = (short)a;
    return v;
}
__device__ __forceinline__ uint32_t perm(uint32_t w, uint32_t sel)
{
    return __builtin_amdgcn_perm(w, w, sel);
}

// int8 message pair of edge e: bytes (A_e, B_e) at byte 2 (e & 1) of record
// word e / 2 -> packed sign-extended int16.  e folds to a constant after
// unrolling.
__device__ __forceinline__ s2 unpack_msg(const uint32_t* w, int e)
{
    const uint32_t sel = (e & 1) ? 0x030c020cu : 0x010c000cu;  // [0, A, 0, B]
    return as_s2(__builtin_amdgcn_perm(w[e >> 1], w[e >> 1], sel)) >> (short)8;
}

// record word of edges (2w, 2w + 1)
__device__ __forceinline__ uint32_t pack_msg(s2 y_even, s2 y_odd)
{
    return __builtin_amdgcn_perm(bits(y_odd), bits(y_even), 0x06040200u);
}

// Scaled and clamped check-node magnitude min((m n + 32) >> 6, 127) of both
// halves (m in 0..32767; the odd-symmetric fixed-point scale of
// ldpc_decode_kernel.cu:75-90 with n in 64ths).  alpha_word = n |
// ceil(8096 / n) << 16 (n in 1..1023, packed by the host): every m >= that
// bound gives m n >= 8096, i.e. at least 127; below it m n + 32 < 8128, so
// after clamping m to the bound the product is exact in 16-bit lanes and
// four packed operations replace two 32-bit multiplies per value.
__device__ __forceinline__ u2 nms_scale_clamp(u2 m, int alpha_word)
{
    const u2 n = (u2)(unsigned short)(alpha_word & 0xffff);
    const u2 bound = (u2)(unsigned short)((unsigned)alpha_word >> 16);
    const u2 mc = __builtin_elementwise_min(m, bound);
    return __builtin_elementwise_min((u2)((mc * n + (u2)32) >> (u2)6), (u2)127);
}

// message record geometry for a kernel variant
template <int CMAX>
struct rec_t {
    static constexpr int NE = CMAX + 2;
    static constexpr int WORDS = (((NE + 1) / 2) + 3) / 4 * 4;
    static constexpr int V4 = WORDS / 4;
};

// LDS int16 index of VN 360 r + jj
__device__ __forceinline__ int vn_index(int v)
{
    const int r = v / CIRC, jj = v - r * CIRC;
    return r * CIRC + (jj < HALF ? 2 * jj : 2 * (jj - HALF) + 1);
}

// Circulant record (host packs it, codes.cc), 4 dwords read with one scalar
// load: sp = s mod 180; app_off / slot_off = LDS byte offset of the pair
// dword for thread 0 without wrap (4 (base - sp)), slot_off < 0 when the
// circulant writes the APP directly; sel = v_perm selector of the pair when
// thread tt does not wrap (tt >= sp); a wrap flips it (SEL_ID <-> SEL_ROT).
struct circ_rec {
    int sp, app_off, slot_off;
    uint32_t sel;
};
constexpr uint32_t SEL_FLIP = SEL_ID ^ SEL_ROT;

// byte offset (to add to the circulant's offsets) and half-swap selector of
// thread tt's VN pair
__device__ __forceinline__ void pair_addr(const circ_rec& c, int tt4, int tt, int& off, uint32_t& sel)
{
    const bool wrap = tt < c.sp;
    off = tt4 + (wrap ? 4 * HALF : 0);
    sel = c.sel ^ (wrap ? SEL_FLIP : 0u);
}

__device__ __forceinline__ uint32_t lds_ld(const unsigned char* base, int byte_off)
{
    return *reinterpret_cast<const uint32_t*>(base + byte_off);
}
__device__ __forceinline__ void lds_st(unsigned char* base, int byte_off, uint32_t v)
{
    *reinterpret_cast<uint32_t*>(base + byte_off) = v;
}

struct kparams {
    int N, K, R, q, n_steps, max_trials, n_param_iters, n_slots, n_groups;
};

// Workgroup-wide OR of a per-thread flag with one barrier (replaces
// __syncthreads_or, whose reduction scratch is static LDS).  flags: 3 ints of
// LDS, zero before the first call; k: a call counter, the same in every
// thread.  Call k writes slot k % 3 and clears slot (k + 1) % 3 -- last read
// in call k - 2, which every thread finished before the barrier of call k - 1.
__device__ __forceinline__ int wg_any(int v, int* flags, int& k)
{
    const int slot = k % 3;
    if (v) flags[slot] = 1;
    if (threadIdx.x == 0) flags[slot == 2 ? 0 : slot + 1] = 0;
    __syncthreads();
    const int r = flags[slot];
    k++;
    return r;
}

// wg_any over two groups of threads (g = 0, 1; flags [3][2]): bit g of the
// result is set if any thread of group g passed v != 0.  The same value in
// every thread of the workgroup.
__device__ __forceinline__ int wg_any2(int v, int g, int* flags, int& k)
{
    const int slot = k % 3;
    if (v) flags[2 * slot + g] = 1;
    if (threadIdx.x == 0) {
        const int nx = slot == 2 ? 0 : slot + 1;
        flags[2 * nx] = 0;
        flags[2 * nx + 1] = 0;
    }
    __syncthreads();
    const int r = (flags[2 * slot] ? 1 : 0) | (flags[2 * slot + 1] ? 2 : 0);
    k++;
    return r;
}

// Collision group record (int32 x 8, host-packed): LDS dword index of the
// VN group's first pair dword, number of slots, the slots' dword indices.
constexpr int GROUP_WORDS = 8;

// Channel LLRs of one frame into the LDS APP layout with the int16-APP input
// clamp -32768 -> -32767 (ldpc_decode_kernel.cu:655-670): info VN v at
// app16[vn_index(v)], parity c = q j + i at dword i * 180 + j mod 180 (low
// half for j < 180).  Whole workgroup (NTHR threads), no barrier.
//
// With an 8-byte-aligned llr (every frame of a packed batch or stream), the
// loads are 8 bytes wide and issued eight at a time before their first use:
// the frame's LLRs then arrive in a few memory round trips instead of one
// round trip per element.  An info unit is 4 LLRs of each half of a VN
// group (x0 .. x0 + 3 and x0 + 180 .. x0 + 183), which are exactly 4 pair
// dwords: one 16-byte LDS write.  A parity unit is 4 consecutive LLRs c;
// (i, j) comes from one division per unit and an increment per element.
template <int NTHR>
__device__ __forceinline__ void load_frame_llr(uint32_t* app32, int16_t* par16, const int16_t* __restrict__ llr,
                                               int K, int R, int q, int t)
{
    int16_t* app16 = reinterpret_cast<int16_t*>(app32);
    if ((reinterpret_cast<uintptr_t>(llr) & 7) != 0) {
        for (int v = t; v < K; v += NTHR) app16[vn_index(v)] = (int16_t)max((int)llr[v], -32767);
        for (int c = t; c < R; c += NTHR) {
            const int i = c % q, j = c / q;
            par16[i * CIRC + (j < HALF ? 2 * j : 2 * (j - HALF) + 1)] = (int16_t)max((int)llr[K + c], -32767);
        }
        return;
    }
    const s2 lo = splat(-32767);
    const uint2* src = reinterpret_cast<const uint2*>(llr);  // 4 LLRs each
    constexpr int UI = 4;  // info units per batch (2 loads each)
    const int nu = K / 8;  // K is a multiple of 360: 45 units per VN group
    for (int u0 = t; u0 < nu; u0 += UI * NTHR) {
        uint2 a[UI], b[UI];
#pragma unroll
        for (int k = 0; k < UI; k++) {
            const int u = u0 + k * NTHR;
            const int r = u / 45, e = 90 * r + (u - 45 * r);  // uint2 index of LLR 360 r + x0
            a[k] = make_uint2(0, 0);
            b[k] = make_uint2(0, 0);
            if (u < nu) {
                a[k] = src[e];
                b[k] = src[e + 45];  // + 180 LLRs
            }
        }
#pragma unroll
        for (int k = 0; k < UI; k++) {
            const int u = u0 + k * NTHR;
            if (u < nu) {
                const int r = u / 45, x0 = 4 * (u - 45 * r);
                // pair dword x = (LLR 360 r + x, LLR 360 r + 180 + x)
                uint4 w;
                w.x = bits(__builtin_elementwise_max(as_s2(__builtin_amdgcn_perm(b[k].x, a[k].x, 0x05040100u)), lo));
                w.y = bits(__builtin_elementwise_max(as_s2(__builtin_amdgcn_perm(b[k].x, a[k].x, 0x07060302u)), lo));
                w.z = bits(__builtin_elementwise_max(as_s2(__builtin_amdgcn_perm(b[k].y, a[k].y, 0x05040100u)), lo));
                w.w = bits(__builtin_elementwise_max(as_s2(__builtin_amdgcn_perm(b[k].y, a[k].y, 0x07060302u)), lo));
                *reinterpret_cast<uint4*>(app32 + 180 * r + x0) = w;
            }
        }
    }
    constexpr int UP = 8;  // parity units per batch
    const int np = R / 4;
    const uint2* psrc = src + K / 4;
    const float inv_q = 1.0f / (float)q;
    for (int u0 = t; u0 < np; u0 += UP * NTHR) {
        uint2 a[UP];
#pragma unroll
        for (int k = 0; k < UP; k++) {
            const int u = u0 + k * NTHR;
            a[k] = u < np ? psrc[u] : make_uint2(0, 0);
        }
#pragma unroll
        for (int k = 0; k < UP; k++) {
            const int u = u0 + k * NTHR;
            if (u < np) {
                const int c0 = 4 * u;
                // j = c0 / q: the float quotient is within one of it (c0 < 2^17)
                int j = (int)((float)c0 * inv_q);
                if (j * q > c0) j--;
                if ((j + 1) * q <= c0) j++;
                int i = c0 - j * q;
                const uint32_t w[2] = {a[k].x, a[k].y};
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int v = (int)(int16_t)(w[e >> 1] >> (16 * (e & 1)));
                    par16[i * CIRC + (j < HALF ? 2 * j : 2 * (j - HALF) + 1)] = (int16_t)max(v, -32767);
                    if (++i == q) {
                        i = 0;
                        j++;
                    }
                }
            }
        }
    }
}

// Hard decisions of the K info VNs, MSB first (ldpc_decode_kernel.cu:468-493):
// byte B holds the APP signs of VNs 8B .. 8B + 7.  Within VN group r (45
// bytes, pair dwords 180 r ..): byte j < 22 is the low halves of dwords
// 8j .. 8j + 7, byte j > 22 the high halves of dwords 8j - 180 .. 8j - 173,
// byte 22 the low halves of dwords 176 .. 179 and the high halves of 0 .. 3.
// Whole workgroup, no barrier.
template <int NTHR>
__device__ __forceinline__ void store_hard(const uint32_t* app32, uint8_t* __restrict__ h, int K, int t)
{
    for (int b = t; b < (K >> 3); b += NTHR) {
        const int r = b / 45, j = b - 45 * r;
        const uint32_t* g = app32 + 180 * r;
        const int d0 = j < 22 ? 8 * j : (j > 22 ? 8 * j - 180 : 176);
        const uint4 w0 = *reinterpret_cast<const uint4*>(g + d0);
        uint4 w1 = *reinterpret_cast<const uint4*>(g + d0 + 4);  // in the LDS for j = 22 too, unused there
        if (j == 22) w1 = *reinterpret_cast<const uint4*>(g);
        const int sa = j <= 22 ? 15 : 31;  // sign bit of the first four VNs' half
        const int sb = j < 22 ? 15 : 31;   // and of the last four
        const uint32_t byte = ((w0.x >> sa) & 1u) << 7 | ((w0.y >> sa) & 1u) << 6 | ((w0.z >> sa) & 1u) << 5 |
                              ((w0.w >> sa) & 1u) << 4 | ((w1.x >> sb) & 1u) << 3 | ((w1.y >> sb) & 1u) << 2 |
                              ((w1.z >> sb) & 1u) << 1 | ((w1.w >> sb) & 1u);
        h[b] = (uint8_t)byte;
    }
}

// Iteration 0's old messages are zero (ldpc_decode_kernel.cu:495-581).  Its
// record prefetches read this never-written, zero-initialised array instead
// of the frame's records: the load stays unconditional (a pointer select,
// no branch and no extra wait), the records need no clearing pass and
// iteration 0 moves no record traffic.  Sized for the largest record of a
// layer (HALF pairs x 4 uint4).
static __device__ uint4 zero_records[HALF * 4];

// word 3 of a raw buffer resource: 32-bit data format, no swizzle / stride;
// loads beyond num_records bytes return 0
constexpr int BUF_RSRC_WORD3 = 0x00020000;

// Decode one frame with the whole workgroup (LANES x 192 threads; a
// workgroup of 384 threads runs a LANES = 1 code with its second lane idle,
// since every step of a one-lane schedule has lane 1 = -1).  llr: the frame's
// N LLRs; rec_base: its message records; h: its K/8 hard bytes.  Returns
// T - t on convergence at trial t, else -1.  Every thread of the workgroup
// must call it (it contains barriers); it starts and ends with the LDS free.
template <int CMAX, bool EXACT, int LANES>
__device__ __forceinline__ int decode_frame(unsigned char* lds_raw, const int16_t* __restrict__ llr,
                                            uint4* __restrict__ rec_base, uint8_t* __restrict__ h,
                                            const circ_rec* __restrict__ circ,
                                            const int* __restrict__ lay_cnt, const int* __restrict__ lay_off,
                                            const int* __restrict__ steps, const int* __restrict__ grp_off,
                                            const int* __restrict__ groups, const int* __restrict__ beta,
                                            const int* __restrict__ alpha, const kparams& p)
{
    using R = rec_t<CMAX>;
    uint32_t* app32 = reinterpret_cast<uint32_t*>(lds_raw);  // K/2 dwords
    uint32_t* par32 = app32 + p.K / 2;                        // R/2 dwords
    uint32_t* dsl32 = par32 + p.R / 2;                        // n_slots * 180 dwords
    // schedule, group offsets and group records, copied to LDS once so that
    // every later read is an LDS read (no vector-memory wait in the loops)
    int* sched_l = reinterpret_cast<int*>(dsl32 + p.n_slots * HALF);  // [2 n_steps]
    int* goff_l = sched_l + 2 * p.n_steps;                            // [n_steps + 1]
    int* grp_l = goff_l + p.n_steps + 1;                              // [n_groups][8]
    int* flg = grp_l + GROUP_WORDS * p.n_groups;                      // [3] wg_any flags
    int16_t* par16 = reinterpret_cast<int16_t*>(par32);

    constexpr int NTHR = LANES * LANE_THREADS;
    const int t = threadIdx.x;
    const int lane = LANES == 2 ? __builtin_amdgcn_readfirstlane(t >= LANE_THREADS ? 1 : 0) : 0;
    const int tt = t - lane * LANE_THREADS;
    const int tt4 = 4 * tt;
    const bool act = tt < HALF;
    const int q = p.q;

    for (int k = t; k < 2 * p.n_steps; k += NTHR) sched_l[k] = steps[k];
    for (int k = t; k <= p.n_steps; k += NTHR) goff_l[k] = grp_off[k];
    for (int k = t; k < GROUP_WORDS * p.n_groups; k += NTHR) grp_l[k] = groups[k];
    if (t < 3) flg[t] = 0;
    int any_k = 0;

    load_frame_llr<NTHR>(app32, par16, llr, p.K, p.R, q, t);
    __syncthreads();

    // Message records are not cleared: iteration 0 reads zero_records
    // instead, and every record is written by its owner thread before it is
    // first read (each layer once per iteration).
    const int tl = act ? tt : HALF - 1;  // idle lanes load a valid record they ignore
    // Double-buffered message records: even steps read ra and prefetch the
    // next step into rb, odd steps the reverse (the host pads the schedule
    // to an even step count).  The prefetch is unconditional, so the only
    // vector-memory wait in a step is for the record loaded one step earlier.
    // learned parameters, loaded one step ahead with the records (vector
    // buffer loads; iterations >= n_param_iters are plain and load nothing)
    const int n_par = p.n_param_iters * q;
    const __amdgpu_buffer_rsrc_t rs_beta =
        __builtin_amdgcn_make_buffer_rsrc((void*)beta, (short)0, beta ? n_par * 4 : 0, BUF_RSRC_WORD3);
    const __amdgpu_buffer_rsrc_t rs_alpha =
        __builtin_amdgcn_make_buffer_rsrc((void*)alpha, (short)0, alpha ? n_par * 4 : 0, BUF_RSRC_WORD3);
    auto par_load = [&](int itx, int ix, int& b, int& a) {
        b = 1;
        a = 0;
        if (itx < p.n_param_iters) {
            if (beta) b = __builtin_amdgcn_raw_buffer_load_b32(rs_beta, (itx * q + ix) * 4, 0, 0);
            if (alpha) a = __builtin_amdgcn_raw_buffer_load_b32(rs_alpha, (itx * q + ix) * 4, 0, 0);
        }
    };
    static_assert(R::V4 <= 4, "a layer's records fit zero_records");
    uint4 ra[R::V4], rb[R::V4];
    int pb[2], pa[2];  // beta, alpha of the step whose record is in ra / rb
    {
        const int i0 = max(__builtin_amdgcn_readfirstlane(sched_l[lane]), 0);
#pragma unroll
        for (int k = 0; k < R::V4; k++) {
            ra[k] = make_uint4(0, 0, 0, 0);  // step 0 of iteration 0
            rb[k] = make_uint4(0, 0, 0, 0);
        }
        par_load(0, i0, pb[0], pa[0]);
        pb[1] = 1;
        pa[1] = 0;
    }

    int res = -1;
    for (int it = 0;; it++) {
        // ---- syndrome over all checks, step by step, early exit ----
        int ok = 1;
        for (int s = 0; s < p.n_steps; s++) {
            const int i = __builtin_amdgcn_readfirstlane(sched_l[2 * s + lane]);
            int bad = 0;
            if (act && i >= 0) {
                uint32_t sg = par32[i * HALF + tt];
                if (i > 0) {
                    sg ^= par32[(i - 1) * HALF + tt];
                } else {
                    const uint32_t w = par32[(q - 1) * HALF + (tt > 0 ? tt - 1 : HALF - 1)];
                    sg ^= tt > 0 ? w : (w << 16);
                }
                const int cnt = EXACT ? CMAX : lay_cnt[i];
                const circ_rec* cc = circ + lay_off[i];
                for (int e = 0; e < cnt; e++) {
                    const circ_rec c = cc[e];
                    int off;
                    uint32_t sel;
                    pair_addr(c, tt4, tt, off, sel);
                    sg ^= perm(lds_ld(lds_raw, c.app_off + off), sel);
                }
                bad = (sg & 0x80008000u) != 0;
            }
            if (wg_any(bad, flg, any_k)) { ok = 0; break; }
        }
        if (ok) { res = p.max_trials - it; break; }
        if (it == p.max_trials) break;

        // ---- one layered iteration ----
        for (int s0 = 0; s0 < p.n_steps; s0 += 2) {
#pragma unroll
        for (int hs = 0; hs < 2; hs++) {
            const int s = s0 + hs;
            uint4* cur = hs == 0 ? ra : rb;
            uint4* nxt = hs == 0 ? rb : ra;
            const int i = __builtin_amdgcn_readfirstlane(sched_l[2 * s + lane]);
            uint32_t rw[R::WORDS];
#pragma unroll
            for (int k = 0; k < R::V4; k++) {
                rw[4 * k] = cur[k].x;
                rw[4 * k + 1] = cur[k].y;
                rw[4 * k + 2] = cur[k].z;
                rw[4 * k + 3] = cur[k].w;
            }
            // prefetch the record of this thread's next layer (next step, or
            // step 0 of the next iteration)
            {
                const int sn = s + 1 < p.n_steps ? s + 1 : 0;
                const int inx = max(__builtin_amdgcn_readfirstlane(sched_l[2 * sn + lane]), 0);
                const int itn = s + 1 < p.n_steps ? it : it + 1;  // iteration of the next step
                const uint4* lay = itn > 0 ? rec_base + (size_t)inx * HALF * R::V4 : zero_records;
                const uint4* src = lay + tl * R::V4;
#pragma unroll
                for (int k = 0; k < R::V4; k++) nxt[k] = src[k];
                par_load(itn, inx, pb[hs ^ 1], pa[hs ^ 1]);
            }
            if (act && i >= 0) {
                const int bt = __builtin_amdgcn_readfirstlane(pb[hs]);
                const int al = __builtin_amdgcn_readfirstlane(pa[hs]);
                const u2 beta2 = as_u2(splat(bt));
                const int cnt = EXACT ? CMAX : lay_cnt[i];
                const circ_rec* cc = circ + lay_off[i];
                const s2 zero = splat(0);

                s2 xv[CMAX], mv[CMAX], av[CMAX], ov[CMAX];
                int wa[CMAX];        // LDS byte address of the edge's write (APP or delta slot)
                uint32_t sels[CMAX];
                uint32_t slotm = 0;  // edges that write a delta slot (wave-uniform)
                s2 mn1 = splat(32767), mn2 = splat(32767);
                uint32_t sg = 0;
#pragma unroll
                for (int e = 0; e < CMAX; e++) {
                    if (EXACT || e < cnt) {
                        const circ_rec c = cc[e];
                        int off;
                        pair_addr(c, tt4, tt, off, sels[e]);
                        const s2 ap = as_s2(perm(lds_ld(lds_raw, c.app_off + off), sels[e]));
                        wa[e] = (c.slot_off < 0 ? c.app_off : c.slot_off) + off;
                        slotm |= (c.slot_off < 0 ? 0u : 1u) << e;
                        const s2 o = unpack_msg(rw, e);
                        const s2 x = __builtin_elementwise_sub_sat(ap, o);
                        const s2 ab = __builtin_elementwise_max(x, __builtin_elementwise_sub_sat(zero, x));
                        const s2 m = as_s2u(__builtin_elementwise_sub_sat(as_u2(ab), beta2));
                        mn2 = __builtin_elementwise_min(mn2, __builtin_elementwise_max(mn1, m));
                        mn1 = __builtin_elementwise_min(mn1, m);
                        sg ^= bits(x);
                        av[e] = ap;
                        ov[e] = o;
                        xv[e] = x;
                        mv[e] = m;
                    }
                }
                // parity edges: own P(i, j) and previous P(i-1, j) / P(q-1, j-1).
                // Check (0, 0) has no second parity edge: its A half reads a
                // +32767 placeholder that cannot change min1/min2 or the sign.
                const int p0 = i * HALF + tt;
                const bool dummy = i == 0 && tt == 0;
                const int p1 = i > 0 ? (i - 1) * HALF + tt : (q - 1) * HALF + (tt > 0 ? tt - 1 : HALF - 1);
                s2 x0, m0, x1, m1;
                {
                    const s2 o = unpack_msg(rw, CMAX);
                    x0 = __builtin_elementwise_sub_sat(as_s2(par32[p0]), o);
                    const s2 ab = __builtin_elementwise_max(x0, __builtin_elementwise_sub_sat(zero, x0));
                    m0 = as_s2u(__builtin_elementwise_sub_sat(as_u2(ab), beta2));
                    mn2 = __builtin_elementwise_min(mn2, __builtin_elementwise_max(mn1, m0));
                    mn1 = __builtin_elementwise_min(mn1, m0);
                    sg ^= bits(x0);
                }
                {
                    const s2 o = unpack_msg(rw, CMAX + 1);
                    uint32_t w = par32[p1];
                    if (dummy) w = (w << 16) | 0x7fffu;
                    x1 = __builtin_elementwise_sub_sat(as_s2(w), o);
                    const s2 ab = __builtin_elementwise_max(x1, __builtin_elementwise_sub_sat(zero, x1));
                    m1 = as_s2u(__builtin_elementwise_sub_sat(as_u2(ab), beta2));
                    mn2 = __builtin_elementwise_min(mn2, __builtin_elementwise_max(mn1, m1));
                    mn1 = __builtin_elementwise_min(mn1, m1);
                    sg ^= bits(x1);
                }
                // outgoing magnitudes: min(scale(min), 127) per half
                u2 c1, c2;
                if (al == 0) {
                    c1 = __builtin_elementwise_min(as_u2(mn1), (u2)127);
                    c2 = __builtin_elementwise_min(as_u2(mn2), (u2)127);
                } else {
                    c1 = nms_scale_clamp(as_u2(mn1), al);
                    c2 = nms_scale_clamp(as_u2(mn2), al);
                }
                // y = sign(extrinsic) * (m == min1 ? c2 : c1): m >= min1, so
                // (min1 + 1) -| m is 1 exactly on the arg-min edge; its
                // negation is the half-word mask that picks c2 (v_bfi)
                // keep the sign parity in one register (stops re-association
                // of the XOR chain into every edge)
                __asm__("" : "+v"(sg));
                // the slot mask is wave-uniform (see ldpc_kernel_split.h)
                const uint32_t slotu = __builtin_amdgcn_readfirstlane(slotm);
                const u2 mn1p1 = as_u2(mn1) + (u2)1;
                const uint32_t c1b = bits(as_s2u(c1)), c2b = bits(as_s2u(c2));
                auto out = [&](s2 m, s2 x) {
                    const u2 hit = __builtin_elementwise_sub_sat(mn1p1, as_u2(m));
                    const uint32_t mask = bits(as_s2u((u2)0 - hit));
                    const s2 mag = as_s2((mask & c2b) | (~mask & c1b));
                    const s2 sgn = as_s2(sg ^ bits(x)) >> (short)15;
                    return (mag ^ sgn) - sgn;
                };
                s2 ys[CMAX + 2];
#pragma unroll
                for (int e = 0; e < CMAX; e++) {
                    ys[e] = zero;
                    if (EXACT || e < cnt) {
                        const s2 y = out(mv[e], xv[e]);
                        ys[e] = y;
                        const s2 d = y - ov[e];
                        if ((slotu >> e) & 1u)
                            lds_st(lds_raw, wa[e], perm(bits(d), sels[e]));
                        else
                            lds_st(lds_raw, wa[e], perm(bits(__builtin_elementwise_add_sat(av[e], d)), sels[e]));
                    }
                }
                {
                    const s2 y0 = out(m0, x0);
                    par32[p0] = bits(__builtin_elementwise_add_sat(x0, y0));
                    s2 y1 = out(m1, x1);
                    const s2 n1 = __builtin_elementwise_add_sat(x1, y1);
                    if (!dummy) {
                        par32[p1] = bits(n1);
                    } else {
                        par16[2 * p1] = n1.y;  // low half of dword p1 = P(q-1, 179)
                        y1.x = 0;
                    }
                    ys[CMAX] = y0;
                    ys[CMAX + 1] = y1;
                }
                uint32_t ow[R::WORDS];
#pragma unroll
                for (int k = 0; k < R::WORDS; k++)
                    ow[k] = 2 * k + 1 < CMAX + 2 ? pack_msg(ys[2 * k], ys[2 * k + 1])
                          : 2 * k < CMAX + 2     ? pack_msg(ys[2 * k], zero)
                                                 : 0u;
                uint4* dst = rec_base + ((size_t)i * HALF + tt) * R::V4;
#pragma unroll
                for (int k = 0; k < R::V4; k++)
                    dst[k] = make_uint4(ow[4 * k], ow[4 * k + 1], ow[4 * k + 2], ow[4 * k + 3]);
            }
            __syncthreads();
            const int g0 = __builtin_amdgcn_readfirstlane(goff_l[s]);
            const int g1 = __builtin_amdgcn_readfirstlane(goff_l[s + 1]);
            if (g1 > g0) {
                // one group per lane at a time (group id wave-uniform); thread
                // tt adds the deltas of pair dword tt in slot order
                for (int gb = g0; gb < g1; gb += LANES) {
                    const int g = gb + lane;
                    if (g < g1 && act) {
                        const int* gr = grp_l + GROUP_WORDS * g;
                        const int base = __builtin_amdgcn_readfirstlane(gr[0]);
                        const int n = __builtin_amdgcn_readfirstlane(gr[1]);
                        s2 val = as_s2(app32[base + tt]);
                        for (int e = 0; e < n; e++)
                            val = __builtin_elementwise_add_sat(
                                val, as_s2(app32[__builtin_amdgcn_readfirstlane(gr[2 + e]) + tt]));
                        app32[base + tt] = bits(val);
                    }
                }
                __syncthreads();
            }
        }
        }
    }

    // hard decisions, K bits MSB first (ldpc_decode_kernel.cu:468-493)
    store_hard<NTHR>(app32, h, p.K, t);
    return res;
}

template <int CMAX, bool EXACT, int LANES>
__global__ void __launch_bounds__(LANES* LANE_THREADS)
    ldpc_kernel(const int16_t* __restrict__ llr_all, uint4* __restrict__ msg_all,
                uint8_t* __restrict__ hard_all, int32_t* __restrict__ result,
                const circ_rec* __restrict__ circ, const int* __restrict__ lay_cnt,
                const int* __restrict__ lay_off, const int* __restrict__ steps,
                const int* __restrict__ grp_off, const int* __restrict__ groups,
                const int* __restrict__ beta, const int* __restrict__ alpha,
                const int32_t* __restrict__ active, kparams p)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int cw = blockIdx.x;
    // frames masked off (active[cw] == 0) keep their outputs; the whole
    // workgroup leaves before its first barrier
    if (active && active[cw] == 0) return;
    const int res = decode_frame<CMAX, EXACT, LANES>(
        lds_raw, llr_all + (size_t)cw * p.N, msg_all + (size_t)cw * p.q * HALF * rec_t<CMAX>::V4,
        hard_all + (size_t)cw * (p.K >> 3), circ, lay_cnt, lay_off, steps, grp_off, groups, beta, alpha, p);
    if (threadIdx.x == 0) result[cw] = res;
}

}  // namespace ldpc_dev

template <int CMAX, bool EXACT, int LANES>
hipError_t launch_ldpc_variant(const ldpc_args& a, int batch, hipStream_t stream)
{
    auto kern = &ldpc_dev::ldpc_kernel<CMAX, EXACT, LANES>;
    const int lds = ldpc_lds_bytes(a.K, a.R, a.n_slots, a.n_steps, a.n_groups);
    static int configured = 0;
    if (lds > __atomic_load_n(&configured, __ATOMIC_ACQUIRE)) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        __atomic_store_n(&configured, lds, __ATOMIC_RELEASE);
    }
    const ldpc_dev::kparams p{a.N, a.K, a.R, a.q, a.n_steps, a.max_trials, a.n_param_iters,
                              a.n_slots, a.n_groups};
    hipLaunchKernelGGL(kern, dim3(batch), dim3(LANES * ldpc_dev::LANE_THREADS), lds, stream, a.llr,
                       reinterpret_cast<uint4*>(a.msg), a.hard, a.result,
                       reinterpret_cast<const ldpc_dev::circ_rec*>(a.circ), a.lay_cnt, a.lay_off,
                       a.steps, a.grp_off, a.groups, a.beta, a.alpha, a.active, p);
    return hipGetLastError();
}

#define CUSAT_LDPC_INSTANTIATE(C, E, L) \
    template hipError_t launch_ldpc_variant<C, E, L>(const ldpc_args&, int, hipStream_t);

}  // namespace cusat
