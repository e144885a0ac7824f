// Split-edge variant of the layered min-sum decode kernel (device code;
// instantiated in ldpc_inst_s.hip).  Same arithmetic, LDS layout and
// collision handling as ldpc_kernel.h, with the work of a check pair divided
// between two groups of waves so that a workgroup runs 12 waves (3 per
// SIMD) instead of 6 (2, 2, 1, 1):
//
//  * A lane (layer of the schedule step) is 6 waves: waves 0-2 ("half 0")
//    evaluate data edges 0 .. n0-1 and the parity edge P(i, j) of the check
//    pairs (t, t + 180), waves 3-5 ("half 1") data edges n0 .. cnt-1 and the
//    second parity edge, where cnt is the layer's number of data circulants
//    (CMAX for an exact-degree code, lay_cnt[i] <= CMAX otherwise) and
//    n0 = ceil(cnt / 2).  Every wave still walks one circulant at a time,
//    so circulant records stay scalar loads.
//  * Each half reduces its edges to (min1, min2, sign) of |x| -- the offset
//    beta is applied once per check, max(min|x| - beta, 0) being the minimum
//    of max(|x| - beta, 0) -- and the halves swap these partials through LDS
//    (one barrier); then each half writes its own edges.
//  * A message record holds one half's edges: [layer][half][pair] x 4, 8 or
//    a multiple of 16 bytes (the bytes its edges take, 2 per edge, rounded
//    up to a load size), so a wave moves consecutive records.
//
// Reads of a step see the step-start APP and every variable node is written
// once per step outside the collision slots, as in ldpc_kernel.h, so the
// results are the same bit for bit.
#pragma once

#include "ldpc_kernel.h"

namespace cusat {
namespace ldpc_dev {

constexpr int SPLIT_LANE_THREADS = 2 * LANE_THREADS;  // 6 waves per lane
constexpr int SPLIT_LDS_STATIC = 160 * 1024;           // the whole LDS of a gfx950 workgroup

// message record of one half: (CMAX + 1) / 2 data slots + 1 parity slot,
// two edges per word; stored as 1 or 2 words when those hold every edge,
// else as a multiple of 4 words (16-byte loads)
template <int CMAX>
struct rec_half_t {
    static constexpr int NE0 = (CMAX + 1) / 2;
    static constexpr int NE1 = CMAX - NE0;
    static constexpr int NEH = NE0 + 1;  // parity edge in the last slot
    static constexpr int USED = (NEH + 1) / 2;
    static constexpr int WORDS = USED <= 2 ? USED : (USED + 3) / 4 * 4;
};

// one half record between global memory and registers, in the widest loads
// its size allows
template <int W>
__device__ __forceinline__ void rec_load(uint32_t* r, const uint32_t* __restrict__ src)
{
    if constexpr (W == 1) {
        r[0] = src[0];
    } else if constexpr (W == 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(src);
        r[0] = v.x;
        r[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < W / 4; k++) {
            const uint4 v = reinterpret_cast<const uint4*>(src)[k];
            r[4 * k] = v.x;
            r[4 * k + 1] = v.y;
            r[4 * k + 2] = v.z;
            r[4 * k + 3] = v.w;
        }
    }
}

template <int W>
__device__ __forceinline__ void rec_store(uint32_t* __restrict__ dst, const uint32_t* r)
{
    if constexpr (W == 1) {
        dst[0] = r[0];
    } else if constexpr (W == 2) {
        *reinterpret_cast<uint2*>(dst) = make_uint2(r[0], r[1]);
    } else {
#pragma unroll
        for (int k = 0; k < W / 4; k++)
            reinterpret_cast<uint4*>(dst)[k] = make_uint4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
    }
}

// One collision group at pair dword tt: APP = sat(...sat(APP + d_1)... + d_n)
// in slot order.  The group record comes in two 16-byte reads; a record's
// unused slot entries are 0 (in-bounds reads never added), so all six slot
// reads are issued before the first add: two LDS round trips per group.
__device__ __forceinline__ void group_apply(uint32_t* app32, const int* rec, int tt)
{
    const int4 ra4 = *reinterpret_cast<const int4*>(rec);
    const int4 rb4 = *reinterpret_cast<const int4*>(rec + 4);
    const int base = __builtin_amdgcn_readfirstlane(ra4.x);
    const int n = __builtin_amdgcn_readfirstlane(ra4.y);
    const int sl[6] = {__builtin_amdgcn_readfirstlane(ra4.z), __builtin_amdgcn_readfirstlane(ra4.w),
                       __builtin_amdgcn_readfirstlane(rb4.x), __builtin_amdgcn_readfirstlane(rb4.y),
                       __builtin_amdgcn_readfirstlane(rb4.z), __builtin_amdgcn_readfirstlane(rb4.w)};
    const uint32_t a0 = app32[base + tt];
    uint32_t dl[6];
#pragma unroll
    for (int e = 0; e < 6; e++) dl[e] = app32[sl[e] + tt];
    s2 val = as_s2(a0);
#pragma unroll
    for (int e = 0; e < 6; e++)
        if (e < n) val = __builtin_elementwise_add_sat(val, as_s2(dl[e]));
    app32[base + tt] = bits(val);
}

// One frame per workgroup (PAIR = false), or, for a code with a one-lane
// schedule, two frames of that code at once (PAIR = true, LANES = 2): lane
// L decodes frame L (llr / rec_base / h, then llr1 / rec1 / h1) in its own
// LDS region, both lanes step through the shared schedule together, and a
// lane whose frame has met its syndrome idles until the other's is done.
// Returns the frame's trial count (PAIR: frame 0's in the low 16 bits,
// frame 1's in the high 16 bits, each an int16).
template <int CMAX, bool EXACT, int LANES, bool PAIR = false>
__device__ __forceinline__ int decode_frame_split(unsigned char* lds_raw, const int16_t* __restrict__ llr,
                                                  uint4* __restrict__ rec_base, uint8_t* __restrict__ h,
                                                  const circ_rec* __restrict__ circ,
                                                  const int* __restrict__ lay_cnt, const int* __restrict__ lay_off,
                                                  const int* __restrict__ steps, const int* __restrict__ grp_off,
                                                  const int* __restrict__ step_defer,
                                                  const int* __restrict__ groups, const int* __restrict__ beta,
                                                  const int* __restrict__ alpha, const kparams& p,
                                                  const int16_t* __restrict__ llr1 = nullptr,
                                                  uint4* __restrict__ rec1 = nullptr, uint8_t* __restrict__ h1 = nullptr)
{
    static_assert(!PAIR || LANES == 2, "a pair takes both lanes");
    using R = rec_half_t<CMAX>;
    constexpr int NE0 = R::NE0, NE1 = R::NE1, NEH = R::NEH;
    constexpr int NTHR = LANES * SPLIT_LANE_THREADS;
    const int t = threadIdx.x;
    const int lane = LANES == 2 ? __builtin_amdgcn_readfirstlane(t >= SPLIT_LANE_THREADS ? 1 : 0) : 0;
    // PAIR: frame region of lane L at L * frw dwords (its APP, parity and
    // slot areas, laid out as for one frame); the shared tables after both
    const int frw = PAIR ? p.K / 2 + p.R / 2 + p.n_slots * HALF : 0;
    unsigned char* lds_f = PAIR ? lds_raw + lane * frw * 4 : lds_raw;  // this lane's frame region
    if constexpr (PAIR) {
        llr = lane ? llr1 : llr;
        rec_base = lane ? rec1 : rec_base;
        h = lane ? h1 : h;
    }
    uint32_t* app32 = reinterpret_cast<uint32_t*>(lds_f);  // K/2 dwords
    uint32_t* par32 = app32 + p.K / 2;                      // R/2 dwords
    uint32_t* dsl32 = par32 + p.R / 2;                      // n_slots * 180 dwords
    uint32_t* xch = PAIR ? reinterpret_cast<uint32_t*>(lds_raw) + 2 * frw
                         : dsl32 + p.n_slots * HALF;        // [3][LANES][2][180] partials
    // per (step, lane): layer (-1 idle), its lay_off and data-circulant
    // count, and the step's first collision group, so that a step's first
    // global access is its circulant records
    int4* sched_l = reinterpret_cast<int4*>(xch + 3 * LANES * 2 * HALF);
    // per (step, lane): the learned (beta, alpha) of the current iteration,
    // written during the iteration's syndrome pass
    int2* parl = reinterpret_cast<int2*>(sched_l + 2 * p.n_steps);
    int* grp_l = reinterpret_cast<int*>(parl + 2 * p.n_steps);  // [n_groups][8], 16-byte aligned
    int* flg = grp_l + GROUP_WORDS * p.n_groups;  // [3] wg_any flags (PAIR: [3][2])
    int16_t* par16 = reinterpret_cast<int16_t*>(par32);

    const int u = t - lane * SPLIT_LANE_THREADS;
    const int half = __builtin_amdgcn_readfirstlane(u >= LANE_THREADS ? 1 : 0);
    const int tt = u - half * LANE_THREADS;
    const int tt4 = 4 * tt;
    const int tt4w = tt4 + 4 * HALF;  // byte offset of the pair dword after a wrap
    const bool act = tt < HALF;
    const int wih = __builtin_amdgcn_readfirstlane(tt >> 6);  // wave index within the half
    const int q = p.q;
    const int quarter = lane * 2 + half;  // which (lane, half) this wave group is
    // the schedule entry of this lane (PAIR: both lanes run lane 0's layers)
    const int sln = PAIR ? 0 : lane;
    // collision groups: wave group gq of gs applies groups gq, gq + gs, ...
    // of the frame (PAIR: the two halves of each lane, on the lane's frame)
    const int gq = PAIR ? half : quarter;
    constexpr int gs = PAIR ? 2 : 2 * LANES;
    // PAIR: bit L set = frame L has met its syndrome (the same in every
    // thread); dn = this lane's frame is done and it only keeps step
    int dmask = 0;
    bool dn = false;
    uint32_t* xm1 = xch;
    uint32_t* xm2 = xch + LANES * 2 * HALF;
    uint32_t* xsg = xch + 2 * LANES * 2 * HALF;

    for (int k = t; k < 2 * p.n_steps; k += NTHR) {
        const int i = steps[k];
        const int ic = max(i, 0);
        // .w: first collision group of the step | deferred-groups flag << 30
        sched_l[k] = make_int4(i, lay_off[ic], EXACT ? CMAX : lay_cnt[ic],
                               grp_off[k >> 1] | (step_defer[k >> 1] ? 1 << 30 : 0));
    }
    for (int k = t; k < GROUP_WORDS * p.n_groups; k += NTHR) grp_l[k] = groups[k];
    if (t < (PAIR ? 6 : 3)) flg[t] = 0;
    int any_k = 0;
    const int n_groups_all = p.n_groups;  // first group past the last step
    if constexpr (PAIR)
        load_frame_llr<SPLIT_LANE_THREADS>(app32, par16, llr, p.K, p.R, q, u);
    else
        load_frame_llr<NTHR>(app32, par16, llr, p.K, p.R, q, t);
    __syncthreads();

    // Message records are not cleared: iteration 0 reads zero_records
    // instead, and every record is written before it is first read (each
    // layer once per iteration).
    const int tl = act ? tt : HALF - 1;
    uint32_t* rec32 = reinterpret_cast<uint32_t*>(rec_base);
    auto rec_at = [&](int i, int pair) { return rec32 + (((size_t)i * 2 + half) * HALF + pair) * R::WORDS; };
    // Learned parameters: the (beta, alpha) of every (step, lane) entry of an
    // iteration are loaded by thread 2 s + lane during that iteration's
    // syndrome pass and stored in parl before the pass's barrier; a step
    // reads its own entry at the exchange.  Plain iterations (>= n_param_iters)
    // load nothing.  (Loaded one step ahead in registers instead, the value
    // was copied at the loop's join points, and each copy waited for every
    // load in flight -- the next step's record prefetch included.)
    static_assert(R::WORDS <= 16, "a layer's records fit zero_records");
    uint32_t ra[R::WORDS], rb[R::WORDS];
    int si[2], so[2], sc[2], sg0[2];  // layer, lay_off, circulant count, first group of the step in ra / rb
    {
        const int4 e0 = sched_l[sln];
        si[0] = __builtin_amdgcn_readfirstlane(e0.x);
        so[0] = __builtin_amdgcn_readfirstlane(e0.y);
        sc[0] = __builtin_amdgcn_readfirstlane(e0.z);
        sg0[0] = __builtin_amdgcn_readfirstlane(e0.w);  // step 0
        si[1] = -1;
        so[1] = 0;
        sc[1] = 0;
        sg0[1] = 0;
#pragma unroll
        for (int k = 0; k < R::WORDS; k++) {
            ra[k] = 0;  // step 0 of iteration 0
            rb[k] = 0;
        }
    }

    int res = -1, res1 = -1;  // res1: frame 1 of a pair
    for (int it = 0;; it++) {
        const bool lrn = it < p.n_param_iters;  // wave-uniform
        int pbv = 1, pav = 0;
        if (lrn && t < 2 * p.n_steps) {
            const int i = sched_l[t].x;
            if (i >= 0) {
                if (beta) pbv = beta[it * q + i];
                if (alpha) pav = alpha[it * q + i];
            }
        }
        // ---- syndrome of the APP signs, early exit (ldpc_decode_kernel.cu:689-749) ----
        // Per barrier, two schedule steps: wave group (lane, half) evaluates
        // the whole check pairs of its lane's layer in step s + half.  The
        // result is the same for any order of the checks.  PAIR: wave group
        // (lane, half) checks step s + half of its lane's frame; the scan
        // ends once each frame is done or has a failed check.
        int ok = 1;
        int failm = 0;  // PAIR: bit L = frame L has a failed check
        for (int s = 0; s < p.n_steps; s += 2) {
            const int sx = min(s + half, p.n_steps - 1);
            const int4 se = sched_l[2 * sx + sln];
            const int i = __builtin_amdgcn_readfirstlane(se.x);
            int bad = 0;
            if (act && !dn && i >= 0 && s + half < p.n_steps) {
                uint32_t sg = par32[i * HALF + tt];
                if (i > 0) {
                    sg ^= par32[(i - 1) * HALF + tt];
                } else {
                    const uint32_t w = par32[(q - 1) * HALF + (tt > 0 ? tt - 1 : HALF - 1)];
                    sg ^= tt > 0 ? w : (w << 16);
                }
                const circ_rec* cc = circ + __builtin_amdgcn_readfirstlane(se.y);
                const int cnt = EXACT ? CMAX : __builtin_amdgcn_readfirstlane(se.z);
#pragma unroll
                for (int e = 0; e < CMAX; e++) {
                    if (EXACT || e < cnt) {
                        const circ_rec c = cc[e];
                        int off;
                        uint32_t sel;
                        pair_addr(c, tt4, tt, off, sel);
                        sg ^= perm(lds_ld(lds_f, c.app_off + off), sel);
                    }
                }
                bad = (sg & 0x80008000u) != 0;
            }
            if constexpr (PAIR) {
                failm |= wg_any2(bad, lane, flg, any_k);
                if ((failm | dmask) == 3) break;  // workgroup-uniform
            } else {
                if (wg_any(bad, flg, any_k)) { ok = 0; break; }
            }
        }
        // the learned parameters, loaded before the scan: read at the
        // exchanges, after the first step's barrier
        if (lrn && t < 2 * p.n_steps) parl[t] = make_int2(pbv, pav);
        // vmcnt(0) in uniform control flow: the compiler cannot tie the
        // parameter loads to the store's branch and would otherwise count
        // them as outstanding in the step loop, waiting at every step
        __builtin_amdgcn_s_waitcnt(0x0F70);
        if constexpr (PAIR) {
            const int okm = ~(failm | dmask) & 3;  // frames that meet their syndrome now
            if (okm & 1) res = p.max_trials - it;
            if (okm & 2) res1 = p.max_trials - it;
            dmask |= okm;
            dn = (dmask >> lane) & 1;
            if (dmask == 3) break;
        } else {
            if (ok) { res = p.max_trials - it; break; }
        }
        if (it == p.max_trials) break;

        // ---- one layered iteration ----
        // collision groups of the previous step left to the current one
        // (none at an iteration's start: its last step is never deferred)
        int prev_g0 = 0, prev_g1 = 0;
        bool prev_def = false;
        for (int s0 = 0; s0 < p.n_steps; s0 += 2) {
#pragma unroll
        for (int hs = 0; hs < 2; hs++) {
            const int s = s0 + hs;
            uint32_t* cur = hs == 0 ? ra : rb;
            uint32_t* nxt = hs == 0 ? rb : ra;
            // this step's schedule entry, read one step earlier
            const int i = si[hs];
            const int loff = so[hs];
            const int lcnt = sc[hs];
            uint32_t rw[R::WORDS];
#pragma unroll
            for (int k = 0; k < R::WORDS; k++) rw[k] = cur[k];
            {
                const int sn = s + 1 < p.n_steps ? s + 1 : 0;
                const int4 en = sched_l[2 * sn + sln];
                si[hs ^ 1] = __builtin_amdgcn_readfirstlane(en.x);
                so[hs ^ 1] = __builtin_amdgcn_readfirstlane(en.y);
                sc[hs ^ 1] = __builtin_amdgcn_readfirstlane(en.z);
                sg0[hs ^ 1] = __builtin_amdgcn_readfirstlane(en.w);
                const int inx = max(si[hs ^ 1], 0);
                const int itn = s + 1 < p.n_steps ? it : it + 1;  // iteration of the next step
                const uint32_t* lay = itn > 0 ? rec32 + ((size_t)inx * 2 + half) * HALF * R::WORDS
                                              : reinterpret_cast<const uint32_t*>(zero_records);
                rec_load<R::WORDS>(nxt, lay + tl * R::WORDS);
            }
            const bool work = act && !dn && i >= 0;
            const s2 zero = splat(0);
            // data edges of the layer and of this half (wave-uniform)
            const int cnt = EXACT ? CMAX : lcnt;
            const int n0 = EXACT ? NE0 : (cnt + 1) >> 1;
            const int ne = half ? cnt - n0 : n0;
            s2 xv[NE0], mv[NE0], av[NE0], ov[NE0];
            int ra_[NE0];  // LDS byte address of the edge's pair dword (its read and, outside
                           // the collision slots, its write)
            uint32_t sels[NE0];
            // set at the start of the read phase and read only under `work`
            // (no moves for the phi of the idle path)
            uint32_t slotm;
            s2 mn1, mn2;
            uint32_t sg;
            // parity edge: half 0 the own P(i, j), half 1 P(i-1, j) / P(q-1, j-1);
            // check (0, 0) has no second parity edge: its A half reads a +32767
            // placeholder that cannot change min1/min2 or the sign
            const int pidx = half == 0 ? i * HALF + tt
                                       : (i > 0 ? (i - 1) * HALF + tt : (q - 1) * HALF + (tt > 0 ? tt - 1 : HALF - 1));
            // Deferred groups of the previous step: their rows are untouched by
            // this step, so wave group gq applies group prev_g0 + gq alongside
            // the step -- reads issued here, the result written before the
            // exchange barrier (more groups than wave groups: the rest after
            // it, synchronously).
            const bool mat = prev_def && !dn && prev_g0 + gq < prev_g1;  // wave-uniform
            // set and read only under mat (no zeroing moves on the common path)
            int m_base, m_n;
            uint32_t m_a0, m_d[6];
            if (mat) {
                const int g = prev_g0 + gq;
                const int4 ga = *reinterpret_cast<const int4*>(grp_l + GROUP_WORDS * g);
                const int4 gb = *reinterpret_cast<const int4*>(grp_l + GROUP_WORDS * g + 4);
                m_base = __builtin_amdgcn_readfirstlane(ga.x);
                m_n = __builtin_amdgcn_readfirstlane(ga.y);
                const int sl[6] = {__builtin_amdgcn_readfirstlane(ga.z), __builtin_amdgcn_readfirstlane(ga.w),
                                   __builtin_amdgcn_readfirstlane(gb.x), __builtin_amdgcn_readfirstlane(gb.y),
                                   __builtin_amdgcn_readfirstlane(gb.z), __builtin_amdgcn_readfirstlane(gb.w)};
                if (act) {
                    m_a0 = app32[m_base + tt];
#pragma unroll
                    for (int e = 0; e < 6; e++) m_d[e] = app32[sl[e] + tt];
                }
            }
            // the wave holding thread tt = 0 of half 1 in layer 0 (wave-uniform):
            // only there does a thread (tt = 0) take the placeholder edge
            const bool dwave = half == 1 && i == 0 && wih == 0;
            const bool dummy = dwave && tt == 0;
            s2 xp, mp;
            const circ_rec* cc = circ + loff + half * n0;
            if (work) {
                slotm = 0;
                mn1 = splat(32767);
                mn2 = splat(32767);
                sg = 0;
                // both halves with every slot used (exact degree, even CMAX):
                // the circulant records and APP dwords of all edges first, so
                // that the loads are in flight together (the other variants
                // keep them inside the edge guard, which measured faster there)
                constexpr bool HOIST = EXACT && NE1 == NE0;
                uint32_t apw[NE0];
                if (HOIST) {
#pragma unroll
                    for (int e = 0; e < NE0; e++) {
                        const circ_rec c = cc[e];
                        const bool wrap = tt < c.sp;
                        const int off = wrap ? tt4w : tt4;
                        sels[e] = (wrap != (c.sel != SEL_ID)) ? SEL_ROT : SEL_ID;
                        ra_[e] = c.app_off + off;
                        apw[e] = lds_ld(lds_f, ra_[e]);
                        slotm |= (c.slot_off < 0 ? 0u : 1u) << e;
                    }
                }
#pragma unroll
                for (int e = 0; e < NE0; e++) {
                    if (e < ne) {
                        if (!HOIST) {
                            const circ_rec c = cc[e];
                            // thread tt's pair dword wraps iff tt < sp; a wrap adds 180
                            // dwords and swaps the halves
                            const bool wrap = tt < c.sp;
                            const int off = wrap ? tt4w : tt4;
                            sels[e] = (wrap != (c.sel != SEL_ID)) ? SEL_ROT : SEL_ID;
                            ra_[e] = c.app_off + off;
                            apw[e] = lds_ld(lds_f, ra_[e]);
                            slotm |= (c.slot_off < 0 ? 0u : 1u) << e;
                        }
                        const s2 ap = as_s2(perm(apw[e], sels[e]));
                        const s2 o = unpack_msg(rw, e);
                        const s2 x = __builtin_elementwise_sub_sat(ap, o);
                        const s2 m = __builtin_elementwise_max(x, __builtin_elementwise_sub_sat(zero, x));
                        mn2 = __builtin_elementwise_min(mn2, __builtin_elementwise_max(mn1, m));
                        mn1 = __builtin_elementwise_min(mn1, m);
                        sg ^= bits(x);
                        av[e] = ap;
                        ov[e] = o;
                        xv[e] = x;
                        mv[e] = m;
                    }
                }
                {
                    const s2 o = unpack_msg(rw, NEH - 1);
                    uint32_t w = par32[pidx];
                    if (dwave && dummy) w = (w << 16) | 0x7fffu;
                    xp = __builtin_elementwise_sub_sat(as_s2(w), o);
                    mp = __builtin_elementwise_max(xp, __builtin_elementwise_sub_sat(zero, xp));
                    mn2 = __builtin_elementwise_min(mn2, __builtin_elementwise_max(mn1, mp));
                    mn1 = __builtin_elementwise_min(mn1, mp);
                    sg ^= bits(xp);
                }
                const int xo = quarter * HALF + tt;
                xm1[xo] = bits(mn1);
                xm2[xo] = bits(mn2);
                xsg[xo] = sg;
            }
            if (mat && act) {
                s2 val = as_s2(m_a0);
#pragma unroll
                for (int e = 0; e < 6; e++)
                    if (e < m_n) val = __builtin_elementwise_add_sat(val, as_s2(m_d[e]));
                app32[m_base + tt] = bits(val);
            }
            __syncthreads();  // every read of the step is done; partials are in LDS
            if (prev_def && act && !dn)
                for (int g = prev_g0 + gq + gs; g < prev_g1; g += gs)
                    group_apply(app32, grp_l + GROUP_WORDS * g, tt);
            if (work) {
                {
                    const int po = (quarter ^ 1) * HALF + tt;  // the other half of the same lane
                    const s2 om1 = as_s2(xm1[po]), om2 = as_s2(xm2[po]);
                    mn2 = __builtin_elementwise_min(__builtin_elementwise_min(mn2, om2),
                                                    __builtin_elementwise_max(mn1, om1));
                    mn1 = __builtin_elementwise_min(mn1, om1);
                    sg ^= xsg[po];
                }
                int bt = 1, al = 0;
                if (lrn) {
                    const int2 pr = parl[2 * s + sln];
                    bt = __builtin_amdgcn_readfirstlane(pr.x);
                    al = __builtin_amdgcn_readfirstlane(pr.y);
                }
                const u2 beta2 = as_u2(splat(bt));
                // offset once per check: max(min|x| - beta, 0)
                const s2 b1 = as_s2u(__builtin_elementwise_sub_sat(as_u2(mn1), beta2));
                const s2 b2 = as_s2u(__builtin_elementwise_sub_sat(as_u2(mn2), beta2));
                u2 c1, c2;
                if (al == 0) {
                    c1 = __builtin_elementwise_min(as_u2(b1), (u2)127);
                    c2 = __builtin_elementwise_min(as_u2(b2), (u2)127);
                } else {
                    c1 = nms_scale_clamp(as_u2(b1), al);
                    c2 = nms_scale_clamp(as_u2(b2), al);
                }
                __asm__("" : "+v"(sg));
                // which edges write a delta slot: the same for every lane of the
                // wave (circulant records are wave-uniform); taking it from the
                // first active lane makes the per-edge choice a scalar branch
                const uint32_t slotu = __builtin_amdgcn_readfirstlane(slotm);
                // |x_e| == min1 exactly on the arg-min edge (all |x| >= min1)
                const u2 mn1p1 = as_u2(mn1) + (u2)1;
                const u2 dc = c2 - c1;  // >= 0: the scale and clamp are monotone
                auto out = [&](s2 m, s2 x) {
                    // hit = 1 on the arg-min edge, else 0: magnitude c1 + hit (c2 - c1)
                    const u2 hit = __builtin_elementwise_sub_sat(mn1p1, as_u2(m));
                    const s2 mag = as_s2u(c1 + hit * dc);
                    const s2 sgn = as_s2(sg ^ bits(x)) >> (short)15;
                    return (mag ^ sgn) - sgn;
                };
                s2 ys[NEH];
                if (slotu == 0) {
                    // no edge of this quarter collides (most quarters): every
                    // edge updates the APP, no per-edge scalar test
#pragma unroll
                    for (int e = 0; e < NE0; e++) {
                        ys[e] = zero;
                        if (e < ne) {
                            const s2 y = out(mv[e], xv[e]);
                            ys[e] = y;
                            lds_st(lds_f, ra_[e],
                                   perm(bits(__builtin_elementwise_add_sat(av[e], y - ov[e])), sels[e]));
                        }
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < NE0; e++) {
                        ys[e] = zero;
                        if (e < ne) {
                            const s2 y = out(mv[e], xv[e]);
                            ys[e] = y;
                            const s2 d = y - ov[e];
                            if ((slotu >> e) & 1u) {
                                const circ_rec c = cc[e];
                                lds_st(lds_f, ra_[e] + (c.slot_off - c.app_off), perm(bits(d), sels[e]));
                            } else {
                                lds_st(lds_f, ra_[e],
                                       perm(bits(__builtin_elementwise_add_sat(av[e], d)), sels[e]));
                            }
                        }
                    }
                }
                {
                    s2 yp = out(mp, xp);
                    const s2 np = __builtin_elementwise_add_sat(xp, yp);
                    if (!dwave) {
                        par32[pidx] = bits(np);
                    } else if (!dummy) {
                        par32[pidx] = bits(np);
                    } else {
                        par16[2 * pidx] = np.y;  // low half of dword P(q-1, 179)
                        yp.x = 0;
                    }
                    ys[NEH - 1] = yp;
                }
                uint32_t ow[R::WORDS];
#pragma unroll
                for (int k = 0; k < R::WORDS; k++)
                    ow[k] = 2 * k + 1 < NEH ? pack_msg(ys[2 * k], ys[2 * k + 1])
                          : 2 * k < NEH     ? pack_msg(ys[2 * k], zero)
                                            : 0u;
                rec_store<R::WORDS>(rec_at(i, tt), ow);
            }
            __syncthreads();
            // collision groups of the step: [first group of this step, first
            // of the next) -- the next step's entry was read at the prefetch
            const int g0 = sg0[hs] & 0x3fffffff;
            const bool defer = (sg0[hs] >> 30) & 1;
            const int g1 = s + 1 < p.n_steps ? (sg0[hs ^ 1] & 0x3fffffff) : n_groups_all;
            prev_g0 = g0;
            prev_g1 = g1;
            prev_def = defer && g1 > g0;
            if (g1 > g0 && !defer) {
                // one group per wave group at a time; thread tt adds the deltas
                // of pair dword tt in slot order.  The group record comes in
                // two 16-byte reads and every slot read is issued before the
                // first add, so a group costs two LDS round trips.
                for (int gb = g0; gb < g1; gb += gs) {
                    const int g = gb + gq;
                    if (g < g1 && act && !dn) group_apply(app32, grp_l + GROUP_WORDS * g, tt);
                }
                __syncthreads();
            }
        }
        }
    }

    if constexpr (PAIR) {
        // a done lane's APP has not changed since its frame met its syndrome
        store_hard<SPLIT_LANE_THREADS>(app32, h, p.K, u);
        return (int)(((uint32_t)res & 0xffffu) | ((uint32_t)res1 << 16));
    } else {
        store_hard<NTHR>(app32, h, p.K, t);
        return res;
    }
}

template <int CMAX, bool EXACT, int LANES>
__global__ void __launch_bounds__(LANES* SPLIT_LANE_THREADS)
    ldpc_kernel_split(const int16_t* __restrict__ llr_all, uint4* __restrict__ msg_all,
                      uint8_t* __restrict__ hard_all, int32_t* __restrict__ result,
                      const circ_rec* __restrict__ circ, const int* __restrict__ lay_cnt,
                      const int* __restrict__ lay_off,
                      const int* __restrict__ steps, const int* __restrict__ grp_off,
                      const int* __restrict__ step_defer, const int* __restrict__ groups,
                      const int* __restrict__ beta, const int* __restrict__ alpha,
                      const int32_t* __restrict__ active, kparams p)
{
    // Normal frames (two layers per step) fill most of the LDS, one
    // workgroup per CU either way: their instantiations take the whole LDS
    // as a static array, whose address (0) is known when the kernel is
    // compiled, so no instruction adds the base of a dynamic allocation to
    // every LDS address.  Short frames keep the dynamic size that lets
    // several workgroups share a CU.
    unsigned char* lds_raw;
    if constexpr (LANES == 2) {
        __shared__ __align__(16) unsigned char lds_fixed[SPLIT_LDS_STATIC];
        lds_raw = lds_fixed;
    } else {
        extern __shared__ __align__(16) unsigned char lds_dyn[];
        lds_raw = lds_dyn;
    }
    const int cw = blockIdx.x;
    if (active && active[cw] == 0) return;
    const int res = decode_frame_split<CMAX, EXACT, LANES>(
        lds_raw, llr_all + (size_t)cw * p.N,
        reinterpret_cast<uint4*>(reinterpret_cast<uint32_t*>(msg_all) + (size_t)cw * p.q * 2 * HALF * rec_half_t<CMAX>::WORDS),
        hard_all + (size_t)cw * (p.K >> 3), circ, lay_cnt, lay_off, steps, grp_off, step_defer, groups, beta,
        alpha, p);
    if (threadIdx.x == 0) result[cw] = res;
}

}  // namespace ldpc_dev

template <int CMAX, bool EXACT, int LANES>
hipError_t launch_ldpc_split_variant(const ldpc_args& a, int batch, hipStream_t stream)
{
    auto kern = &ldpc_dev::ldpc_kernel_split<CMAX, EXACT, LANES>;
    int lds = ldpc_lds_bytes_split(a.K, a.R, a.n_slots, a.n_steps, a.n_groups, LANES);
    if (LANES == 2) {  // static LDS (see ldpc_kernel_split)
        if (lds > ldpc_dev::SPLIT_LDS_STATIC) return hipErrorInvalidValue;
        lds = 0;
    }
    static int configured = 0;
    if (lds > __atomic_load_n(&configured, __ATOMIC_ACQUIRE)) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        __atomic_store_n(&configured, lds, __ATOMIC_RELEASE);
    }
    const ldpc_dev::kparams p{a.N, a.K, a.R, a.q, a.n_steps, a.max_trials, a.n_param_iters,
                              a.n_slots, a.n_groups};
    hipLaunchKernelGGL(kern, dim3(batch), dim3(LANES * ldpc_dev::SPLIT_LANE_THREADS), lds, stream, a.llr,
                       reinterpret_cast<uint4*>(a.msg), a.hard, a.result,
                       reinterpret_cast<const ldpc_dev::circ_rec*>(a.circ), a.lay_cnt, a.lay_off, a.steps, a.grp_off,
                       a.step_defer, a.groups, a.beta, a.alpha, a.active, p);
    return hipGetLastError();
}

#define CUSAT_LDPC_SPLIT_INSTANTIATE(C, E, L) \
    template hipError_t launch_ldpc_split_variant<C, E, L>(const ldpc_args&, int, hipStream_t);

}  // namespace cusat
