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

#define CUSAT_AS_CONST __attribute__((address_space(4)))

// LLRs of one symbol of an APSK table of 2^NM points (pts: the
// constellation, read through scalar loads, eight points per load):
// two-pass log-MAP.  Pass 1: per bit b the smallest squared distance to a
// point with bit b = 0 (mn0) and = 1 (mn1).  Pass 2, in point order: for
// each bit the sum over the points of that bit value of exp(-dl),
// dl = (d - min) g/4, over the points with dl < 16.  LLR = (mn1 - mn0) g/2
// + 2 (log sum0 - log sum1).
//
// Pass 1 takes the points eight at a time.  Within a group the top NM - 3
// label bits are fixed, so those bits' minima take the group's minimum once,
// and the three low bits take minima of four; a minimum is exact in any
// order (the distances are never NaN unless the symbol is, and then every
// minimum keeps its start value either way).
//
// Pass 2 skips a point, for the whole wave, when no lane is near it: a
// lane's bound for a group is the largest minimum over the fixed bits'
// values and both values of the three low bits, + 20 / (g/4); a point
// beyond it has dl >= 20 (1 - a few ulps) > 16 for every bit and adds
// nothing.  COHERENT (the sorted path, where a wave's symbols are
// neighbours): a bit whose minimum is the lane's overall minimum d0 (the
// bit's value at the nearest point) has dl = (d - d0) g/4, the same number
// for all such bits, so its exp is computed once per point; the other bits
// take their own exp, skipped when no lane of the wave needs one.  Inside,
// the per-bit update is branch-free: a term left out adds +0.0f, an
// identity here (the sums start at +0 and every term is > 0).  The values
// therefore equal the plain two-pass loop over every point.
template <int NM, bool COHERENT>
__device__ __forceinline__ void demap_apsk_symbol(CUSAT_AS_CONST const float* pts, float re, float im, float gain,
                                                  const uint32_t* col, int N, int j, int16_t* __restrict__ llr)
{
    static_assert(NM >= 3, "groups of eight points");
    constexpr int NP = 1 << NM;
    constexpr int TOP = NM - 3;  // label bits fixed within a group of eight
    float mn0[NM], mn1[NM], s0[NM], s1[NM];
#pragma unroll
    for (int b = 0; b < NM; b++) {
        mn0[b] = 1e30f;
        mn1[b] = 1e30f;
        s0[b] = 0.0f;
        s1[b] = 0.0f;
    }
    // The sorted path unrolls pass 1 whole: each group's fixed bits are then
    // known when it is compiled (no select between the two minima of a bit)
    // and no running minimum crosses a loop edge, where fminf would have to
    // quiet it first (one v_max per minimum per iteration).  256APSK:
    // 9.2 k -> 5.1 k VALU cycles per wave for pass 1.  The unsorted path
    // (16/32APSK) keeps two groups per iteration, which measured faster there.
#pragma unroll(COHERENT ? NP / 8 : 2)
    for (int i0 = 0; i0 < NP; i0 += 8) {
        float d[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float dr = re - pts[2 * (i0 + k)], di = im - pts[2 * (i0 + k) + 1];
            d[k] = fmaf(dr, dr, di * di);
        }
        // label bit 2 (b = NM - 3): points 0-3 / 4-7; bit 1: {0,1,4,5} /
        // {2,3,6,7}; bit 0 (b = NM - 1): even / odd
        const float p01 = fminf(d[0], d[1]), p23 = fminf(d[2], d[3]);
        const float p45 = fminf(d[4], d[5]), p67 = fminf(d[6], d[7]);
        const float lo2 = fminf(p01, p23), hi2 = fminf(p45, p67);
        const float lo1 = fminf(p01, p45), hi1 = fminf(p23, p67);
        const float lo0 = fminf(fminf(d[0], d[2]), fminf(d[4], d[6]));
        const float hi0 = fminf(fminf(d[1], d[3]), fminf(d[5], d[7]));
        const float g = fminf(lo2, hi2);
        mn0[NM - 3] = fminf(mn0[NM - 3], lo2);
        mn1[NM - 3] = fminf(mn1[NM - 3], hi2);
        mn0[NM - 2] = fminf(mn0[NM - 2], lo1);
        mn1[NM - 2] = fminf(mn1[NM - 2], hi1);
        mn0[NM - 1] = fminf(mn0[NM - 1], lo0);
        mn1[NM - 1] = fminf(mn1[NM - 1], hi0);
#pragma unroll
        for (int b = 0; b < TOP; b++) {
            if ((i0 >> (NM - 1 - b)) & 1) mn1[b] = fminf(mn1[b], g);
            else                          mn0[b] = fminf(mn0[b], g);
        }
    }
    const float quarter = gain * 0.25f, half = gain * 0.5f;
    const float margin = 20.0f / quarter;
    const float d0 = fminf(mn0[NM - 1], mn1[NM - 1]);
    float low_max = 0.0f;
#pragma unroll
    for (int b = TOP; b < NM; b++) low_max = fmaxf(low_max, fmaxf(mn0[b], mn1[b]));
    if constexpr (COHERENT) {
        // per (bit, value): a point at d >= min + margin adds nothing to
        // that bit's sum (dl >= 20 (1 - a few ulps) > 16), for any lane
        float th0[NM], th1[NM];
#pragma unroll
        for (int b = 0; b < NM; b++) {
            th0[b] = mn0[b] + margin;
            th1[b] = mn1[b] + margin;
        }
        // two groups per iteration: the lowest fixed bit is known when the
        // loop is compiled, and the second group reuses the first's
        // selections of the bits above it
#pragma unroll 2
        for (int i0 = 0; i0 < NP; i0 += 8) {
            float tg = low_max;
#pragma unroll
            for (int b = 0; b < TOP; b++) tg = fmaxf(tg, ((i0 >> (NM - 1 - b)) & 1) ? mn1[b] : mn0[b]);
            tg += margin;
            float px[8], py[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                px[k] = pts[2 * (i0 + k)];
                py[k] = pts[2 * (i0 + k) + 1];
            }
            // the group's fixed bits: their minimum, threshold and running
            // sum (the sum of that bit value), written back after the group
            float mt[TOP > 0 ? TOP : 1], tt[TOP > 0 ? TOP : 1], at[TOP > 0 ? TOP : 1];
#pragma unroll
            for (int b = 0; b < TOP; b++) {
                const bool pol = (i0 >> (NM - 1 - b)) & 1;
                mt[b] = pol ? mn1[b] : mn0[b];
                tt[b] = pol ? th1[b] : th0[b];
                at[b] = pol ? s1[b] : s0[b];
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float dr = re - px[k], di = im - py[k];
                const float d = fmaf(dr, dr, di * di);
                if (d < tg) {
                    const float dl0 = (d - d0) * quarter;
                    float e0 = 0.0f;
                    if (__builtin_amdgcn_ballot_w64(dl0 < 16.0f)) e0 = dl0 < 16.0f ? exp_neg(-dl0) : 0.0f;
#pragma unroll
                    for (int b = 0; b < NM; b++) {
                        const bool lpol = (k >> (NM - 1 - b)) & 1;  // low bits: fixed by k
                        const float m = b < TOP ? mt[b < TOP ? b : 0] : (lpol ? mn1[b] : mn0[b]);
                        const float thr = b < TOP ? tt[b < TOP ? b : 0] : (lpol ? th1[b] : th0[b]);
                        if (__builtin_amdgcn_ballot_w64(d < thr)) {
                            // A lane with dl == dl0 (every lane whose minimum
                            // for this bit value is d0) has exp(-dl) == e0
                            // bit for bit; the exp runs, for every lane, only
                            // when some lane in range has another dl.
                            const float dl = (d - m) * quarter;
                            const bool in = dl < 16.0f;
                            float ev = e0;
                            if (__builtin_amdgcn_ballot_w64(in && dl != dl0)) ev = exp_neg(-dl);
                            const float e = in ? ev : 0.0f;
                            if (b < TOP)   at[b < TOP ? b : 0] += e;
                            else if (lpol) s1[b] += e;
                            else           s0[b] += e;
                        }
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < TOP; b++) {
                if ((i0 >> (NM - 1 - b)) & 1) s1[b] = at[b];
                else                          s0[b] = at[b];
            }
        }
    } else {
        for (int i0 = 0; i0 < NP; i0 += 8) {
            float tg = low_max;
#pragma unroll
            for (int b = 0; b < TOP; b++) tg = fmaxf(tg, ((i0 >> (NM - 1 - b)) & 1) ? mn1[b] : mn0[b]);
            tg += margin;
            float px[8], py[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                px[k] = pts[2 * (i0 + k)];
                py[k] = pts[2 * (i0 + k) + 1];
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int i = i0 + k;
                const float dr = re - px[k], di = im - py[k];
                const float d = fmaf(dr, dr, di * di);
                if (d < tg) {
#pragma unroll
                    for (int b = 0; b < NM; b++) {
                        const bool pol = (i >> (NM - 1 - b)) & 1;
                        const float dl = (d - (pol ? mn1[b] : mn0[b])) * quarter;
                        const float e = dl < 16.0f ? exp_neg(-dl) : 0.0f;
                        s1[b] = pol ? s1[b] + e : s1[b];
                        s0[b] = pol ? s0[b] : s0[b] + e;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int b = 0; b < NM; b++) {
        const float corr = log_pos(fmaxf(s0[b], 1e-30f)) - log_pos(fmaxf(s1[b], 1e-30f));
        const float v = fmaf(2.0f, corr, (mn1[b] - mn0[b]) * half);
        const uint32_t pos = col[b] + (uint32_t)j;
        if (pos < (uint32_t)N) llr[pos] = to_llr(v);
    }
}

template <bool COHERENT>
__device__ __forceinline__ void demap_apsk_any(int n_mod, const float* pts_g, float re, float im, float gain,
                                               const uint32_t* col, int N, int j, int16_t* llr)
{
    const auto* pts = (CUSAT_AS_CONST const float*)pts_g;
    switch (n_mod) {
    // the sorted path takes tables of 64 points and more
    case 3: if constexpr (!COHERENT) demap_apsk_symbol<3, false>(pts, re, im, gain, col, N, j, llr); break;
    case 4: if constexpr (!COHERENT) demap_apsk_symbol<4, false>(pts, re, im, gain, col, N, j, llr); break;
    case 5: if constexpr (!COHERENT) demap_apsk_symbol<5, false>(pts, re, im, gain, col, N, j, llr); break;
    case 6: demap_apsk_symbol<6, COHERENT>(pts, re, im, gain, col, N, j, llr); break;
    case 7: demap_apsk_symbol<7, COHERENT>(pts, re, im, gain, col, N, j, llr); break;
    case 8: demap_apsk_symbol<8, COHERENT>(pts, re, im, gain, col, N, j, llr); break;
    default: break;
    }
}

// LLRs of symbol j of one frame (sym = the frame's symbols, llr = its N
// LLRs): the QPSK and 8PSK closed forms, else the APSK table a.pts.
__device__ __forceinline__ void demap_symbol(const demap_args& a, int j, const float2* sym, int16_t* llr)
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
        demap_apsk_any<false>(a.n_mod, a.pts, re, im, a.gain, a.col, a.N, j, llr);
    }
}

__global__ void __launch_bounds__(256) demap_kernel(demap_args a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int cw = blockIdx.y;
    if (j >= a.dlen) return;
    demap_symbol(a, j, reinterpret_cast<const float2*>(a.sym) + (size_t)cw * a.xfec, a.llr + (size_t)cw * a.N);
}

// ---- sorted APSK demap (constellations of 64 points and more) ----
//
// demap_apsk_symbol skips a point for a wave only when no lane is near
// it, so it pays off when a wave's 64 symbols lie close together: one
// workgroup per frame sorts the frame's symbols by a 32 x 32 grid cell
// (Morton order, counting sort in the LDS) and each wave takes 64
// consecutive sorted symbols.  Which lane demaps which symbol changes only
// the speed; the values are demap_kernel's.

constexpr int SORT_THREADS = DEMAP_BLOCK;
constexpr int GRID_CELLS = 1024;  // 32 x 32

__device__ __forceinline__ unsigned spread5(unsigned v)
{
    v &= 31u;
    v = (v | (v << 4)) & 0x10Fu;
    v = (v | (v << 2)) & 0x133u;
    v = (v | (v << 1)) & 0x155u;
    return v;
}

// grid cell of a symbol (Morton order of the 32 x 32 cells); NaN and
// symbols outside the grid land in the border cells
__device__ __forceinline__ unsigned grid_cell(float2 y, float scale)
{
    const float qx = fminf(fmaxf(fmaf(y.x, scale, 16.0f), 0.0f), 31.0f);
    const float qy = fminf(fmaxf(fmaf(y.y, scale, 16.0f), 0.0f), 31.0f);
    return spread5((unsigned)qx) | (spread5((unsigned)qy) << 1);
}

// LDS of the sorted path: cell counts / bases, the sorted symbol indices
struct sort_lds {
    int cell[GRID_CELLS];
    uint16_t order[SORTED_DEMAP_MAX_SYMBOLS];
    int wave_sum[SORT_THREADS / 64];
    int range[2];  // sorted positions of this workgroup's part
};

// Whole workgroup: part `part` of `parts` of one frame of n symbols y (LLRs
// to llr), constellation pts in global memory, n_mod bits per symbol.  The
// workgroup sorts the whole frame and demaps the cells that start in its
// share of the sorted positions.  The order of the symbols inside a cell
// depends on the LDS atomics' timing, so the parts split at cell starts:
// which symbols a cell holds, and where it starts, are the same in every
// workgroup that sorts the frame.
__device__ __forceinline__ void demap_frame_sorted(sort_lds& L, int n_mod, const float* pts,
                                                   const float2* __restrict__ y, int n, int N, float gain,
                                                   float scale, const uint32_t* col, int16_t* __restrict__ llr,
                                                   int part, int parts)
{
    constexpr int PER = GRID_CELLS / SORT_THREADS;  // cells per thread in the scan
    static_assert(PER * SORT_THREADS == GRID_CELLS, "whole cells per thread");
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < PER; k++) L.cell[PER * t + k] = 0;
    if (t < 2) L.range[t] = n;
    __syncthreads();
    for (int j = t; j < n; j += SORT_THREADS) atomicAdd(&L.cell[grid_cell(y[j], scale)], 1);
    __syncthreads();
    // exclusive scan of the counts: PER per thread, within each wave, then
    // over the waves
    int v[PER], run = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        v[k] = L.cell[PER * t + k];
        run += v[k];
    }
    const int lane = t & 63, w = t >> 6;
    int inc = run;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) L.wave_sum[w] = inc;
    __syncthreads();
    int base = inc - run;
    for (int k = 0; k < w; k++) base += L.wave_sum[k];
    const int share = (n + parts - 1) / parts;
    const int want_lo = part * share, want_hi = part == parts - 1 ? n + 1 : (part + 1) * share;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        L.cell[PER * t + k] = base;
        // the part runs from the first cell start at or after its share's
        // start to the first at or after its end (cell 0 starts at 0)
        if (base >= want_lo) atomicMin(&L.range[0], base);
        if (base >= want_hi) atomicMin(&L.range[1], base);
        base += v[k];
    }
    __syncthreads();
    for (int j = t; j < n; j += SORT_THREADS) {
        const int p = atomicAdd(&L.cell[grid_cell(y[j], scale)], 1);
        L.order[p] = (uint16_t)j;
    }
    __syncthreads();
    const int p_lo = L.range[0], p_hi = L.range[1];
    for (int p = p_lo + t; p < p_hi; p += SORT_THREADS) {
        const int j = L.order[p];
        const float2 s = y[j];
        demap_apsk_any<true>(n_mod, pts, s.x, s.y, gain, col, N, j, llr);
    }
    __syncthreads();  // the LDS is reused by the next frame
}

// cusat_fec_demap of a constellation of the sorted path: `parts` workgroups
// per frame
__global__ void __launch_bounds__(SORT_THREADS) demap_sorted_kernel(demap_args a, int parts)
{
    __shared__ sort_lds L;
    const int f = blockIdx.x / parts;
    demap_frame_sorted(L, a.n_mod, a.pts, reinterpret_cast<const float2*>(a.sym) + (size_t)f * a.xfec, a.dlen, a.N,
                       a.gain, a.grid_scale, a.col, a.llr + (size_t)f * a.N, (int)blockIdx.x - f * parts, parts);
}

// Mixed-MODCOD batch: a fixed grid takes work items from a queue in device
// memory (hdr[8], reset by the last workgroup out through hdr[9], so a
// captured launch replays any plan).  Items 0 .. hdr[3] - 1 are the parts
// of the sorted-path frames in dlist order, one workgroup each; the next
// hdr[1] are blocks of rx_code::span symbols of the other frames (the
// entry k of dlist at or after hdr[3] for frame dlist[k].x owns blocks
// dlist[k].y .. dlist[k + 1].y - 1).  Both are in descending cost, so the queue hands
// out the heavy work first and the tail is cheap.  Every symbol gets
// exactly the arithmetic of demap_kernel.
__global__ void __launch_bounds__(SORT_THREADS)
    demap_mixed_kernel(const rx_code* __restrict__ codes, const rx_rec* __restrict__ recs,
                       const int2* __restrict__ dlist, const rx_scale* __restrict__ snap, int* hdr,
                       const float* __restrict__ sym, int16_t* __restrict__ llr)
{
    __shared__ sort_lds L;
    __shared__ int item_s;
    const int nh = hdr[3], n_dl = hdr[4];
    const int total = nh + hdr[1];
    unsigned* qc = reinterpret_cast<unsigned*>(hdr + 8);
    int k_lo = nh;  // a workgroup's block items only grow: their frame only moves forward
    for (;;) {
        if (threadIdx.x == 0) item_s = (int)atomicAdd(&qc[0], 1u);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();
        if (item >= total) break;
        if (item < nh) {
            const int2 e = dlist[item];
            const int f = e.x;
            const rx_rec& r = recs[f];
            const rx_code& c = codes[r.code];
            const float gain = (r.flags & RX_SCALE_STATE) ? snap[r.code].gain : r.gain;
            uint32_t col[8];
#pragma unroll
            for (int b = 0; b < 8; b++) col[b] = c.col[b];
            demap_frame_sorted(L, c.n_mod, c.pts, reinterpret_cast<const float2*>(sym) + r.sym_off, c.dlen, c.N,
                               gain, c.grid_scale, col, llr + r.llr_off, e.y & 0xff, e.y >> 8);
            continue;
        }
        const int blk = item - nh;
        int lo = k_lo, hi = n_dl - 1;  // last entry with first block <= blk
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (dlist[mid].y <= blk) lo = mid;
            else hi = mid - 1;
        }
        k_lo = lo;
        const int2 e = dlist[lo];
        const rx_rec& r = recs[e.x];
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
        const int j0 = (blk - e.y) * c.span, j1 = min(a.dlen, j0 + c.span);
        for (int j = j0 + (int)threadIdx.x; j < j1; j += DEMAP_BLOCK)
            demap_symbol(a, j, reinterpret_cast<const float2*>(sym) + r.sym_off, llr + r.llr_off);
    }
    // the last workgroup to leave resets the queue for the next launch
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&qc[1], 1u) == gridDim.x - 1) {
            atomicExch(&qc[0], 0u);
            atomicExch(&qc[1], 0u);
        }
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
        // the next chunk's symbol is loaded before this chunk is summed, so
        // its latency hides behind the two lanes' adds
        float2 next = t < m ? y[(size_t)t * stride] : make_float2(0.0f, 0.0f);
        for (int c0 = 0; c0 < m; c0 += 256) {
            const int j = c0 + t;
            const float2 v = next;
            if (j < m) {
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
            if (j + 256 < m) next = y[(size_t)(j + 256) * stride];
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
    if (demap_sorted(a.pts != nullptr, a.n_points, a.dlen)) {
        const int parts = demap_parts(a.dlen, a.n_points);
        hipLaunchKernelGGL(demap_sorted_kernel, dim3(batch * parts), dim3(SORT_THREADS), 0, stream, a, parts);
        return hipGetLastError();
    }
    dim3 grid((a.dlen + 255) / 256, batch);
    hipLaunchKernelGGL(demap_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_demap_mixed(const rx_code* codes, const rx_rec* recs, const int2* dlist, const rx_scale* snap,
                              int* hdr, const float* sym, int16_t* llr, int n_cu, hipStream_t stream)
{
    hipLaunchKernelGGL(demap_mixed_kernel, dim3(n_cu * 4), dim3(SORT_THREADS), 0, stream, codes, recs, dlist, snap,
                       hdr, sym, llr);
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
