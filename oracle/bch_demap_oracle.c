/*
 * CPU oracle: BCH remainder check and soft demapper -- TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * BCH restates lib/cusat_dsp/fec/bch_code.cc:14-19,99-122 (generator as the
 * product of the EN 302 307 Table 6a/6b minimal polynomials) and the
 * acceptance test of lib/cusat_dsp/cuda/bch_common.cuh:91-158
 * (valid = c(x) x^p mod g(x) == 0, which is c(x) mod g(x) == 0 because
 * g(0) = 1), here as a plain bit-serial long division.
 *
 * The demapper restates lib/cusat_dsp/cuda/demapper_kernel.cu:31-184 with
 * an explicit evaluation order (no implicit FMA contraction; build with
 * -ffp-contract=off) that the HIP kernel repeats operation for operation.
 * The APSK log-MAP correction uses this file's own exp/log (Cody-Waite
 * range reduction + Taylor series for exp; frexp reduction + least-squares
 * Chebyshev-node polynomial for log1p), not any vendor fast-math routine.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* ---- BCH ---- */

static const uint32_t MINPOLY[3][12] = {
    /* normal FECFRAME, GF(2^16), EN 302 307-1 Table 6a */
    { 0x1002D, 0x10173, 0x10FBD, 0x15A55, 0x11F2F, 0x1F7B5,
      0x1AF65, 0x17367, 0x10EA1, 0x175A7, 0x13A2D, 0x11AE3 },
    /* short FECFRAME, GF(2^14), Table 6b */
    { 0x402B, 0x4941, 0x4647, 0x5591, 0x6B55, 0x6389,
      0x6CE5, 0x4F21, 0x460F, 0x5A49, 0x5811, 0x65EF },
    /* medium FECFRAME, GF(2^15), EN 302 307-2 */
    { 0x802D, 0x8C93, 0xB555, 0x8D6D, 0x94D7, 0xB0D1,
      0xD8B5, 0xCB55, 0xBAB7, 0xBC9F, 0x8A17, 0xED15 },
};

static int poly_deg(const uint64_t g[4])
{
    for (int i = 255; i >= 0; i--)
        if ((g[i >> 6] >> (i & 63)) & 1) return i;
    return -1;
}

int oc_bch_generator(int framesize, int t, uint64_t g[4])
{
    if (framesize < 0 || framesize > 2 || t < 1 || t > 12) return -1;
    uint64_t a[4] = { 1, 0, 0, 0 };
    for (int i = 0; i < t; i++) {
        uint64_t r[4] = { 0, 0, 0, 0 };
        uint32_t b = MINPOLY[framesize][i];
        for (int s = 0; s < 32; s++) {
            if (!((b >> s) & 1)) continue;
            /* r ^= a << s */
            for (int w = 3; w >= 0; w--) {
                int ws = w - (s >> 6), bs = s & 63;
                uint64_t v = 0;
                if (ws >= 0) v = a[ws] << bs;
                if (bs && ws - 1 >= 0) v |= a[ws - 1] >> (64 - bs);
                r[w] ^= v;
            }
        }
        memcpy(a, r, sizeof a);
    }
    memcpy(g, a, sizeof a);
    return poly_deg(g);
}

int oc_bch_check(const uint8_t* bytes, int nbits, int framesize, int t)
{
    uint64_t g[4];
    int p = oc_bch_generator(framesize, t, g);
    if (p <= 0 || p > 250) return -1;
    /* r holds p+1 bits; shift in one message bit at a time, reduce by g. */
    uint64_t r[4] = { 0, 0, 0, 0 };
    for (int i = 0; i < nbits; i++) {
        int bit = (bytes[i >> 3] >> (7 - (i & 7))) & 1;
        for (int w = 3; w > 0; w--) r[w] = (r[w] << 1) | (r[w - 1] >> 63);
        r[0] = (r[0] << 1) | (uint64_t)bit;
        if ((r[p >> 6] >> (p & 63)) & 1)
            for (int w = 0; w < 4; w++) r[w] ^= g[w];
    }
    return (r[0] | r[1] | r[2] | r[3]) == 0;
}

/* ---- exp / log for the log-MAP correction ---- */

/* e^x for x in (-16, 0]: n = rint(x log2 e), r = x - n ln2 in two steps
 * (Cody-Waite split of ln2), Taylor polynomial of degree 7, scale by 2^n. */
float oc_exp_neg(float x)
{
    const float n = rintf(x * 1.44269502e+00f);
    float r = fmaf(-n, 6.93359375e-01f, x);
    r = fmaf(-n, -2.12194442e-04f, r);
    float p = 1.98412701e-04f;          /* 1/7! */
    p = fmaf(p, r, 1.38888892e-03f);    /* 1/6! */
    p = fmaf(p, r, 8.33333377e-03f);    /* 1/5! */
    p = fmaf(p, r, 4.16666679e-02f);    /* 1/4! */
    p = fmaf(p, r, 1.66666672e-01f);    /* 1/3! */
    p = fmaf(p, r, 5.00000000e-01f);
    p = fmaf(p, r, 1.00000000e+00f);
    p = fmaf(p, r, 1.00000000e+00f);
    return ldexpf(p, (int)n);
}

/* ln x for x >= 1e-30: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln x = e ln2 + f P(f), f = m - 1, P of degree 9 fitted to log1p(f)/f. */
float oc_log_pos(float x)
{
    int e;
    float m = frexpf(x, &e);
    if (m < 7.07106769e-01f) { m = m + m; e -= 1; }
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

static inline int16_t to_llr(float v)
{
    v = rintf(v);
    v = fminf(fmaxf(v, -32768.0f), 32767.0f);
    return (int16_t)v;
}

void oc_demap(const float* sym, int n_sym, int n_mod, const float* pts,
              int n_points, const uint32_t* col, int fec_len, float qpsk_scale,
              float psk8_scale, float gain, int16_t* llr)
{
    for (int j = 0; j < n_sym; j++) {
        const float re = sym[2 * j], im = sym[2 * j + 1];
        if (n_mod == 2) {
            /* QPSK (demapper_kernel.cu:53-75): no interleaving */
            if (2 * j + 1 < fec_len) {
                llr[2 * j] = to_llr(re * qpsk_scale);
                llr[2 * j + 1] = to_llr(im * qpsk_scale);
            }
        } else if (n_mod == 3 && !pts) {
            /* 8PSK (:76-107): derotate by e^{-j pi/8}, closed-form max-log */
            const float C = 9.23879504e-01f, S = 3.82683426e-01f;
            const float cr = fmaf(re, C, im * S);
            const float ci = fmaf(im, C, -(re * S));
            const float l0 = (7.07106769e-01f * (fabsf(cr) - fabsf(ci))) * psk8_scale;
            llr[col[0] + j] = to_llr(l0);
            llr[col[1] + j] = to_llr(cr * psk8_scale);
            llr[col[2] + j] = to_llr(ci * psk8_scale);
        } else {
            /* APSK log-MAP (:108-183) */
            float mn0[8], mn1[8], s0[8], s1[8];
            for (int b = 0; b < 8; b++) { mn0[b] = 1e30f; mn1[b] = 1e30f; s0[b] = 0; s1[b] = 0; }
            for (int i = 0; i < n_points; i++) {
                const float dr = re - pts[2 * i], di = im - pts[2 * i + 1];
                const float d = fmaf(dr, dr, di * di);
                for (int b = 0; b < n_mod; b++) {
                    if (i & (1 << (n_mod - 1 - b))) { if (d < mn1[b]) mn1[b] = d; }
                    else                            { if (d < mn0[b]) mn0[b] = d; }
                }
            }
            const float quarter = gain * 0.25f, half = gain * 0.5f;
            for (int i = 0; i < n_points; i++) {
                const float dr = re - pts[2 * i], di = im - pts[2 * i + 1];
                const float d = fmaf(dr, dr, di * di);
                for (int b = 0; b < n_mod; b++) {
                    if (i & (1 << (n_mod - 1 - b))) {
                        const float dl = (d - mn1[b]) * quarter;
                        if (dl < 16.0f) s1[b] += oc_exp_neg(-dl);
                    } else {
                        const float dl = (d - mn0[b]) * quarter;
                        if (dl < 16.0f) s0[b] += oc_exp_neg(-dl);
                    }
                }
            }
            for (int b = 0; b < n_mod; b++) {
                const float corr = oc_log_pos(fmaxf(s0[b], 1e-30f)) - oc_log_pos(fmaxf(s1[b], 1e-30f));
                const float v = fmaf(2.0f, corr, (mn1[b] - mn0[b]) * half);
                const uint32_t pos = col[b] + (uint32_t)j;
                if (pos < (uint32_t)fec_len) llr[pos] = to_llr(v);
            }
        }
    }
}
