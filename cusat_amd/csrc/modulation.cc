// Constellations and bit interleaver.
//
// Geometry and label maps are those of EN 302 307-1 clause 5.4 (figures
// 9-12, Table 10) and EN 302 307-2 clause 5.4 (Tables 13, 14, 15); the float
// evaluation order follows modulation/constellation.cc so the points agree
// bit for bit with the reference's transmitter.
#include "modulation.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "_gen/apsk256_inc.h"  // k_apsk256_points_txt

namespace cusat {

namespace {

enum { QPSK, PSK8, APSK8, APSK16, APSK32, APSK64, APSK128, APSK256 };
// rate ids (include/cusat_fec.h)
enum {
    R1_4, R1_3, R2_5, R1_2, R3_5, R2_3, R3_4, R4_5, R5_6, R8_9, R9_10,
    R11_45, R4_15, R13_45, R14_45, R9_20, R7_15, R90_180, R8_15, R96_180, R11_20,
    R100_180, R26_45, R104_180, R18_30, R28_45, R23_36, R116_180, R20_30, R124_180,
    R25_36, R32_45, R128_180, R13_18, R22_30, R132_180, R135_180, R7_9, R140_180,
    R154_180
};

void set(constellation* c, int label, float re, float im)
{
    c->xy[2 * label] = re;
    c->xy[2 * label + 1] = im;
}

// radius * (cos, sin) of the float angle k pi / d, all in single precision
void s2_point(constellation* c, int label, float radius, int k, int d)
{
    const float pi = (float)M_PI;
    const float angle = (float)(k * (double)pi / d);
    set(c, label, radius * std::cos(angle), radius * std::sin(angle));
}

// double-precision polar point rounded to float
void polar(constellation* c, int label, double radius, double angle)
{
    set(c, label, (float)(radius * std::cos(angle)), (float)(radius * std::sin(angle)));
}

float gamma_16apsk(int rate)
{
    switch (rate) {
    case R2_3: return 3.15f;
    case R3_4: return 2.85f;
    case R4_5: return 2.75f;
    case R5_6: return 2.70f;
    case R8_9: return 2.60f;
    case R9_10: return 2.57f;
    case R26_45: return 3.7f;
    case R3_5: return 3.7f;
    case R28_45: return 3.5f;
    case R23_36: return 3.1f;
    case R25_36: return 3.1f;
    case R13_18: return 2.85f;
    case R140_180: return 3.6f;
    case R154_180: return 3.2f;
    case R7_15: return 3.32f;
    case R8_15: return 3.50f;
    default: return 2.85f;
    }
}

struct ring_angle {
    int ring, k, d;
};

void build_16apsk(constellation* c, int rate)
{
    // 4+12APSK, figure 11: label -> (ring, angle k pi / d)
    static const ring_angle L[16] = {
        {1, 1, 4},  {1, -1, 4},  {1, 3, 4},   {1, -3, 4},  {1, 1, 12}, {1, -1, 12},
        {1, 11, 12}, {1, -11, 12}, {1, 5, 12}, {1, -5, 12}, {1, 7, 12}, {1, -7, 12},
        {0, 1, 4},  {0, -1, 4},  {0, 3, 4},   {0, -3, 4},
    };
    const float g = gamma_16apsk(rate);
    const float r1 = std::sqrt(16.0 / (4.0 + 12.0 * g * g));
    const float rad[2] = {r1, g * r1};
    for (int l = 0; l < 16; l++) s2_point(c, l, rad[L[l].ring], L[l].k, L[l].d);
}

void build_32apsk(constellation* c, int rate)
{
    // 4+12+16APSK, figure 12, ring ratios of Table 10
    double g1 = 2.84, g2 = 5.27;
    switch (rate) {
    case R4_5: g1 = 2.72; g2 = 4.87; break;
    case R5_6: g1 = 2.64; g2 = 4.64; break;
    case R8_9: g1 = 2.54; g2 = 4.33; break;
    case R9_10: g1 = 2.53; g2 = 4.30; break;
    default: break;
    }
    const double r1 = std::sqrt(8.0 / (1.0 + 3.0 * g1 * g1 + 4.0 * g2 * g2));
    const float rad[3] = {(float)r1, (float)(g1 * r1), (float)(g2 * r1)};
    static const ring_angle L[32] = {
        {1, 1, 4},  {1, 5, 12},  {1, -1, 4},  {1, -5, 12},  {1, 3, 4},  {1, 7, 12},
        {1, -3, 4}, {1, -7, 12}, {2, 1, 8},   {2, 3, 8},    {2, -1, 4}, {2, -1, 2},
        {2, 3, 4},  {2, 1, 2},   {2, -7, 8},  {2, -5, 8},   {1, 1, 12}, {0, 1, 4},
        {1, -1, 12}, {0, -1, 4}, {1, 11, 12}, {0, 3, 4},    {1, -11, 12}, {0, -3, 4},
        {2, 0, 1},  {2, 1, 4},   {2, -1, 8},  {2, -3, 8},   {2, 7, 8},  {2, 5, 8},
        {2, 1, 1},  {2, -3, 4},
    };
    for (int l = 0; l < 32; l++) s2_point(c, l, rad[L[l].ring], L[l].k, L[l].d);
}

struct ring_row {
    int ring;
    double a[4];  // angle / pi for columns (p, q) = 00, 01, 10, 11
};

bool build_64apsk(constellation* c, int rate)
{
    // EN 302 307-2 Tables 13a-13f: three ring structures.
    // 4+12+20+28 (label = row p q), 8+16+20+20 (label = p X q bbb),
    // 16+16+16+16 (label = RR q p bb)
    static const ring_row R_4_12_20_28[16] = {
        {3, {1.0 / 4, 7.0 / 4, 3.0 / 4, 5.0 / 4}},       {3, {13.0 / 28, 43.0 / 28, 15.0 / 28, 41.0 / 28}},
        {3, {1.0 / 28, 55.0 / 28, 27.0 / 28, 29.0 / 28}}, {0, {1.0 / 4, 7.0 / 4, 3.0 / 4, 5.0 / 4}},
        {3, {9.0 / 28, 47.0 / 28, 19.0 / 28, 37.0 / 28}}, {3, {11.0 / 28, 45.0 / 28, 17.0 / 28, 39.0 / 28}},
        {2, {1.0 / 20, 39.0 / 20, 19.0 / 20, 21.0 / 20}}, {1, {1.0 / 12, 23.0 / 12, 11.0 / 12, 13.0 / 12}},
        {3, {5.0 / 28, 51.0 / 28, 23.0 / 28, 33.0 / 28}}, {2, {9.0 / 20, 31.0 / 20, 11.0 / 20, 29.0 / 20}},
        {3, {3.0 / 28, 53.0 / 28, 25.0 / 28, 31.0 / 28}}, {1, {5.0 / 12, 19.0 / 12, 7.0 / 12, 17.0 / 12}},
        {2, {1.0 / 4, 7.0 / 4, 3.0 / 4, 5.0 / 4}},       {2, {7.0 / 20, 33.0 / 20, 13.0 / 20, 27.0 / 20}},
        {2, {3.0 / 20, 37.0 / 20, 17.0 / 20, 23.0 / 20}}, {1, {1.0 / 4, 7.0 / 4, 3.0 / 4, 5.0 / 4}},
    };
    static const ring_row R_8_16_20_20[16] = {
        {1, {25.0 / 16, 23.0 / 16, 7.0 / 16, 9.0 / 16}},  {3, {7.0 / 4, 5.0 / 4, 1.0 / 4, 3.0 / 4}},
        {1, {27.0 / 16, 21.0 / 16, 5.0 / 16, 11.0 / 16}}, {2, {7.0 / 4, 5.0 / 4, 1.0 / 4, 3.0 / 4}},
        {3, {31.0 / 20, 29.0 / 20, 9.0 / 20, 11.0 / 20}}, {3, {33.0 / 20, 27.0 / 20, 7.0 / 20, 13.0 / 20}},
        {2, {31.0 / 20, 29.0 / 20, 9.0 / 20, 11.0 / 20}}, {2, {33.0 / 20, 27.0 / 20, 7.0 / 20, 13.0 / 20}},
        {0, {13.0 / 8, 11.0 / 8, 3.0 / 8, 5.0 / 8}},      {3, {37.0 / 20, 23.0 / 20, 3.0 / 20, 17.0 / 20}},
        {1, {29.0 / 16, 19.0 / 16, 3.0 / 16, 13.0 / 16}}, {2, {37.0 / 20, 23.0 / 20, 3.0 / 20, 17.0 / 20}},
        {0, {15.0 / 8, 9.0 / 8, 1.0 / 8, 7.0 / 8}},       {3, {39.0 / 20, 21.0 / 20, 1.0 / 20, 19.0 / 20}},
        {1, {31.0 / 16, 17.0 / 16, 1.0 / 16, 15.0 / 16}}, {2, {39.0 / 20, 21.0 / 20, 1.0 / 20, 19.0 / 20}},
    };
    static const ring_row R_16_16_16_16[16] = {
        {0, {1.0 / 16, 31.0 / 16, 15.0 / 16, 17.0 / 16}}, {0, {3.0 / 16, 29.0 / 16, 13.0 / 16, 19.0 / 16}},
        {0, {7.0 / 16, 25.0 / 16, 9.0 / 16, 23.0 / 16}},  {0, {5.0 / 16, 27.0 / 16, 11.0 / 16, 21.0 / 16}},
        {1, {1.0 / 16, 31.0 / 16, 15.0 / 16, 17.0 / 16}}, {1, {3.0 / 16, 29.0 / 16, 13.0 / 16, 19.0 / 16}},
        {1, {7.0 / 16, 25.0 / 16, 9.0 / 16, 23.0 / 16}},  {1, {5.0 / 16, 27.0 / 16, 11.0 / 16, 21.0 / 16}},
        {3, {1.0 / 16, 31.0 / 16, 15.0 / 16, 17.0 / 16}}, {3, {3.0 / 16, 29.0 / 16, 13.0 / 16, 19.0 / 16}},
        {3, {7.0 / 16, 25.0 / 16, 9.0 / 16, 23.0 / 16}},  {3, {5.0 / 16, 27.0 / 16, 11.0 / 16, 21.0 / 16}},
        {2, {1.0 / 16, 31.0 / 16, 15.0 / 16, 17.0 / 16}}, {2, {3.0 / 16, 29.0 / 16, 13.0 / 16, 19.0 / 16}},
        {2, {7.0 / 16, 25.0 / 16, 9.0 / 16, 23.0 / 16}},  {2, {5.0 / 16, 27.0 / 16, 11.0 / 16, 21.0 / 16}},
    };
    int layout = 0;  // 0: 4+12+20+28, 1: 8+16+20+20, 2: 16+16+16+16
    double g1 = 2.4, g2 = 4.3, g3 = 7.0;
    if (rate == R4_5 || rate == R7_9 || rate == R140_180) { layout = 1; g1 = 2.2; g2 = 3.6; g3 = 5.2; }
    else if (rate == R5_6) { layout = 1; g1 = 2.2; g2 = 3.5; g3 = 5.0; }
    else if (rate == R128_180) { layout = 2; g1 = 1.88; g2 = 2.72; g3 = 3.95; }
    static const int COUNT[3][4] = {{4, 12, 20, 28}, {8, 16, 20, 20}, {16, 16, 16, 16}};
    const ring_row* rows = layout == 0 ? R_4_12_20_28 : layout == 1 ? R_8_16_20_20 : R_16_16_16_16;
    const double ratio[4] = {1.0, g1, g2, g3};
    double energy = 0.0;
    for (int i = 0; i < 4; i++) energy += COUNT[layout][i] * ratio[i] * ratio[i];
    const double scale = std::sqrt(64.0 / energy);
    float radius[4];
    for (int i = 0; i < 4; i++) radius[i] = (float)(ratio[i] * scale);
    for (int row = 0; row < 16; row++)
        for (int col = 0; col < 4; col++) {
            const int p = col >> 1, q = col & 1;
            int label;
            if (layout == 0) label = (row << 2) | (p << 1) | q;
            else if (layout == 1) label = (p << 5) | ((row >> 3) << 4) | (q << 3) | (row & 7);
            else label = ((row >> 2) << 4) | (q << 3) | (p << 2) | (row & 3);
            polar(c, label, radius[rows[row].ring], rows[row].a[col] * M_PI);
        }
    return true;
}

void build_8apsk(constellation* c, int rate)
{
    // 2+4+2APSK, EN 302 307-2 Tables 10a/10b
    const double g1 = (rate == R100_180) ? 5.32 : 6.39;
    const double g2 = (rate == R100_180) ? 6.8 : 8.0;
    const double r1 = std::sqrt(8.0 / (2.0 + 4.0 * g1 * g1 + 2.0 * g2 * g2));
    const double radius[8] = {r1, g1 * r1, g1 * r1, g2 * r1, r1, g1 * r1, g1 * r1, g2 * r1};
    const double angle[8] = {1.0, 1.0 + 0.352, 1.0 - 0.352, 1.0, 0.0, -0.352, 0.352, 0.0};
    for (int l = 0; l < 8; l++) polar(c, l, radius[l], angle[l] * M_PI);
}

void build_16apsk_8_8(constellation* c, int rate)
{
    // 8+8APSK: explicit coordinates (Table 11e) for 18/30 and 20/30, else the
    // ring form of Tables 11c/11d
    static const double C18[16][2] = {
        {0.4718, 0.2606},  {0.2606, 0.4718},  {-0.4718, 0.2606},  {-0.2606, 0.4718},
        {0.4718, -0.2606}, {0.2606, -0.4718}, {-0.4718, -0.2606}, {-0.2606, -0.4718},
        {1.2088, 0.4984},  {0.4984, 1.2088},  {-1.2088, 0.4984},  {-0.4984, 1.2088},
        {1.2088, -0.4984}, {0.4984, -1.2088}, {-1.2088, -0.4984}, {-0.4984, -1.2088},
    };
    static const double C20[16][2] = {
        {0.5061, 0.2474},  {0.2474, 0.5061},  {-0.5061, 0.2474},  {-0.2474, 0.5061},
        {0.5061, -0.2474}, {0.2474, -0.5061}, {-0.5061, -0.2474}, {-0.2474, -0.5061},
        {1.2007, 0.4909},  {0.4909, 1.2007},  {-1.2007, 0.4909},  {-0.4909, 1.2007},
        {1.2007, -0.4909}, {0.4909, -1.2007}, {-1.2007, -0.4909}, {-0.4909, -1.2007},
    };
    const double(*xy)[2] = rate == R18_30 ? C18 : rate == R20_30 ? C20 : nullptr;
    if (xy) {
        double energy = 0.0;
        for (int i = 0; i < 16; i++) energy += xy[i][0] * xy[i][0] + xy[i][1] * xy[i][1];
        const double scale = std::sqrt(16.0 / energy);
        for (int i = 0; i < 16; i++) set(c, i, (float)(xy[i][0] * scale), (float)(xy[i][1] * scale));
        return;
    }
    const double gamma = 2.19;
    const double r1 = std::sqrt(2.0 / (1.0 + gamma * gamma));
    static const int NN[2][2][2] = {{{1, 15}, {7, 9}}, {{3, 13}, {5, 11}}};
    for (int l = 0; l < 16; l++) {
        const int ring = (l >> 3) & 1, q = (l >> 2) & 1, pp = (l >> 1) & 1, b0 = l & 1;
        polar(c, l, ring ? gamma * r1 : r1, NN[b0][pp][q] * (M_PI / 8.0));
    }
}

void build_32apsk_rb(constellation* c, int rate, int framesize)
{
    // 4+12+16rbAPSK, EN 302 307-2 Tables 12a/12b; row = (b3, b2, b0), column = (b4, b1)
    double g1 = 2.85, g2 = 5.55;
    if (rate == R32_45) { g1 = 2.839; g2 = 5.258; }
    else if (framesize == 1) { g1 = 2.84; g2 = 5.54; }
    const double r1 = std::sqrt(32.0 / (4.0 + 12.0 * g1 * g1 + 16.0 * g2 * g2));
    const double radius[4] = {0.0, r1, g1 * r1, g2 * r1};
    static const int RING[8] = {3, 3, 2, 2, 3, 3, 2, 1};
    static const double ANG[8][4] = {
        {11.0 / 16, 5.0 / 16, 21.0 / 16, 27.0 / 16}, {9.0 / 16, 7.0 / 16, 23.0 / 16, 25.0 / 16},
        {3.0 / 4, 1.0 / 4, 5.0 / 4, 7.0 / 4},         {7.0 / 12, 5.0 / 12, 17.0 / 12, 19.0 / 12},
        {13.0 / 16, 3.0 / 16, 19.0 / 16, 29.0 / 16}, {15.0 / 16, 1.0 / 16, 17.0 / 16, 31.0 / 16},
        {11.0 / 12, 1.0 / 12, 13.0 / 12, 23.0 / 12}, {3.0 / 4, 1.0 / 4, 5.0 / 4, 7.0 / 4},
    };
    for (int l = 0; l < 32; l++) {
        const int row = (((l >> 3) & 1) << 2) | (((l >> 2) & 1) << 1) | (l & 1);
        const int col = (((l >> 4) & 1) << 1) | ((l >> 1) & 1);
        polar(c, l, radius[RING[row]], ANG[row][col] * M_PI);
    }
}

void build_32apsk_4_8_4_16(constellation* c, int rate)
{
    // EN 302 307-2 Table 12e; row = (b4, b3, b0), column = (b2, b1)
    double g1 = 2.6, g2 = 2.99, g3 = 5.6;
    if (rate == R132_180) g2 = 2.86;
    else if (rate == R7_9 || rate == R140_180) { g1 = 2.8; g2 = 3.08; }
    else if (rate == R32_45) { g1 = 2.84; g2 = 2.84; g3 = 5.26; }
    const double r1 = std::sqrt(32.0 / (4.0 + 8.0 * g1 * g1 + 4.0 * g2 * g2 + 16.0 * g3 * g3));
    const double radius[5] = {0.0, r1, g1 * r1, g2 * r1, g3 * r1};
    static const int RING[8] = {1, 4, 2, 4, 2, 4, 3, 4};
    static const double ANG[8][4] = {
        {1.0 / 4, 7.0 / 4, 3.0 / 4, 5.0 / 4},         {7.0 / 16, 25.0 / 16, 9.0 / 16, 23.0 / 16},
        {1.0 / 12, 23.0 / 12, 11.0 / 12, 13.0 / 12}, {1.0 / 16, 31.0 / 16, 15.0 / 16, 17.0 / 16},
        {5.0 / 12, 19.0 / 12, 7.0 / 12, 17.0 / 12},  {5.0 / 16, 27.0 / 16, 11.0 / 16, 21.0 / 16},
        {1.0 / 4, 7.0 / 4, 3.0 / 4, 5.0 / 4},         {3.0 / 16, 29.0 / 16, 13.0 / 16, 19.0 / 16},
    };
    for (int l = 0; l < 32; l++) {
        const int row = (((l >> 4) & 1) << 2) | (((l >> 3) & 1) << 1) | (l & 1);
        const int col = (((l >> 2) & 1) << 1) | ((l >> 1) & 1);
        polar(c, l, radius[RING[row]], ANG[row][col] * M_PI);
    }
}

void build_128apsk(constellation* c, int rate)
{
    // EN 302 307-2 Tables 14a/14b: 32 rows on six rings, label = q p b4..b0
    static const ring_row ROWS[32] = {
        {0, {83.0 / 1260, 2437.0 / 1260, 1177.0 / 1260, 1343.0 / 1260}},
        {5, {11.0 / 105, 199.0 / 105, 94.0 / 105, 116.0 / 105}},
        {5, {37.0 / 1680, 3323.0 / 1680, 1643.0 / 1680, 1717.0 / 1680}},
        {5, {11.0 / 168, 325.0 / 168, 157.0 / 168, 179.0 / 168}},
        {1, {121.0 / 2520, 4919.0 / 2520, 2399.0 / 2520, 2641.0 / 2520}},
        {2, {23.0 / 280, 537.0 / 280, 257.0 / 280, 303.0 / 280}},
        {4, {19.0 / 720, 1421.0 / 720, 701.0 / 720, 739.0 / 720}},
        {3, {61.0 / 720, 1379.0 / 720, 659.0 / 720, 781.0 / 720}},
        {0, {103.0 / 560, 1017.0 / 560, 457.0 / 560, 663.0 / 560}},
        {5, {61.0 / 420, 779.0 / 420, 359.0 / 420, 481.0 / 420}},
        {5, {383.0 / 1680, 2977.0 / 1680, 1297.0 / 1680, 2063.0 / 1680}},
        {5, {929.0 / 5040, 9151.0 / 5040, 4111.0 / 5040, 5969.0 / 5040}},
        {1, {113.0 / 560, 1007.0 / 560, 447.0 / 560, 673.0 / 560}},
        {2, {169.0 / 1008, 1847.0 / 1008, 839.0 / 1008, 1177.0 / 1008}},
        {4, {563.0 / 2520, 4477.0 / 2520, 1957.0 / 2520, 3083.0 / 2520}},
        {3, {139.0 / 840, 1541.0 / 840, 701.0 / 840, 979.0 / 840}},
        {0, {243.0 / 560, 877.0 / 560, 317.0 / 560, 803.0 / 560}},
        {5, {1993.0 / 5040, 8087.0 / 5040, 3047.0 / 5040, 7033.0 / 5040}},
        {5, {43.0 / 90, 137.0 / 90, 47.0 / 90, 133.0 / 90}},
        {5, {73.0 / 168, 263.0 / 168, 95.0 / 168, 241.0 / 168}},
        {1, {1139.0 / 2520, 3901.0 / 2520, 1381.0 / 2520, 3659.0 / 2520}},
        {2, {117.0 / 280, 443.0 / 280, 163.0 / 280, 397.0 / 280}},
        {4, {341.0 / 720, 1099.0 / 720, 379.0 / 720, 1061.0 / 720}},
        {3, {349.0 / 840, 1331.0 / 840, 491.0 / 840, 1189.0 / 840}},
        {0, {177.0 / 560, 943.0 / 560, 383.0 / 560, 737.0 / 560}},
        {5, {1789.0 / 5040, 8291.0 / 5040, 3251.0 / 5040, 6829.0 / 5040}},
        {5, {49.0 / 180, 311.0 / 180, 131.0 / 180, 229.0 / 180}},
        {5, {53.0 / 168, 283.0 / 168, 115.0 / 168, 221.0 / 168}},
        {1, {167.0 / 560, 953.0 / 560, 393.0 / 560, 727.0 / 560}},
        {2, {239.0 / 720, 1201.0 / 720, 481.0 / 720, 959.0 / 720}},
        {4, {199.0 / 720, 1241.0 / 720, 521.0 / 720, 919.0 / 720}},
        {3, {281.0 / 840, 1399.0 / 840, 559.0 / 840, 1121.0 / 840}},
    };
    const double ratio[6] = {1.0, 1.715, 2.118, 2.681, 2.75, rate == R135_180 ? 3.819 : 3.733};
    const int counts[6] = {16, 16, 16, 16, 16, 48};
    double energy = 0;
    for (int i = 0; i < 6; i++) energy += counts[i] * ratio[i] * ratio[i];
    const double scale = std::sqrt(128.0 / energy);
    const double PI = 3.14159265358979323846;
    for (int row = 0; row < 32; row++) {
        const float radius = (float)(ratio[ROWS[row].ring] * scale);
        for (int col = 0; col < 4; col++) {
            const int p = col >> 1, q = col & 1;
            polar(c, (q << 6) | (p << 5) | row, radius, ROWS[row].a[col] * PI);
        }
    }
}

// The 256 (I, Q) pairs of an explicit-coordinate 256APSK table
// (cusat_amd/data/apsk256_points.txt, EN 302 307-2 Table 15d), label order.
bool apsk256_table(const char* code_id, float (*xy)[2])
{
    char tag[32];
    std::snprintf(tag, sizeof tag, "@points %s\n", code_id);
    const char* p = std::strstr(k_apsk256_points_txt, tag);
    if (!p) return false;
    p += std::strlen(tag);
    for (int i = 0; i < 256; i++) {
        char* end = nullptr;
        xy[i][0] = std::strtof(p, &end);
        if (end == p) return false;
        p = end;
        xy[i][1] = std::strtof(p, &end);
        if (end == p) return false;
        p = end;
    }
    return true;
}

bool build_256apsk(constellation* c, int rate)
{
    // Explicit coordinates (Table 15d) for code ids 20/30 and 22/30 (the
    // latter also for 132/180, as the reference: constellation.cc:400-414),
    // scaled to unit mean energy (energy summed in double, one float scale).
    if (rate == R20_30 || rate == R22_30 || rate == R132_180) {
        float xy[256][2];
        if (!apsk256_table(rate == R20_30 ? "20_30" : "22_30", xy)) return false;
        double energy = 0.0;
        for (int i = 0; i < 256; i++) energy += (double)xy[i][0] * xy[i][0] + (double)xy[i][1] * xy[i][1];
        const float scale = (float)std::sqrt(256.0 / energy);
        for (int i = 0; i < 256; i++) set(c, i, xy[i][0] * scale, xy[i][1] * scale);
        return true;
    }
    // Ring form (EN 302 307-2 Tables 15a-15c): eight rings of 32 points,
    // label = XXX q p aaa, angle (2k - 1) pi / 32 mirrored by (p, q).
    double g[7] = {1.791, 2.405, 2.980, 3.569, 4.235, 5.078, 6.536};
    if (rate == R128_180 || rate == R135_180) {
        const double common[5] = {1.794, 2.409, 2.986, 3.579, 4.045};
        for (int i = 0; i < 5; i++) g[i] = common[i];
        g[5] = rate == R128_180 ? 4.6 : 4.5;
        g[6] = rate == R128_180 ? 5.4 : 5.2;
    }
    const double ratio[8] = {1.0, g[0], g[1], g[2], g[3], g[4], g[5], g[6]};
    double energy = 0.0;
    for (int i = 0; i < 8; i++) energy += 32.0 * ratio[i] * ratio[i];
    const double scale = std::sqrt(256.0 / energy);
    static const int RING_OF_XXX[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    static const int K_OF_AAA[8] = {1, 2, 4, 3, 8, 7, 5, 6};
    const double pi = M_PI;
    for (int xxx = 0; xxx < 8; xxx++) {
        const float radius = (float)(ratio[RING_OF_XXX[xxx]] * scale);
        for (int q = 0; q < 2; q++)
            for (int p = 0; p < 2; p++)
                for (int aaa = 0; aaa < 8; aaa++) {
                    const double phi = (2 * K_OF_AAA[aaa] - 1) * pi / 32.0;
                    const double ang = p ? (q ? pi + phi : pi - phi) : (q ? -phi : phi);
                    polar(c, (xxx << 5) | (q << 4) | (p << 3) | aaa, radius, ang);
                }
    }
    return true;
}

bool is_16apsk_8_8(int r)
{
    return r == R90_180 || r == R96_180 || r == R100_180 || r == R18_30 || r == R20_30;
}
bool is_32apsk_other(int r)
{
    return r == R2_3 || r == R32_45 || r == R128_180 || r == R132_180 || r == R22_30 ||
           r == R140_180 || r == R7_9;
}

const char* column_order(int m, int r)
{
    switch (m) {
    case PSK8:
        if (r == R3_5) return "210";
        if (r == R25_36 || r == R13_18 || r == R7_15 || r == R8_15 || r == R26_45) return "102";
        return "012";
    case APSK8: return "012";
    case APSK16:
        switch (r) {
        case R26_45: return "3201";
        case R3_5: return "3210";
        case R28_45: return "3012";
        case R23_36: return "3021";
        case R25_36: return "2310";
        case R13_18: return "3021";
        case R7_9:
        case R140_180: return "3210";
        case R154_180: return "0321";
        case R90_180: return "3210";
        case R96_180: return "2310";
        case R100_180: return "2301";
        default: return "0123";
        }
    case APSK32:
        if (r == R2_3 || r == R32_45) return "21430";
        if (r == R128_180 || r == R132_180 || r == R22_30) return "40312";
        if (r == R7_9 || r == R140_180) return "40213";
        return "01234";
    case APSK64:
        if (r == R7_9 || r == R140_180) return "201543";
        if (r == R4_5) return "124053";
        if (r == R5_6) return "421053";
        if (r == R128_180) return "305214";
        return "520143";
    case APSK128: return (r == R7_9 || r == R140_180) ? "4130256" : "4250316";
    case APSK256:
        switch (r) {
        case R135_180: return "50743612";
        case R128_180: return "75642301";
        case R116_180: return "40372156";
        case R124_180: return "46320571";
        default: return "01234567";
        }
    default: return nullptr;
    }
}

}  // namespace

bool make_constellation(int modulation, int rate, int framesize, constellation* c)
{
    *c = constellation();
    switch (modulation) {
    case QPSK: {
        c->bits_per_sym = 2;
        const float a = (float)M_SQRT1_2;
        for (int l = 0; l < 4; l++) set(c, l, (l & 2) ? -a : a, (l & 1) ? -a : a);
        break;
    }
    case PSK8: {
        c->bits_per_sym = 3;
        const float a = (float)M_SQRT1_2;
        const float P[8][2] = {{a, a}, {1, 0}, {-1, 0}, {-a, -a}, {0, 1}, {a, -a}, {-a, a}, {0, -1}};
        for (int l = 0; l < 8; l++) set(c, l, P[l][0], P[l][1]);
        break;
    }
    case APSK8:
        c->bits_per_sym = 3;
        build_8apsk(c, rate);
        break;
    case APSK16:
        c->bits_per_sym = 4;
        if (is_16apsk_8_8(rate)) build_16apsk_8_8(c, rate);
        else build_16apsk(c, rate);
        break;
    case APSK32:
        c->bits_per_sym = 5;
        if (rate == R2_3 || rate == R32_45) build_32apsk_rb(c, rate, framesize);
        else if (is_32apsk_other(rate)) build_32apsk_4_8_4_16(c, rate);
        else build_32apsk(c, rate);
        break;
    case APSK64:
        c->bits_per_sym = 6;
        if (!build_64apsk(c, rate)) return false;
        break;
    case APSK128:
        c->bits_per_sym = 7;
        build_128apsk(c, rate);
        break;
    case APSK256:
        c->bits_per_sym = 8;
        if (!build_256apsk(c, rate)) return false;
        break;
    default:
        return false;
    }
    c->n_points = 1 << c->bits_per_sym;
    return true;
}

bool make_interleaver(int modulation, int rate, int framesize, interleaver* il)
{
    *il = interleaver();
    if (modulation == QPSK) return true;
    const char* order = column_order(modulation, rate);
    if (!order) return false;
    const unsigned n_fec = framesize == 0 ? 64800 : framesize == 1 ? 16200 : 32400;
    il->n_cols = (int)strlen(order);
    const unsigned stuffed = modulation == APSK128 ? n_fec + 6 : n_fec;
    il->rows = (int)(stuffed / il->n_cols);
    for (int c = 0; c < il->n_cols; c++) il->col[c] = (uint32_t)il->rows * (uint32_t)(order[c] - '0');
    return true;
}

void xfec_lengths(int modulation, int n_ldpc, int n_mod, int* xfec_len, int* demap_len)
{
    if (modulation == APSK128) {
        *demap_len = (n_ldpc + 6) / n_mod;
        *xfec_len = *demap_len + 12;
    } else {
        *xfec_len = *demap_len = n_ldpc / n_mod;
    }
}

float llr_dmax(int modulation, const constellation& c)
{
    if (modulation == QPSK) return 2.0f;
    if (modulation == PSK8) return 4.0f;
    float dmax = 0.0f;
    const int n = c.n_points, m = c.bits_per_sym;
    for (int i = 0; i < n; i++)
        for (int b = 0; b < m; b++) {
            const int mask = 1 << (m - 1 - b);
            float dmin = 1e30f;
            for (int j = 0; j < n; j++) {
                if (((i ^ j) & mask) == 0) continue;
                const float dr = c.xy[2 * i] - c.xy[2 * j];
                const float di = c.xy[2 * i + 1] - c.xy[2 * j + 1];
                const float d2r = dr * dr;
                const float d2i = di * di;
                const float d = d2r + d2i;
                if (d < dmin) dmin = d;
            }
            if (dmin > dmax) dmax = dmin;
        }
    return dmax;
}

}  // namespace cusat
