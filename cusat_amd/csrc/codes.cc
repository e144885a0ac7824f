// Code data and decoder geometry.
//
// The LDPC accumulator addresses and the parallel-layer schedules are the
// reference's data, embedded from cusat_amd/data/*.txt at build time (the
// table text is the checksum-canonical form of
// scripts/check_ldpc_tables.py:87-95).  The (k_bch, n_bch, t) table and the
// (frame, rate) -> table map restate fec/code_params.cc:17-84 and
// ldpc_table_factory.cc:17-95.
#include "codes.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>

#include "_gen/data_inc.h"  // k_ldpc_tables_txt, k_ldpc_schedules_txt

namespace cusat {

namespace {

struct fec_row {
    int rate, k, n, t;
};

// rate ids follow include/cusat_fec.h (= code_rate_t of modcod_types.h)
enum {
    R1_4, R1_3, R2_5, R1_2, R3_5, R2_3, R3_4, R4_5, R5_6, R8_9, R9_10,
    R11_45, R4_15, R13_45, R14_45, R9_20, R7_15, R90_180, R8_15, R96_180, R11_20,
    R100_180, R26_45, R104_180, R18_30, R28_45, R23_36, R116_180, R20_30, R124_180,
    R25_36, R32_45, R128_180, R13_18, R22_30, R132_180, R135_180, R7_9, R140_180,
    R154_180
};

const fec_row NORMAL[] = {
    {R1_4, 16008, 16200, 12},   {R1_3, 21408, 21600, 12},   {R2_5, 25728, 25920, 12},
    {R1_2, 32208, 32400, 12},   {R3_5, 38688, 38880, 12},   {R2_3, 43040, 43200, 10},
    {R3_4, 48408, 48600, 12},   {R4_5, 51648, 51840, 12},   {R5_6, 53840, 54000, 10},
    {R8_9, 57472, 57600, 8},    {R9_10, 58192, 58320, 8},   {R13_45, 18528, 18720, 12},
    {R9_20, 28968, 29160, 12},  {R90_180, 32208, 32400, 12}, {R96_180, 34368, 34560, 12},
    {R11_20, 35448, 35640, 12}, {R100_180, 35808, 36000, 12}, {R104_180, 37248, 37440, 12},
    {R26_45, 37248, 37440, 12}, {R18_30, 38688, 38880, 12}, {R28_45, 40128, 40320, 12},
    {R23_36, 41208, 41400, 12}, {R116_180, 41568, 41760, 12}, {R20_30, 43008, 43200, 12},
    {R124_180, 44448, 44640, 12}, {R25_36, 44808, 45000, 12}, {R128_180, 45888, 46080, 12},
    {R13_18, 46608, 46800, 12}, {R132_180, 47328, 47520, 12}, {R22_30, 47328, 47520, 12},
    {R135_180, 48408, 48600, 12}, {R140_180, 50208, 50400, 12}, {R7_9, 50208, 50400, 12},
    {R154_180, 55248, 55440, 12},
};
const fec_row SHORT[] = {
    {R1_4, 3072, 3240, 12},   {R1_3, 5232, 5400, 12},   {R2_5, 6312, 6480, 12},
    {R1_2, 7032, 7200, 12},   {R3_5, 9552, 9720, 12},   {R2_3, 10632, 10800, 12},
    {R3_4, 11712, 11880, 12}, {R4_5, 12432, 12600, 12}, {R5_6, 13152, 13320, 12},
    {R8_9, 14232, 14400, 12}, {R11_45, 3792, 3960, 12}, {R4_15, 4152, 4320, 12},
    {R14_45, 4872, 5040, 12}, {R7_15, 7392, 7560, 12},  {R8_15, 8472, 8640, 12},
    {R26_45, 9192, 9360, 12}, {R32_45, 11352, 11520, 12},
};

struct table_map {
    int rate;
    const char* name;
};
const table_map NORMAL_TABLES[] = {
    {R1_4, "s2_b1"}, {R1_3, "s2_b2"}, {R2_5, "s2_b3"}, {R1_2, "s2_b4"}, {R3_5, "s2_b5"},
    {R2_3, "s2_b6"}, {R3_4, "s2_b7"}, {R4_5, "s2_b8"}, {R5_6, "s2_b9"}, {R8_9, "s2_b10"},
    {R9_10, "s2_b11"}, {R13_45, "s2x_b2"}, {R9_20, "s2x_b3"}, {R90_180, "s2x_b11"},
    {R96_180, "s2x_b12"}, {R11_20, "s2x_b4"}, {R100_180, "s2x_b13"}, {R104_180, "s2x_b14"},
    {R26_45, "s2x_b5"}, {R18_30, "s2x_b22"}, {R28_45, "s2x_b6"}, {R23_36, "s2x_b7"},
    {R116_180, "s2x_b15"}, {R20_30, "s2x_b23"}, {R124_180, "s2x_b16"}, {R25_36, "s2x_b8"},
    {R128_180, "s2x_b17"}, {R13_18, "s2x_b9"}, {R132_180, "s2x_b18"}, {R22_30, "s2x_b24"},
    {R135_180, "s2x_b19"}, {R140_180, "s2x_b20"}, {R7_9, "s2x_b10"}, {R154_180, "s2x_b21"},
};
const table_map SHORT_TABLES[] = {
    {R1_4, "s2_c1"}, {R1_3, "s2_c2"}, {R2_5, "s2_c3"}, {R1_2, "s2_c4"}, {R3_5, "s2_c5"},
    {R2_3, "s2_c6"}, {R3_4, "s2_c7"}, {R4_5, "s2_c8"}, {R5_6, "s2_c9"}, {R8_9, "s2_c10"},
    {R11_45, "s2x_c1"}, {R4_15, "s2x_c2"}, {R14_45, "s2x_c3"}, {R7_15, "s2x_c4"},
    {R8_15, "s2x_c5"}, {R26_45, "s2x_c6"}, {R32_45, "s2x_c7"},
};

std::once_flag g_parse_once;
std::map<std::string, ldpc_table>* g_tables;
std::map<std::string, std::vector<std::array<int, 2>>>* g_scheds;
std::map<std::string, int>* g_sched_q;

void parse_all()
{
    g_tables = new std::map<std::string, ldpc_table>();
    g_scheds = new std::map<std::string, std::vector<std::array<int, 2>>>();
    g_sched_q = new std::map<std::string, int>();
    std::istringstream in(k_ldpc_tables_txt);
    std::string line;
    ldpc_table* cur = nullptr;
    bool header = false;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        if (line == "@table") { header = true; cur = nullptr; continue; }
        if (line.compare(0, 6, "@links") == 0) {
            if (cur) sscanf(line.c_str(), "@links %d %d", &cur->links_total, &cur->links_max_cn);
            cur = nullptr;
            continue;
        }
        if (header) {
            char name[32];
            int n, k, m;
            if (sscanf(line.c_str(), "%31s N=%d K=%d M=%d", name, &n, &k, &m) != 4 || m != CIRC)
                throw std::runtime_error("bad LDPC table header");
            cur = &(*g_tables)[name];
            cur->name = name;
            cur->N = n;
            cur->K = k;
            header = false;
            continue;
        }
        if (cur) {
            std::istringstream ls(line);
            std::vector<int> row;
            int a;
            while (ls >> a) row.push_back(a);
            cur->rows.push_back(row);
        }
    }
    std::istringstream sin(k_ldpc_schedules_txt);
    std::vector<std::array<int, 2>>* sc = nullptr;
    while (std::getline(sin, line)) {
        if (line.empty() || line[0] == '#') continue;
        if (line.compare(0, 6, "@sched") == 0) {
            char name[32], rate[32];
            int q, n;
            sscanf(line.c_str(), "@sched %31s %31s q=%d steps=%d", name, rate, &q, &n);
            sc = &(*g_scheds)[name];
            (*g_sched_q)[name] = q;
            continue;
        }
        if (sc) {
            int a, b;
            if (sscanf(line.c_str(), "%d %d", &a, &b) == 2) sc->push_back({a, b});
        }
    }
}

}  // namespace

int fecframe_bits(int framesize)
{
    return framesize == 0 ? 64800 : framesize == 1 ? 16200 : 32400;
}

bool get_fec_params(int framesize, int rate, fec_params* out)
{
    const fec_row* b = framesize == 0 ? NORMAL : SHORT;
    size_t n = framesize == 0 ? sizeof(NORMAL) / sizeof(NORMAL[0])
             : framesize == 1 ? sizeof(SHORT) / sizeof(SHORT[0]) : 0;
    for (size_t i = 0; i < n; i++)
        if (b[i].rate == rate) {
            *out = {b[i].k, b[i].n, b[i].t};
            return true;
        }
    return false;
}

const char* ldpc_table_name(int framesize, int rate)
{
    const table_map* b = framesize == 0 ? NORMAL_TABLES : SHORT_TABLES;
    size_t n = framesize == 0 ? sizeof(NORMAL_TABLES) / sizeof(NORMAL_TABLES[0])
             : framesize == 1 ? sizeof(SHORT_TABLES) / sizeof(SHORT_TABLES[0]) : 0;
    for (size_t i = 0; i < n; i++)
        if (b[i].rate == rate) return b[i].name;
    return nullptr;
}

const ldpc_table* find_table(const std::string& name)
{
    std::call_once(g_parse_once, parse_all);
    auto it = g_tables->find(name);
    return it == g_tables->end() ? nullptr : &it->second;
}

std::vector<std::array<int, 2>> find_schedule(const std::string& table_name, int q)
{
    std::call_once(g_parse_once, parse_all);
    auto it = g_scheds->find(table_name);
    if (it == g_scheds->end() || (*g_sched_q)[table_name] != q) return {};
    return it->second;
}

decoder_geometry build_geometry(const ldpc_table& t, bool use_schedule, bool defer_groups)
{
    decoder_geometry g;
    g.N = t.N;
    g.K = t.K;
    g.R = t.N - t.K;
    g.q = g.R / CIRC;
    g.cnl = t.links_max_cn - 2;
    const int q = g.q;

    // Address a of row r belongs to layer a mod q with shift a div q: check j
    // of that layer reads VN 360 r + (j - s) mod 360 (ldpc_decode_kernel.cu:1006-1031).
    std::vector<std::vector<std::pair<int, int>>> lay(q);
    for (size_t r = 0; r < t.rows.size(); r++)
        for (int a : t.rows[r]) lay[a % q].push_back({(int)r, a / q});
    g.lay_cnt.resize(q);
    g.lay_off.resize(q);
    g.lay_moff.resize(q);
    int off = 0, moff = 0;
    for (int i = 0; i < q; i++) {
        g.lay_cnt[i] = (int)lay[i].size();
        g.lay_off[i] = off;
        g.lay_moff[i] = moff;
        off += g.lay_cnt[i];
        moff += CIRC * (g.lay_cnt[i] + 2);
    }
    g.lt_alloc = moff;
    for (int i = 0; i < q; i++)
        for (auto& rs : lay[i]) g.circ_rs.push_back((rs.first << 16) | rs.second);

    std::vector<std::array<int, 2>> sched;
    if (use_schedule) sched = find_schedule(t.name, q);
    if (sched.empty()) {
        for (int i = 0; i < q; i++) sched.push_back({i, -1});
        g.lanes = 1;
    } else {
        g.lanes = 2;
    }
    g.n_steps = (int)sched.size();

    // Every layer exactly once; the two layers of a step must not share a
    // parity variable node (layers i and i+1 mod q), since parity writes are
    // not serialised.
    std::vector<int> seen(q, 0);
    for (auto& st : sched) {
        for (int l = 0; l < 2; l++)
            if (st[l] >= 0) {
                if (st[l] >= q) throw std::runtime_error("schedule layer out of range");
                seen[st[l]]++;
            }
        if (st[0] >= 0 && st[1] >= 0) {
            int d = (st[0] - st[1] + q) % q;
            if (d == 1 || d == q - 1) throw std::runtime_error("schedule pairs parity-adjacent layers");
        }
    }
    for (int i = 0; i < q; i++)
        if (seen[i] != 1) throw std::runtime_error("schedule does not cover every layer once");

    g.circ.resize(4 * (size_t)off);
    g.grp_off.push_back(0);
    const int S = g.n_steps;
    // all (lane, k) circulants of every step in application order, by row
    std::vector<std::map<int, std::vector<int>>> by_row(S);
    for (int s = 0; s < S; s++)
        for (int l = 0; l < 2; l++) {
            int i = sched[s][l];
            if (i < 0) continue;
            for (int k = 0; k < g.lay_cnt[i]; k++) by_row[s][lay[i][k].first].push_back(g.lay_off[i] + k);
        }
    // Deferred collision groups: a step whose collided rows the next step
    // does not touch may leave its groups to be applied during the next
    // step (no extra barrier); the two steps then use disjoint slot areas
    // (slot numbers offset by step parity).  The last step of an iteration
    // is never deferred: the syndrome check that follows reads every row.
    std::vector<int> n_slot(S, 0);
    g.step_defer.assign(S, 0);
    bool any_defer = false;
    for (int s = 0; s < S; s++) {
        bool has = false, ok = s + 1 < S;
        for (auto& kv : by_row[s]) {
            if (kv.second.size() < 2) continue;
            has = true;
            n_slot[s] += (int)kv.second.size();
            if (ok && by_row[s + 1].count(kv.first)) ok = false;
        }
        g.step_defer[s] = (defer_groups && has && ok) ? 1 : 0;
        any_defer |= g.step_defer[s] != 0;
    }
    int max_even = 0, max_odd = 0;
    for (int s = 0; s < S; s++) (s & 1 ? max_odd : max_even) = std::max(s & 1 ? max_odd : max_even, n_slot[s]);
    g.slot_parity_base = any_defer ? max_even : 0;
    g.n_slots = any_defer ? max_even + max_odd : std::max(max_even, max_odd);

    for (int s = 0; s < S; s++) {
        int slot = any_defer && (s & 1) ? max_even : 0;
        std::map<int, int> slot_of;
        for (auto& kv : by_row[s]) {
            if (kv.second.size() < 2) continue;
            if (kv.second.size() > 6) throw std::runtime_error("collision group too large");
            collision_group cg{};
            cg.row = (int16_t)kv.first;
            cg.n = (int16_t)kv.second.size();
            for (size_t e = 0; e < kv.second.size(); e++) {
                cg.slot[e] = (int8_t)slot;
                slot_of[kv.second[e]] = slot++;
            }
            g.groups.push_back(cg);
        }
        if (slot > 63) throw std::runtime_error("too many collision slots in a step");
        for (int l = 0; l < 2; l++) {
            int i = sched[s][l];
            if (i < 0) continue;
            for (int k = 0; k < g.lay_cnt[i]; k++) {
                int id = g.lay_off[i] + k;
                auto it = slot_of.find(id);
                uint32_t sl = it == slot_of.end() ? 0u : (uint32_t)(it->second + 1);
                // LDS: VN group r is 180 dwords at dword 180 r; dword x holds
                // VNs (x, x + 180).  Thread t's checks (t, t + 180) read dword
                // (t - s) mod 180, halves swapped iff (t - s) mod 360 >= 180.
                const int sh = lay[i][k].second, r = lay[i][k].first;
                const int sp = sh % 180;
                const int slot_base = (g.K + g.R) / 2;  // dwords before the delta slots
                uint32_t* cr = &g.circ[4 * (size_t)id];
                cr[0] = (uint32_t)sp;
                cr[1] = (uint32_t)(4 * (180 * r - sp));
                cr[2] = sl ? (uint32_t)(4 * (slot_base + 180 * (int)(sl - 1) - sp)) : 0xffffffffu;
                cr[3] = sh >= 180 ? 0x01000302u : 0x03020100u;  // v_perm selector without wrap
            }
        }
        g.grp_off.push_back((int)g.groups.size());
        g.steps.push_back((int16_t)sched[s][0]);
        g.steps.push_back((int16_t)sched[s][1]);
    }
    return g;
}

}  // namespace cusat
