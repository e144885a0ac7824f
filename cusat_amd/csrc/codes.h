// Code data and the decoder geometry derived from it (host side).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace cusat {

struct fec_params {
    int k_bch, n_bch, t;
};

// (k_bch, n_bch, t) of EN 302 307-1 Table 5a/5b and EN 302 307-2 (the table
// of fec/code_params.cc:17-84).  False for an undefined (frame, rate).
bool get_fec_params(int framesize, int rate, fec_params* out);
int fecframe_bits(int framesize);

// ETSI table name ("s2_b7", "s2x_c3", ...) of (frame, rate) as selected by
// ldpc_table_factory.cc:17-95, or nullptr.
const char* ldpc_table_name(int framesize, int rate);

struct ldpc_table {
    std::string name;
    int N = 0, K = 0;
    int links_total = 0, links_max_cn = 0;
    std::vector<std::vector<int>> rows;  // accumulator addresses per 360-bit row
};
// Parsed from the embedded canonical table text; nullptr if unknown.
const ldpc_table* find_table(const std::string& name);

// 2-lane schedule for a normal-frame table (ldpc_schedule_registry.h:62-99);
// empty if none.
std::vector<std::array<int, 2>> find_schedule(const std::string& table_name, int q);

constexpr int CIRC = 360;

// One VN group touched by several circulants of one schedule step: deltas
// are accumulated in LDS slots and applied in slot order after the step.
struct collision_group {
    int16_t row;        // VN group r (VNs 360 r .. 360 r + 359)
    int16_t n;          // number of slots
    int8_t slot[6];     // LDS delta slots in application order
};

struct decoder_geometry {
    int N = 0, K = 0, R = 0, q = 0, cnl = 0;
    int lt_alloc = 0;       // message bytes per frame
    int n_steps = 0, lanes = 1;
    int n_slots = 0;        // max delta slots used by any step
    std::vector<int> lay_cnt;    // [q] info circulants per layer
    std::vector<int> lay_off;    // [q] first circulant of the layer
    std::vector<int> lay_moff;   // [q] message offset of the layer in a frame
    // circulant records, 4 words each: s mod 180; LDS byte offset of the VN
    // pair dword of thread 0 (no wrap); same for its delta slot or ~0 when
    // the circulant writes the APP directly; v_perm half-swap selector
    std::vector<uint32_t> circ;
    std::vector<int> circ_rs;    // the same circulants as (row << 16) | shift
    std::vector<int16_t> steps;  // [n_steps][2], -1 = idle lane
    std::vector<int> grp_off;    // [n_steps + 1]
    std::vector<collision_group> groups;
    // [n_steps] 1: the step's collision groups may be applied during the
    // next step, which touches none of their rows (no barrier of their own)
    std::vector<int> step_defer;
    int slot_parity_base = 0;   // first slot of odd steps (0: one slot area)
};

// Per-layer circulant lists in table order (row asc, address index asc),
// schedule, and the collision plan of every step.  defer_groups: plan
// deferred collision groups (disjoint slot areas for consecutive steps).
decoder_geometry build_geometry(const ldpc_table& t, bool use_schedule, bool defer_groups = true);

}  // namespace cusat
