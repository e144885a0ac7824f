// LDPC decode: variant selection and dispatch (kernel in ldpc_kernel.h,
// instantiations in ldpc_inst_*.hip).
#include "kernels.h"

namespace cusat {

namespace {

// Variants compiled: exact-degree kernels for the uniform-degree codes of the
// DVB-S2/S2X tables, guarded kernels (cnt <= CMAX) for the rest.
constexpr int EXACT2[] = {2, 3, 4, 5, 7, 8, 9, 11, 12, 16, 20, 25, 28};  // 2 lanes
constexpr int EXACT1[] = {2, 3, 4, 8, 9, 11, 25};                         // 1 lane
constexpr int GUARDED[] = {6, 12, 18, 24, 30};                             // both

constexpr int rec_words(int cmax) { return ((((cmax + 2) + 1) / 2) + 3) / 4 * 4; }

}  // namespace

bool ldpc_select_variant(int cnt_min, int cnt_max, int lanes, int* cmax, bool* exact)
{
    if (cnt_min == cnt_max) {
        const int* b = lanes == 2 ? EXACT2 : EXACT1;
        const int n = lanes == 2 ? (int)(sizeof(EXACT2) / sizeof(int)) : (int)(sizeof(EXACT1) / sizeof(int));
        for (int k = 0; k < n; k++)
            if (b[k] == cnt_max) {
                *cmax = cnt_max;
                *exact = true;
                return true;
            }
    }
    for (int g : GUARDED)
        if (cnt_max <= g) {
            *cmax = g;
            *exact = false;
            return true;
        }
    return false;
}

size_t ldpc_msg_bytes(int q, int cmax) { return (size_t)q * 180 * rec_words(cmax) * 4; }

int ldpc_lds_bytes(int K, int R, int n_slots, int n_steps, int n_groups)
{
    return ((K + R) * 2 + n_slots * CIRC * 2 + 4 * (3 * n_steps + 1 + 8 * n_groups) + 15) & ~15;
}

hipError_t launch_ldpc_decode(const ldpc_args& a, int batch, hipStream_t stream)
{
#define CASE(C, E, L) \
    if (a.cmax == C && a.exact == E && a.lanes == L) return launch_ldpc_variant<C, E, L>(a, batch, stream);
    CASE(2, true, 2) CASE(3, true, 2) CASE(4, true, 2) CASE(5, true, 2) CASE(7, true, 2)
    CASE(8, true, 2) CASE(9, true, 2) CASE(11, true, 2) CASE(12, true, 2) CASE(16, true, 2)
    CASE(20, true, 2) CASE(25, true, 2) CASE(28, true, 2)
    CASE(2, true, 1) CASE(3, true, 1) CASE(4, true, 1) CASE(8, true, 1) CASE(9, true, 1)
    CASE(11, true, 1) CASE(25, true, 1)
    CASE(6, false, 2) CASE(12, false, 2) CASE(18, false, 2) CASE(24, false, 2) CASE(30, false, 2)
    CASE(6, false, 1) CASE(12, false, 1) CASE(18, false, 1) CASE(24, false, 1) CASE(30, false, 1)
#undef CASE
    return hipErrorInvalidValue;
}

}  // namespace cusat
