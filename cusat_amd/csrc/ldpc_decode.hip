// LDPC decode: variant selection and dispatch (kernel in ldpc_kernel.h,
// instantiations in ldpc_inst_*.hip), and the decode entry point that
// allocates nothing (message records in a static device array).
#include <mutex>

#include "ctx.h"
#include "cusat_fec.h"
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
    // + 4 ints: the triple-buffered flag of wg_any
    return ((K + R) * 2 + n_slots * CIRC * 2 + 4 * (3 * n_steps + 1 + 8 * n_groups + 4) + 15) & ~15;
}

bool ldpc_split_available(int cmax, bool exact, int lanes)
{
    // every variant of the 6-wave kernel has a split-edge form
    const int* b = !exact ? GUARDED : lanes == 2 ? EXACT2 : EXACT1;
    const int n = !exact        ? (int)(sizeof(GUARDED) / sizeof(int))
                  : lanes == 2 ? (int)(sizeof(EXACT2) / sizeof(int))
                               : (int)(sizeof(EXACT1) / sizeof(int));
    for (int k = 0; k < n; k++)
        if (b[k] == cmax) return true;
    return false;
}

int ldpc_lds_bytes_split(int K, int R, int n_slots, int n_steps, int n_groups, int lanes)
{
    // + the partial exchange area, + 10 ints per step: the schedule entries
    // are (layer, lay_off, count, group) per lane instead of a layer id,
    // and each lane's learned (beta, alpha)
    return ldpc_lds_bytes(K, R, n_slots, n_steps, n_groups) + 3 * lanes * 2 * 180 * 4 + 10 * n_steps * 4;
}

size_t ldpc_msg_bytes_split(int q, int cmax)
{
    // rec_half_t<cmax>::WORDS per half record
    const int neh = (cmax + 1) / 2 + 1;
    const int used = (neh + 1) / 2;
    const int words = used <= 2 ? used : (used + 3) / 4 * 4;
    return (size_t)q * 2 * 180 * words * 4;
}

int ldpc_lds_bytes_pair(int K, int R, int n_slots, int n_steps, int n_groups)
{
    // two frame regions (APP, parity, delta slots), the exchange area of
    // both lanes, the per-(step, lane) entries and learned parameters, the
    // group records and the [3][2] flags of wg_any2
    const int frame = (K + R) * 2 + n_slots * CIRC * 2;
    return (2 * frame + 3 * 2 * 2 * 180 * 4 + 48 * n_steps + 4 * (8 * n_groups + 6) + 15) & ~15;
}

hipError_t launch_ldpc_decode(const ldpc_args& a, int batch, hipStream_t stream)
{
    // the split kernel only where its LDS fits a workgroup (the 6-wave
    // kernel needs less: no partial-exchange area)
    if (a.split && ldpc_split_available(a.cmax, a.exact, a.lanes) &&
        ldpc_lds_bytes_split(a.K, a.R, a.n_slots, a.n_steps, a.n_groups, a.lanes) <= 160 * 1024) {
#define SCASE(C, E, L) \
    if (a.cmax == C && a.exact == E && a.lanes == L) return launch_ldpc_split_variant<C, E, L>(a, batch, stream);
        SCASE(2, true, 2) SCASE(3, true, 2) SCASE(4, true, 2) SCASE(5, true, 2) SCASE(7, true, 2)
        SCASE(8, true, 2) SCASE(9, true, 2) SCASE(11, true, 2) SCASE(12, true, 2) SCASE(16, true, 2)
        SCASE(20, true, 2) SCASE(25, true, 2) SCASE(28, true, 2)
        SCASE(2, true, 1) SCASE(3, true, 1) SCASE(4, true, 1) SCASE(8, true, 1) SCASE(9, true, 1)
        SCASE(11, true, 1) SCASE(25, true, 1)
        SCASE(6, false, 2) SCASE(12, false, 2) SCASE(18, false, 2) SCASE(24, false, 2) SCASE(30, false, 2)
        SCASE(6, false, 1) SCASE(12, false, 1) SCASE(18, false, 1) SCASE(24, false, 1) SCASE(30, false, 1)
#undef SCASE
    }
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

// ---- cusat_ldpc_decode_noalloc: message records in a static device array ----
//
// The decoder's only scratch is the check-to-variable message records of the
// frames in flight (one record set per frame of the batch; the kernel writes
// every record before it reads it, so nothing needs clearing).  This entry
// point keeps them in msg_scratch, a file-scope device array that every GPU
// the library runs on holds (its own copy) until the process ends, instead
// of the context's workspace.  In host code the array's name is not a device
// address: the address of the current device's copy comes from
// hipGetSymbolAddress, looked up once per device and used everywhere the
// host hands the array to a kernel.  All calls on one GPU share its copy, so
// calls that overlap on one GPU (two streams, two threads) race.

namespace cusat {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char msg_scratch[NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of the current device's msg_scratch, looked up on first use
int scratch_address(void** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_last_error("message scratch: no current device (%s)", hipGetErrorString(e));
        return CUSAT_E_NODEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(msg_scratch));
        if (e != hipSuccess || !p) {
            set_last_error("message scratch: hipGetSymbolAddress: %s", hipGetErrorString(e));
            return CUSAT_E_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CUSAT_OK;
}

}  // namespace

int ldpc_scratch_setup()
{
    void* p = nullptr;
    return scratch_address(&p);
}

}  // namespace cusat

using namespace cusat;

extern "C" int cusat_ldpc_decode_noalloc(cusat_fec_ctx* c, const int16_t* llr, uint8_t* hard, int32_t* trials,
                                         int batch, void* stream)
{
    if (!c || !llr || !hard || !trials) {
        set_last_error("bad argument");
        return CUSAT_E_INVALID;
    }
    const bool split = !(c->flags & CUSAT_CTX_NO_SPLIT_KERNEL);
    // the records of the kernel the dispatch will pick (the split kernel's
    // layout where it is used, else the 6-wave kernel's)
    const size_t rec = std::max(ldpc_msg_bytes(c->geo.q, c->cmax),
                                split ? ldpc_msg_bytes_split(c->geo.q, c->cmax) : (size_t)0);
    const long max_batch = (long)(NOALLOC_SCRATCH_BYTES / rec);
    if (batch < 1 || batch > max_batch) {
        set_last_error("batch %d outside 1..%ld (%zu bytes of message records per frame, %zu-byte static scratch)",
                       batch, max_batch, rec, NOALLOC_SCRATCH_BYTES);
        return CUSAT_E_BATCH;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != c->device) {
        set_last_error("context of device %d used on device %d", c->device, dev);
        return CUSAT_E_INVALID;
    }
    void* scratch = nullptr;
    const int rc = scratch_address(&scratch);
    if (rc != CUSAT_OK) return rc;
    ldpc_args a = make_ldpc_args(c, llr, hard, trials);
    a.msg = static_cast<int8_t*>(scratch);
    const hipError_t e = launch_ldpc_decode(a, batch, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_last_error("cusat_ldpc_decode_noalloc launch: %s", hipGetErrorString(e));
        return CUSAT_E_HIP;
    }
    return CUSAT_OK;
}
