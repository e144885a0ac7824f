// LDPC kernel instantiations: guarded degree (cnt <= CMAX), both lane counts.
#include "ldpc_kernel.h"

namespace cusat {
CUSAT_LDPC_INSTANTIATE(6, false, 2)
CUSAT_LDPC_INSTANTIATE(12, false, 2)
CUSAT_LDPC_INSTANTIATE(18, false, 2)
CUSAT_LDPC_INSTANTIATE(24, false, 2)
CUSAT_LDPC_INSTANTIATE(30, false, 2)
CUSAT_LDPC_INSTANTIATE(6, false, 1)
CUSAT_LDPC_INSTANTIATE(12, false, 1)
CUSAT_LDPC_INSTANTIATE(18, false, 1)
CUSAT_LDPC_INSTANTIATE(24, false, 1)
CUSAT_LDPC_INSTANTIATE(30, false, 1)
}  // namespace cusat
