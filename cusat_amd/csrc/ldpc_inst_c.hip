// LDPC kernel instantiations: exact-degree, one lane (sequential schedules).
#include "ldpc_kernel.h"

namespace cusat {
CUSAT_LDPC_INSTANTIATE(2, true, 1)
CUSAT_LDPC_INSTANTIATE(3, true, 1)
CUSAT_LDPC_INSTANTIATE(4, true, 1)
CUSAT_LDPC_INSTANTIATE(8, true, 1)
CUSAT_LDPC_INSTANTIATE(9, true, 1)
CUSAT_LDPC_INSTANTIATE(11, true, 1)
CUSAT_LDPC_INSTANTIATE(25, true, 1)
}  // namespace cusat
