// LDPC kernel instantiations: exact-degree, two lanes (high degrees).
#include "ldpc_kernel.h"

namespace cusat {
CUSAT_LDPC_INSTANTIATE(11, true, 2)
CUSAT_LDPC_INSTANTIATE(12, true, 2)
CUSAT_LDPC_INSTANTIATE(16, true, 2)
CUSAT_LDPC_INSTANTIATE(20, true, 2)
CUSAT_LDPC_INSTANTIATE(25, true, 2)
CUSAT_LDPC_INSTANTIATE(28, true, 2)
}  // namespace cusat
