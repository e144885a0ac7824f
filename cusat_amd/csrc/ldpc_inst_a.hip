// LDPC kernel instantiations: exact-degree, two lanes (low degrees).
#include "ldpc_kernel.h"

namespace cusat {
CUSAT_LDPC_INSTANTIATE(2, true, 2)
CUSAT_LDPC_INSTANTIATE(3, true, 2)
CUSAT_LDPC_INSTANTIATE(4, true, 2)
CUSAT_LDPC_INSTANTIATE(5, true, 2)
CUSAT_LDPC_INSTANTIATE(7, true, 2)
CUSAT_LDPC_INSTANTIATE(8, true, 2)
CUSAT_LDPC_INSTANTIATE(9, true, 2)
}  // namespace cusat
