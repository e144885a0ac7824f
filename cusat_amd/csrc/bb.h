// Baseband helpers shared by the C ABI and the kernels (bb.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cusat {

// PRBS bytes of one BBFRAME (EN 302 307-1 clause 5.2.2), MSB first.
void bb_prbs(uint8_t* out, size_t n_bytes);
// BBHEADER CRC-8 (clause 5.1.6): poly 0xD5, register cleared, MSB first.
__host__ __device__ uint8_t bb_crc8(const uint8_t* d, int n);
hipError_t launch_bb_extract(const uint8_t* hard, int hard_stride, const int32_t* trials, const int32_t* valid,
                             const uint8_t* prbs, int kbch_bytes, uint8_t* out, int32_t* dfl_out, int batch,
                             hipStream_t stream);

}  // namespace cusat
