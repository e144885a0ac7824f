/*
 * Standalone LDPC decoder, C++ API of lib/cusat_dsp/ldpc_decode.h:22-69
 * (same class, constructor and methods), running the gfx950 decoder of the
 * C ABI (cusat_ldpc_decode, include/cusat_fec.h).
 *
 * Errors: an undefined (frame size, rate) throws std::invalid_argument (as
 * get_fec_params does in the reference); a batch outside 1..cuda_max_batch
 * returns without decoding (ldpc_decode_kernel.cu:1291-1292); a HIP error
 * throws std::runtime_error (the reference aborts the process).
 */
#ifndef CUSAT_AMD_DSP_LDPC_DECODE_H
#define CUSAT_AMD_DSP_LDPC_DECODE_H

#include <cstdint>

#include "modcod_types.h"

struct cusat_fec_ctx;
struct ihipStream_t;

namespace cusat_dsp {

class ldpc_decode
{
public:
    ldpc_decode(fecframe_size_t framesize, code_rate_t rate, int max_trials = 25, int cuda_max_batch = 64);
    ~ldpc_decode();
    ldpc_decode(const ldpc_decode&) = delete;
    ldpc_decode& operator=(const ldpc_decode&) = delete;

    /** Host soft_in (batch x n_bits int16) -> host hard_out (batch x k_bytes,
     *  MSB first) and trials_out (remaining trials, < 0 = failed). */
    void decode_batch_cuda(const int16_t* soft_in, uint8_t* hard_out, int* trials_out, int batch_size);

    /** Device soft_in -> host hard_out / trials_out. */
    void decode_batch_cuda_d2h(const int16_t* soft_in_dev, uint8_t* hard_out, int* trials_out, int batch_size);

    /** The decoder's HIP stream (for external event recording). */
    ihipStream_t* cuda_stream() const { return d_stream; }

    unsigned int n_bits() const { return d_n_bits; }
    unsigned int k_bits() const { return d_k_bits; }
    unsigned int n_bytes() const { return d_n_bits / 8; }
    unsigned int k_bytes() const { return d_k_bits / 8; }

private:
    unsigned int d_n_bits = 0, d_k_bits = 0;
    int d_max_batch = 0;
    cusat_fec_ctx* d_ctx = nullptr;
    ihipStream_t* d_stream = nullptr;
    int16_t* d_llr_dev = nullptr;
    uint8_t* d_hard_dev = nullptr;
    int32_t* d_trials_dev = nullptr;
};

}  // namespace cusat_dsp

#endif
