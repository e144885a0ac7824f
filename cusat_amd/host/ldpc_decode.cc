// cusat_dsp::ldpc_decode over the C ABI (lib/cusat_dsp/ldpc_decode.cc:19-70
// is the interface this mirrors).
#include "cusat_dsp/ldpc_decode.h"

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

#include "cusat_fec.h"

namespace cusat_dsp {

namespace {

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string("ldpc_decode: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

ldpc_decode::ldpc_decode(fecframe_size_t framesize, code_rate_t rate, int max_trials, int cuda_max_batch)
{
    cusat_fec_info inf;
    // the code's dimensions do not depend on the modulation; QPSK exists for every code
    if (cusat_fec_query((int)framesize, (int)rate, CUSAT_MOD_QPSK, 0, &inf) != CUSAT_OK)
        throw std::invalid_argument(std::string("ldpc_decode: ") + cusat_last_error());
    if (cuda_max_batch < 1 || max_trials < 0) throw std::invalid_argument("ldpc_decode: bad max_trials / batch");
    d_n_bits = (unsigned)inf.n_ldpc;
    d_k_bits = (unsigned)inf.k_ldpc;
    d_max_batch = cuda_max_batch;
    const int rc = cusat_fec_ctx_create(&d_ctx, (int)framesize, (int)rate, CUSAT_MOD_QPSK, max_trials,
                                        cuda_max_batch, 0);
    if (rc != CUSAT_OK) throw std::runtime_error(std::string("ldpc_decode: ") + cusat_last_error());
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
    d_stream = s;
    hip_check(hipMalloc(&d_llr_dev, (size_t)cuda_max_batch * d_n_bits * sizeof(int16_t)), "buffers");
    hip_check(hipMalloc(&d_hard_dev, (size_t)cuda_max_batch * (d_k_bits / 8)), "buffers");
    hip_check(hipMalloc(&d_trials_dev, (size_t)cuda_max_batch * sizeof(int32_t)), "buffers");
}

ldpc_decode::~ldpc_decode()
{
    if (d_stream) hipStreamSynchronize(d_stream);
    if (d_llr_dev) hipFree(d_llr_dev);
    if (d_hard_dev) hipFree(d_hard_dev);
    if (d_trials_dev) hipFree(d_trials_dev);
    if (d_stream) hipStreamDestroy(d_stream);
    cusat_fec_ctx_destroy(d_ctx);
}

void ldpc_decode::decode_batch_cuda(const int16_t* soft_in, uint8_t* hard_out, int* trials_out, int batch_size)
{
    if (batch_size < 1 || batch_size > d_max_batch) return;
    hip_check(hipMemcpyAsync(d_llr_dev, soft_in, (size_t)batch_size * d_n_bits * sizeof(int16_t),
                             hipMemcpyHostToDevice, d_stream),
              "copy in");
    decode_batch_cuda_d2h(d_llr_dev, hard_out, trials_out, batch_size);
}

void ldpc_decode::decode_batch_cuda_d2h(const int16_t* soft_in_dev, uint8_t* hard_out, int* trials_out,
                                        int batch_size)
{
    if (batch_size < 1 || batch_size > d_max_batch) return;
    if (cusat_ldpc_decode(d_ctx, soft_in_dev, d_hard_dev, d_trials_dev, batch_size, d_stream) != CUSAT_OK)
        throw std::runtime_error(std::string("ldpc_decode: ") + cusat_last_error());
    hip_check(hipMemcpyAsync(hard_out, d_hard_dev, (size_t)batch_size * (d_k_bits / 8), hipMemcpyDeviceToHost,
                             d_stream),
              "copy out");
    hip_check(hipMemcpyAsync(trials_out, d_trials_dev, (size_t)batch_size * sizeof(int32_t),
                             hipMemcpyDeviceToHost, d_stream),
              "copy out");
    hip_check(hipStreamSynchronize(d_stream), "sync");
}

}  // namespace cusat_dsp
