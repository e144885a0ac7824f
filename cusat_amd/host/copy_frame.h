// Copy of one frame's symbols into a pinned flight slot, shared by the
// chain and the stream pool.
#pragma once

#include <immintrin.h>

#include <cstddef>
#include <cstring>

namespace cusat_dsp {
namespace detail {

// The slot (32-byte aligned) is written once and then read by the DMA
// engine, so streaming stores skip the read-for-ownership of every
// destination line; frames are far below the size at which memcpy switches
// to such stores itself.
__attribute__((target("avx2"))) inline void copy_stream_avx2(void* dst, const void* src, size_t bytes)
{
    char* d = static_cast<char*>(dst);
    const char* s = static_cast<const char*>(src);
    size_t i = 0;
    for (; i + 128 <= bytes; i += 128) {
        const __m256i a = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(s + i));
        const __m256i b = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(s + i + 32));
        const __m256i c = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(s + i + 64));
        const __m256i e = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(s + i + 96));
        _mm256_stream_si256(reinterpret_cast<__m256i*>(d + i), a);
        _mm256_stream_si256(reinterpret_cast<__m256i*>(d + i + 32), b);
        _mm256_stream_si256(reinterpret_cast<__m256i*>(d + i + 64), c);
        _mm256_stream_si256(reinterpret_cast<__m256i*>(d + i + 96), e);
    }
    _mm_sfence();  // the streaming stores are visible before the slot is handed to the GPU
    if (i < bytes) memcpy(d + i, s + i, bytes - i);
}

// dst must be 32-byte aligned
inline void copy_frame(void* dst, const void* src, size_t bytes)
{
    static const bool avx2 = __builtin_cpu_supports("avx2");
    if (avx2)
        copy_stream_avx2(dst, src, bytes);
    else
        memcpy(dst, src, bytes);
}

}  // namespace detail
}  // namespace cusat_dsp
