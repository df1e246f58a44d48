// Streaming-copy probe for the benchmark's roofline: the read + write bandwidth a plain
// gfx950 copy kernel reaches, as the practical ceiling next to the 8 TB/s specification.
// Not part of the FFT library (built into its own libbbfft_stream.so).
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));

// One 16-byte element per work-item, one work-group per 256 elements, streaming hints.
__global__ void __launch_bounds__(256) copy_v4(const v4f *__restrict__ in, v4f *__restrict__ out,
                                               std::size_t n) {
    std::size_t const i = blockIdx.x * std::size_t(256) + threadIdx.x;
    if (i < n) {
        __builtin_nontemporal_store(__builtin_nontemporal_load(in + i), out + i);
    }
}

// A 16 KiB tile per work-group, 8 loads of 8 bytes in flight per work-item before the stores:
// the access pattern of an FFT tile of 2048 complex fp32 values.
__global__ void __launch_bounds__(256) copy_tile(const v2f *__restrict__ in, v2f *__restrict__ out,
                                                 std::size_t n) {
    std::size_t const base = blockIdx.x * std::size_t(2048) + threadIdx.x;
    v2f v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        std::size_t const i = base + 256u * j;
        v[j] = i < n ? __builtin_nontemporal_load(in + i) : v2f{0.0f, 0.0f};
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        std::size_t const i = base + 256u * j;
        if (i < n) {
            __builtin_nontemporal_store(v[j], out + i);
        }
    }
}

} // namespace

/// Copies `bytes` (a multiple of 16; both pointers 16-byte aligned) from in to out on `stream`
/// with kernel `kind` (0: one 16-byte element per work-item, 1: 16 KiB tiles). Returns a
/// hipError_t value.
extern "C" __attribute__((visibility("default"))) int
bbfft_stream_copy(const void *in, void *out, std::size_t bytes, int kind, hipStream_t stream) {
    if (bytes % 16 != 0 || reinterpret_cast<std::size_t>(in) % 16 != 0 ||
        reinterpret_cast<std::size_t>(out) % 16 != 0 || kind < 0 || kind > 1) {
        return static_cast<int>(hipErrorInvalidValue);
    }
    if (bytes == 0) {
        return static_cast<int>(hipSuccess);
    }
    if (kind == 0) {
        std::size_t const n = bytes / 16;
        copy_v4<<<dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream>>>(
            static_cast<const v4f *>(in), static_cast<v4f *>(out), n);
    } else {
        std::size_t const n = bytes / 8;
        copy_tile<<<dim3(static_cast<unsigned>((n + 2047) / 2048)), dim3(256), 0, stream>>>(
            static_cast<const v2f *>(in), static_cast<v2f *>(out), n);
    }
    return static_cast<int>(hipGetLastError());
}
