// synth.hip -- seeded uniform[-1,1) inputs generated in HBM (SplitMix64,
// u = (x >> 11) 2^-53, value 2u - 1), draw i of the stream at state
// seed + (i + 1) * 0x9E3779B97F4A7C15.
#include <string>

#include "../../include/nrfft.h"
#include "internal.hpp"

namespace nrfft {

__device__ __forceinline__ double splitmix_u(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * 0x1.0p-53;
}

template <typename T>
__global__ void k_gen_uniform(T* out, uint64_t n, uint64_t seed, uint64_t start) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        out[i] = (T)(2.0 * splitmix_u(seed, start + i) - 1.0);
}

template <typename T>
static int gen(void* d, size_t n, uint64_t seed, uint64_t start, void* stream) {
    if (!d && n) return fail(NRFFT_E_INVALID, "null buffer");
    if (n == 0) return NRFFT_OK;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 1u << 16);
    hipLaunchKernelGGL((k_gen_uniform<T>), dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (T*)d,
                       (uint64_t)n, seed, start);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(NRFFT_E_HIP, std::string("gen: ") + hipGetErrorString(e));
    return NRFFT_OK;
}

}  // namespace nrfft

extern "C" {
int nrfft_gen_uniform_f32(void* d, size_t n, uint64_t seed, uint64_t start, void* stream) {
    nrfft::g_err.clear();
    return nrfft::gen<float>(d, n, seed, start, stream);
}
int nrfft_gen_uniform_f64(void* d, size_t n, uint64_t seed, uint64_t start, void* stream) {
    nrfft::g_err.clear();
    return nrfft::gen<double>(d, n, seed, start, stream);
}
}
