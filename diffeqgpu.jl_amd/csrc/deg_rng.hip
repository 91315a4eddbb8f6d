// deg_rng.hip -- device-side seeded parameter draws for ensemble sweeps.
//
// The counter-based generator is SplitMix64 (Steele, Lea and Flood 2014) evaluated at counter
// k of stream `seed`: z = seed + (k + 1) * 0x9E3779B97F4A7C15 followed by the SplitMix64
// finaliser; U[0,1) = (z >> 40) * 2^-24, exact in Float32. The host generator
// (diffeqgpu.jl_amd/rng.py) is the same function, so a draw on the device equals the
// host draw bit for bit. Entry (row r, trajectory i) of a table with `rows` rows uses
// counter i * rows + r, which makes a trajectory's inputs independent of batch size and order.
//
// out[i] = lo + scale * U is evaluated in Float32 without contraction (a separately rounded
// multiply and add), matching NumPy's float32 `lo + scale * U`.
//
// deg_draw_normal exposes the SDE kernels' Wiener-increment generator (Philox4x32-10 and a
// deterministic Box-Muller, device/deg_philox.h) on its own.
#include "deghip.h"

#include <hip/hip_runtime.h>

#include <string>

#include "device/deg_philox.h"

namespace {

__device__ __forceinline__ unsigned long long deg_splitmix64(unsigned long long seed, unsigned long long k) {
    unsigned long long z = seed + (k + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) deg_draw_uniform_kernel(float* __restrict__ out, long long n, long long id_offset,
                                                               unsigned long long seed, int rows, int row, float lo,
                                                               float scale) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = (unsigned long long)(id_offset + i) * (unsigned long long)rows + (unsigned long long)row;
    const float u = (float)(deg_splitmix64(seed, k) >> 40) * 5.9604644775390625e-08f;  // 2^-24
    out[i] = lo + scale * u;
}

// out[c * n + i] = normal c of step `step`, stream id_offset + i (the SDE kernels' increments / sqrt(dt))
__global__ void __launch_bounds__(256) deg_draw_normal_kernel(float* __restrict__ out, long long n, long long id_offset,
                                                              unsigned long long seed, unsigned int step, int m) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long id = (unsigned long long)(id_offset + i);
    for (int b = 0; 4 * b < m; ++b) {
        const deg_u4 r = deg_philox4x32_10(deg_u4{step, (unsigned int)b, (unsigned int)id, (unsigned int)(id >> 32)},
                                           (unsigned int)seed, (unsigned int)(seed >> 32));
        float z[4];
        deg_bm_pair(r.x, r.y, z[0], z[1]);
        deg_bm_pair(r.z, r.w, z[2], z[3]);
        for (int j = 0; j < 4 && 4 * b + j < m; ++j) out[(long long)(4 * b + j) * n + i] = z[j];
    }
}

}  // namespace

int deg_rng_fail(int code, const char* msg);  // defined in deghip.cpp (thread-local last error)

extern "C" int deg_draw_uniform(float* out, int64_t n, int64_t id_offset, uint64_t seed, int32_t rows, int32_t row,
                                float lo, float scale, void* stream) {
    if (n < 0 || id_offset < 0) return deg_rng_fail(DEG_E_INVALID, "n and id_offset must be >= 0");
    if (rows < 1 || row < 0 || row >= rows) return deg_rng_fail(DEG_E_INVALID, "need 0 <= row < rows");
    if (n == 0) return DEG_OK;
    if (!out) return deg_rng_fail(DEG_E_INVALID, "out is NULL");
    const long long blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffll) return deg_rng_fail(DEG_E_INVALID, "n too large for one launch");
    hipLaunchKernelGGL(deg_draw_uniform_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out,
                       (long long)n, (long long)id_offset, (unsigned long long)seed, (int)rows, (int)row, lo, scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return deg_rng_fail(DEG_E_HIP, (std::string("draw launch: ") + hipGetErrorString(e)).c_str());
    return DEG_OK;
}

extern "C" int deg_draw_normal(float* out, int64_t n, int64_t id_offset, uint64_t seed, uint32_t step, int32_t m,
                               void* stream) {
    if (n < 0 || id_offset < 0) return deg_rng_fail(DEG_E_INVALID, "n and id_offset must be >= 0");
    if (m < 1 || m > 4096) return deg_rng_fail(DEG_E_INVALID, "m must be in 1..4096");
    if (n == 0) return DEG_OK;
    if (!out) return deg_rng_fail(DEG_E_INVALID, "out is NULL");
    const long long blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffll) return deg_rng_fail(DEG_E_INVALID, "n too large for one launch");
    hipLaunchKernelGGL(deg_draw_normal_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out,
                       (long long)n, (long long)id_offset, (unsigned long long)seed, (unsigned int)step, (int)m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return deg_rng_fail(DEG_E_HIP, (std::string("normal draw launch: ") + hipGetErrorString(e)).c_str());
    return DEG_OK;
}
