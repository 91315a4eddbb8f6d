// deg_order.hip -- longest-first start order for the persistent queue.
//
// A counting sort of the trajectories by their span tf_i - t0_i, descending, into 2048
// buckets given by the span's leading float bits (sign 0, 8 exponent bits, 3 mantissa bits:
// non-negative floats order like their bit patterns). Four kernels on one stream:
// clear the counters, histogram, exclusive scan in descending bucket order (one block),
// scatter. Histogram and scatter count in LDS per block of 4096 trajectories, so the global
// atomics are one per non-empty bucket per block (spans often fall into a few dozen buckets). Order within a bucket follows the atomics and is not fixed; the solver's results do
// not depend on it (each lane's arithmetic is independent of where and when it runs).
#include "deghip.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

int deg_rng_fail(int code, const char* msg);  // thread-local last error (deghip.cpp)

namespace {

constexpr int kBuckets = 2048;

__device__ __forceinline__ int span_bucket(const float* t0s, const float* tfs, float t0, float tf, long long i) {
    const float s = (tfs ? tfs[i] : tf) - (t0s ? t0s[i] : t0);
    if (!(s > 0.0f)) return 0;                       // zero, negative or NaN spans go last
    const unsigned int b = __builtin_bit_cast(unsigned int, s) >> 20;  // < 2048 for finite s > 0
    return b < (unsigned int)kBuckets ? (int)b : kBuckets - 1;
}

constexpr int kPerBlock = 4096;  // trajectories per block in the histogram and scatter passes

__global__ void __launch_bounds__(256) clear_kernel(int* __restrict__ work) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < 2 * kBuckets) work[j] = 0;
}

// per-block histogram in LDS, then one global atomic per non-empty bucket of the block
__global__ void __launch_bounds__(256) histogram_kernel(const float* t0s, const float* tfs, float t0, float tf,
                                                        long long n, int* __restrict__ count) {
    __shared__ int h[kBuckets];
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kPerBlock;
    for (int j = threadIdx.x; j < kPerBlock; j += blockDim.x) {
        const long long i = base + j;
        if (i < n) atomicAdd(&h[span_bucket(t0s, tfs, t0, tf, i)], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x)
        if (h[b] != 0) atomicAdd(&count[b], h[b]);
}

// cursor[b] = number of trajectories in buckets above b (descending order); one block of 1024
__global__ void __launch_bounds__(1024) scan_kernel(const int* __restrict__ count, int* __restrict__ cursor) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    // each thread sums two buckets, taken from the top: thread t covers b = 2047-2t, 2046-2t
    const int b0 = kBuckets - 1 - 2 * t, b1 = b0 - 1;
    const int c0 = count[b0], c1 = count[b1];
    part[t] = c0 + c1;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        const int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const int before = part[t] - (c0 + c1);
    cursor[b0] = before;
    cursor[b1] = before + c0;
}

// each block reserves one contiguous range per non-empty bucket with a single global atomic,
// then places its trajectories inside the ranges through LDS atomics
__global__ void __launch_bounds__(256) scatter_kernel(const float* t0s, const float* tfs, float t0, float tf,
                                                      long long n, int* __restrict__ cursor, int* __restrict__ order) {
    __shared__ int h[kBuckets];
    __shared__ int start[kBuckets];
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kPerBlock;
    for (int j = threadIdx.x; j < kPerBlock; j += blockDim.x) {
        const long long i = base + j;
        if (i < n) atomicAdd(&h[span_bucket(t0s, tfs, t0, tf, i)], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x) {
        start[b] = (h[b] != 0) ? atomicAdd(&cursor[b], h[b]) : 0;
        h[b] = 0;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kPerBlock; j += blockDim.x) {
        const long long i = base + j;
        if (i < n) {
            const int b = span_bucket(t0s, tfs, t0, tf, i);
            order[start[b] + atomicAdd(&h[b], 1)] = (int)i;
        }
    }
}

}  // namespace

// Scratch of deg_start_order_noalloc: the cost keys of up to DEG_NOALLOC_MAX_TRAJ trajectories,
// then the 2 x 2048 bucket counters of the counting sort (cleared by clear_kernel in every call).
// One copy per device, held until the process ends. In host code `scratch` names a host-side
// stand-in: its device address comes from hipGetSymbolAddress (deg_noalloc_scratch).
static __device__ __align__(256) float scratch[DEG_NOALLOC_MAX_TRAJ + 2 * 2048];

// The device address of `scratch` on `device` (which must be the current device), looked up
// once per device. Called by deg_ctx_create.
int deg_noalloc_scratch(int device, void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* ptrs[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return deg_rng_fail(DEG_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return deg_rng_fail(DEG_E_HIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
        ptrs[device] = ptr;
    }
    *out = ptrs[device];
    return DEG_OK;
}

extern "C" int deg_order_by_span(const float* t0s, const float* tfs, float t0, float tf, int64_t n, int32_t* order,
                                 int32_t* work, void* stream) {
    if (n < 0 || n > 0x7fffffffll) return deg_rng_fail(DEG_E_INVALID, "n must be in 0..2^31-1");
    if (n == 0) return DEG_OK;
    if (!order || !work) return deg_rng_fail(DEG_E_INVALID, "order and work are required");
    hipStream_t st = (hipStream_t)stream;
    int* count = work;
    int* cursor = work + kBuckets;
    const unsigned blocks = (unsigned)((n + kPerBlock - 1) / kPerBlock);
    hipLaunchKernelGGL(clear_kernel, dim3((2 * kBuckets + 255) / 256), dim3(256), 0, st, work);
    hipLaunchKernelGGL(histogram_kernel, dim3(blocks), dim3(256), 0, st, t0s, tfs, t0, tf, (long long)n, count);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, count, cursor);
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(256), 0, st, t0s, tfs, t0, tf, (long long)n, cursor,
                       order);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return deg_rng_fail(DEG_E_HIP, (std::string("deg_order_by_span: ") + hipGetErrorString(e)).c_str());
    return DEG_OK;
}
