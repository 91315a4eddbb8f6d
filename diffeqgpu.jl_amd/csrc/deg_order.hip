// deg_order.hip -- longest-first start order for the persistent queue.
//
// A counting sort of the trajectories by their span tf_i - t0_i, descending. Non-negative floats order
// like their bit patterns, so the buckets are ranges of bit patterns: bucket 0 holds zero, negative and
// NaN spans (started last); buckets 1..2047 split [bmin, bmax], the bit patterns of the smallest and
// largest positive span, into ranges of 2^shift patterns, with shift the smallest that fits the range
// into 2047 buckets. Spans within a factor of 2 of each other thus get a resolution of about 1/2000
// octave whatever their magnitude (a fixed split by the leading bits would give 1/8 octave, and a
// narrow key range would fall into a handful of buckets). Five kernels on one stream: clear the
// counters, the range (one atomic min/max pair per block), histogram, exclusive scan in descending
// bucket order (one block), scatter. Histogram and scatter count in LDS per block of 8192 trajectories,
// so the global atomics are one per non-empty bucket per block. Order within a bucket follows the
// atomics and is not fixed; the solver's results do not depend on it (each lane's arithmetic is
// independent of where and when it runs).
//
// work (4096 int32): count[2048], then cursor[2048]. Until the scan, cursor[0] and cursor[1] hold the
// range (bmin, bmax); after it, count[0] and count[1] hold (bmin, shift) for the scatter.
#include "deghip.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

int deg_rng_fail(int code, const char* msg);  // thread-local last error (deghip.cpp)

namespace {

constexpr int kBuckets = 2048;
constexpr int kThreads = 256;                    // block size of the range, histogram and scatter passes
constexpr int kPerThread = 32;                   // trajectories per thread in those passes
constexpr int kPerBlock = kThreads * kPerThread;  // trajectories per block

// The spans of this thread's kPerThread trajectories (base + threadIdx.x + m * kThreads), with every
// load issued before the first is used: a loop that used each value as it arrived waited out one
// memory latency per trajectory. There is no branch around a load: which arrays exist is a template
// parameter (kT0: t0s, kTF: tfs), and an index at or past n reads entry n - 1 (n >= 1) and yields NaN,
// which falls into bucket 0 and must not be counted.
template <bool kT0, bool kTF>
__device__ __forceinline__ void load_spans(const float* t0s, const float* tfs, float t0, float tf, long long base,
                                           long long n, float (&s)[kPerThread]) {
    float a[kPerThread], b[kPerThread];
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) {
        const long long i = base + threadIdx.x + (long long)m * kThreads;
        const long long ic = i < n ? i : n - 1;
        a[m] = kTF ? tfs[ic] : tf;
        b[m] = kT0 ? t0s[ic] : t0;
    }
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) {
        const long long i = base + threadIdx.x + (long long)m * kThreads;
        s[m] = i < n ? a[m] - b[m] : __builtin_nanf("");
    }
}

// the smallest shift with (bmax - bmin) >> shift <= 2046; 0 when there is at most one positive value
__device__ __forceinline__ unsigned bucket_shift(unsigned bmin, unsigned bmax) {
    if (bmax <= bmin) return 0u;
    const unsigned r = bmax - bmin;
    unsigned sh = 0u;
    while ((r >> sh) > (unsigned)(kBuckets - 2)) ++sh;
    return sh;
}

__device__ __forceinline__ int span_bucket(float s, unsigned bmin, unsigned shift) {
    if (!(s > 0.0f)) return 0;  // zero, negative or NaN spans go last
    const unsigned d = (__builtin_bit_cast(unsigned, s) - bmin) >> shift;
    return 1 + (int)(d < (unsigned)(kBuckets - 2) ? d : (unsigned)(kBuckets - 2));
}

__global__ void __launch_bounds__(256) clear_kernel(int* __restrict__ work) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < kBuckets) work[j] = 0;                          // counters
    if (j == 0) {
        work[kBuckets] = (int)0xFFFFFFFFu;                  // range minimum (as unsigned)
        work[kBuckets + 1] = 0;                             // range maximum
    }
}

// bit patterns of the smallest and largest positive span: a block reduction, then one atomic pair
template <bool kT0, bool kTF>
__global__ void __launch_bounds__(kThreads) range_kernel(const float* t0s, const float* tfs, float t0, float tf,
                                                         long long n, unsigned* __restrict__ mm) {
    __shared__ unsigned wlo[kThreads / 64], whi[kThreads / 64];
    float sp[kPerThread];
    load_spans<kT0, kTF>(t0s, tfs, t0, tf, (long long)blockIdx.x * kPerBlock, n, sp);
    unsigned lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) {
        if (sp[m] > 0.0f) {  // false for NaN, so entries past n drop out
            const unsigned b = __builtin_bit_cast(unsigned, sp[m]);
            lo = b < lo ? b : lo;
            hi = b > hi ? b : hi;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned l2 = __shfl_xor(lo, off), h2 = __shfl_xor(hi, off);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wlo[w] = lo;
        whi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kThreads / 64; ++k) {
            lo = wlo[k] < lo ? wlo[k] : lo;
            hi = whi[k] > hi ? whi[k] : hi;
        }
        if (lo <= hi) {
            atomicMin(&mm[0], lo);
            atomicMax(&mm[1], hi);
        }
    }
}

// per-block histogram in LDS, then one global atomic per non-empty bucket of the block
template <bool kT0, bool kTF>
__global__ void __launch_bounds__(kThreads) histogram_kernel(const float* t0s, const float* tfs, float t0, float tf,
                                                             long long n, int* __restrict__ count,
                                                             const unsigned* __restrict__ mm) {
    __shared__ int h[kBuckets];
    const long long base = (long long)blockIdx.x * kPerBlock;
    float sp[kPerThread];
    load_spans<kT0, kTF>(t0s, tfs, t0, tf, base, n, sp);
    const unsigned bmin = mm[0], shift = bucket_shift(mm[0], mm[1]);
    for (int b = threadIdx.x; b < kBuckets; b += kThreads) h[b] = 0;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kPerThread; ++m)
        if (base + threadIdx.x + (long long)m * kThreads < n) atomicAdd(&h[span_bucket(sp[m], bmin, shift)], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < kBuckets; b += kThreads)
        if (h[b] != 0) atomicAdd(&count[b], h[b]);
}

// cursor[b] = number of trajectories in buckets above b (descending order); one block of 1024. It
// reads the range from cursor[0..1] before overwriting them and leaves (bmin, shift) in count[0..1].
__global__ void __launch_bounds__(1024) scan_kernel(int* __restrict__ count, int* __restrict__ cursor) {
    __shared__ int part[1024];
    __shared__ unsigned range[2];
    const int t = threadIdx.x;
    if (t == 0) {
        range[0] = (unsigned)cursor[0];
        range[1] = (unsigned)cursor[1];
    }
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
    if (t == 0) {  // every thread read its counts before the first barrier
        count[0] = (int)range[0];
        count[1] = (int)bucket_shift(range[0], range[1]);
    }
}

// Optional payload of the scatter: columns of per-trajectory values (two source arrays of ncol0 and
// ncol1 columns, leading dimension src_stride, entry i of column c at src[c * src_stride + i]) copied
// into dst in queue-major order position: order position k goes to column index (k mod 8) * slab + k / 8
// with slab = ceil(n / 8), columns 8 * slab apart. The persistent solver's queue q serves positions
// q, q + 8, q + 16, ...; in this layout each queue reads its trajectories' inputs contiguously.
struct order_payload {
    const float* src0;
    const float* src1;
    long long src_stride;
    int ncol0, ncol1;
    float* dst;
};

// each block reserves one contiguous range per non-empty bucket with a single global atomic,
// then places its trajectories inside the ranges through LDS atomics
template <bool kT0, bool kTF>
__global__ void __launch_bounds__(kThreads) scatter_kernel(const float* t0s, const float* tfs, float t0, float tf,
                                                           long long n, int* __restrict__ cursor,
                                                           const int* __restrict__ count, int* __restrict__ order,
                                                           const order_payload pl) {
    __shared__ int h[kBuckets];
    __shared__ int start[kBuckets];
    const long long base = (long long)blockIdx.x * kPerBlock;
    float sp[kPerThread];
    load_spans<kT0, kTF>(t0s, tfs, t0, tf, base, n, sp);
    const unsigned bmin = (unsigned)count[0], shift = (unsigned)count[1];
    int bk[kPerThread];  // bucket of each of this thread's trajectories (both passes use it)
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) bk[m] = span_bucket(sp[m], bmin, shift);
    for (int b = threadIdx.x; b < kBuckets; b += kThreads) h[b] = 0;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kPerThread; ++m)
        if (base + threadIdx.x + (long long)m * kThreads < n) atomicAdd(&h[bk[m]], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < kBuckets; b += kThreads) {
        start[b] = (h[b] != 0) ? atomicAdd(&cursor[b], h[b]) : 0;
        h[b] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) {
        const long long i = base + threadIdx.x + (long long)m * kThreads;
        if (i < n) {
            const int pos = start[bk[m]] + atomicAdd(&h[bk[m]], 1);
            order[pos] = (int)i;
            if (pl.dst != nullptr) {
                // coalesced reads of this block's trajectories, one store per column at their positions
                const long long slab = (n + 7) / 8;
                float* d = pl.dst + (long long)(pos & 7) * slab + (pos >> 3);
                for (int c = 0; c < pl.ncol0; ++c) d[(long long)c * 8 * slab] = pl.src0[(long long)c * pl.src_stride + i];
                for (int c = 0; c < pl.ncol1; ++c)
                    d[(long long)(pl.ncol0 + c) * 8 * slab] = pl.src1[(long long)c * pl.src_stride + i];
            }
        }
    }
}

// the range, histogram, scan and scatter passes for one combination of span arrays
template <bool kT0, bool kTF>
void launch_passes(hipStream_t st, unsigned blocks, const float* t0s, const float* tfs, float t0, float tf, long long n,
                   int* work, int32_t* order, const order_payload& pl) {
    int* count = work;
    int* cursor = work + kBuckets;
    unsigned* range = (unsigned*)cursor;  // (bmin, bmax) until the scan
    hipLaunchKernelGGL((range_kernel<kT0, kTF>), dim3(blocks), dim3(kThreads), 0, st, t0s, tfs, t0, tf, n, range);
    hipLaunchKernelGGL((histogram_kernel<kT0, kTF>), dim3(blocks), dim3(kThreads), 0, st, t0s, tfs, t0, tf, n, count,
                       (const unsigned*)range);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, count, cursor);
    hipLaunchKernelGGL((scatter_kernel<kT0, kTF>), dim3(blocks), dim3(kThreads), 0, st, t0s, tfs, t0, tf, n, cursor,
                       (const int*)count, order, pl);
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

// deg_order_by_span, and with payload columns (src0: ncol0 columns, src1: ncol1 columns, leading
// dimension src_stride) also copied to dst ((ncol0 + ncol1) columns of 8 * ceil(n / 8) floats) in
// queue-major order position (see order_payload). Used by the start orders of deghip.cpp to lay out
// each trajectory's inputs in the order the persistent solver takes them.
int deg_order_payload(const float* t0s, const float* tfs, float t0, float tf, int64_t n, int32_t* order,
                      int32_t* work, void* stream, const float* src0, int ncol0, const float* src1, int ncol1,
                      long long src_stride, float* dst) {
    if (n < 0 || n > 0x7fffffffll) return deg_rng_fail(DEG_E_INVALID, "n must be in 0..2^31-1");
    if (n == 0) return DEG_OK;
    if (!order || !work) return deg_rng_fail(DEG_E_INVALID, "order and work are required");
    hipStream_t st = (hipStream_t)stream;
    const unsigned blocks = (unsigned)((n + kPerBlock - 1) / kPerBlock);
    order_payload pl;
    pl.src0 = src0;
    pl.src1 = src1;
    pl.src_stride = src_stride;
    pl.ncol0 = src0 ? ncol0 : 0;
    pl.ncol1 = src1 ? ncol1 : 0;
    pl.dst = dst;
    hipLaunchKernelGGL(clear_kernel, dim3(kBuckets / 256), dim3(256), 0, st, work);
    if (t0s && tfs)
        launch_passes<true, true>(st, blocks, t0s, tfs, t0, tf, n, work, order, pl);
    else if (t0s)
        launch_passes<true, false>(st, blocks, t0s, tfs, t0, tf, n, work, order, pl);
    else if (tfs)
        launch_passes<false, true>(st, blocks, t0s, tfs, t0, tf, n, work, order, pl);
    else
        launch_passes<false, false>(st, blocks, t0s, tfs, t0, tf, n, work, order, pl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return deg_rng_fail(DEG_E_HIP, (std::string("deg_order_by_span: ") + hipGetErrorString(e)).c_str());
    return DEG_OK;
}

extern "C" int deg_order_by_span(const float* t0s, const float* tfs, float t0, float tf, int64_t n, int32_t* order,
                                 int32_t* work, void* stream) {
    return deg_order_payload(t0s, tfs, t0, tf, n, order, work, stream, nullptr, 0, nullptr, 0, 0, nullptr);
}
