// deg_mlp.hip -- training and inference kernels of the learned step-size proposer.
//
// The solver writes one step-log row per attempt (DEG_F_STEPLOG, DEG_STEPLOG_COLS floats):
//   [x0..x5 at attempt n, log2(dt_{n+1} / dt_n), log2 EEst_{n+1}, z_PI at n, accepted at n, 0, 0]
// The regression target of a row is the correction to the PI controller's log2 step factor
// z_PI that would have hit the target error ratio rho at attempt n+1, if the error scaled as
// dt^(1/expo):
//   y = clamp(log2(dt_{n+1} / dt_n) + expo * (log2 rho - log2 EEst_{n+1}), -ycap, ycap) - z_PI.
// The loss is the mean squared error of the network output against y.
//
// Kernels (FP32 unless stated):
//   deg_mlp_grad   one wavefront per block, 64-row tiles staged in LDS; every thread owns a
//                  fixed set of parameters and accumulates their gradient over the block's
//                  rows; one partial gradient row per block (deterministic, no atomics)
//   deg_mlp_adam   sums the partial rows (four interleaved partial sums per column, added in a
//                  fixed order) and applies one Adam step
//   deg_mlp_forward  FP32 forward pass (evaluation)
//   deg_mlp_pack   FP32 master parameters -> BF16 MFMA fragments (struct deg_mlp_packed)
//   deg_mlp_infer  the solver's MFMA evaluation (deg_mlp_eval) over an array of feature rows
#include "deghip.h"

#include <hip/hip_runtime.h>

#include <string>

#include "device/deg_mlp.h"

int deg_rng_fail(int code, const char* msg);  // thread-local last error (deghip.cpp)

namespace {

constexpr int kNP = DEG_MLP_NPARAM;
constexpr int kTile = 64;
constexpr int kOwn = (kNP + 63) / 64;  // parameters owned per thread (21)
constexpr int kPartial = kNP + 2;      // per-block gradient row + loss sum + row count

struct TrainCfg {
    float mean[DEG_MLP_IN];
    float inv_std[DEG_MLP_IN];
    float expo, log2_rho, ycap, pad;
};

__device__ __forceinline__ float target_of(const float* row, const TrainCfg& c) {
    const float y = row[6] + c.expo * (c.log2_rho - row[7]);
    return __builtin_fminf(__builtin_fmaxf(y, -c.ycap), c.ycap) - row[8];
}

__global__ void __launch_bounds__(64) deg_mlp_grad_kernel(const float* __restrict__ rows, const int* __restrict__ idx,
                                                          long long nrows, long long per_block,
                                                          const float* __restrict__ P, TrainCfg cfg,
                                                          float* __restrict__ partial) {
    __shared__ float sP[kNP];
    __shared__ float sx[kTile][DEG_MLP_IN + 1];
    __shared__ float sh1[kTile][DEG_MLP_HID + 1];
    __shared__ float sh2[kTile][DEG_MLP_HID + 1];
    __shared__ float sd1[kTile][DEG_MLP_HID + 1];
    __shared__ float sd2[kTile][DEG_MLP_HID + 1];
    __shared__ float sdo[kTile];
    __shared__ float sred[64];
    const int t = threadIdx.x;
    for (int j = t; j < kNP; j += 64) sP[j] = P[j];
    __syncthreads();
    float acc[kOwn];
#pragma unroll
    for (int q = 0; q < kOwn; ++q) acc[q] = 0.0f;
    float loss = 0.0f, count = 0.0f;
    const long long begin = (long long)blockIdx.x * per_block;
    const long long end = min(begin + per_block, nrows);
    const float inv_n = 1.0f / (float)nrows;
    for (long long base = begin; base < end; base += kTile) {
        const long long r = base + t;
        const bool valid = r < end;
        float x[DEG_MLP_IN], h1[DEG_MLP_HID], h2[DEG_MLP_HID], d2[DEG_MLP_HID], d1[DEG_MLP_HID];
        float dout = 0.0f;
#pragma unroll
        for (int k = 0; k < DEG_MLP_IN; ++k) x[k] = 0.0f;
        if (valid) {
            const long long src = idx ? (long long)idx[r] : r;
            const float* row = rows + src * DEG_STEPLOG_COLS;
#pragma unroll
            for (int k = 0; k < DEG_MLP_IN; ++k) x[k] = (row[k] - cfg.mean[k]) * cfg.inv_std[k];
            const float y = target_of(row, cfg);
#pragma unroll
            for (int m = 0; m < DEG_MLP_HID; ++m) {
                float s = sP[DEG_MLP_OFF_B1 + m];
#pragma unroll
                for (int k = 0; k < DEG_MLP_IN; ++k) s = __builtin_fmaf(x[k], sP[DEG_MLP_OFF_W1 + k * DEG_MLP_HID + m], s);
                h1[m] = __builtin_fmaxf(s, 0.0f);
            }
            float o = sP[DEG_MLP_OFF_B3];
            for (int m = 0; m < DEG_MLP_HID; ++m) {
                float s = sP[DEG_MLP_OFF_B2 + m];
#pragma unroll
                for (int k = 0; k < DEG_MLP_HID; ++k) s = __builtin_fmaf(h1[k], sP[DEG_MLP_OFF_W2 + k * DEG_MLP_HID + m], s);
                h2[m] = __builtin_fmaxf(s, 0.0f);
                o = __builtin_fmaf(sP[DEG_MLP_OFF_W3 + m], h2[m], o);
            }
            const float diff = o - y;
            loss += diff * diff;
            count += 1.0f;
            dout = 2.0f * diff * inv_n;
#pragma unroll
            for (int m = 0; m < DEG_MLP_HID; ++m) d2[m] = (h2[m] > 0.0f) ? dout * sP[DEG_MLP_OFF_W3 + m] : 0.0f;
            for (int k = 0; k < DEG_MLP_HID; ++k) {
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < DEG_MLP_HID; ++m) s = __builtin_fmaf(d2[m], sP[DEG_MLP_OFF_W2 + k * DEG_MLP_HID + m], s);
                d1[k] = (h1[k] > 0.0f) ? s : 0.0f;
            }
        } else {
#pragma unroll
            for (int m = 0; m < DEG_MLP_HID; ++m) h1[m] = h2[m] = d1[m] = d2[m] = 0.0f;
        }
#pragma unroll
        for (int k = 0; k < DEG_MLP_IN; ++k) sx[t][k] = x[k];
#pragma unroll
        for (int m = 0; m < DEG_MLP_HID; ++m) {
            sh1[t][m] = h1[m];
            sh2[t][m] = h2[m];
            sd1[t][m] = d1[m];
            sd2[t][m] = d2[m];
        }
        sdo[t] = dout;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kOwn; ++q) {
            const int p = t + 64 * q;
            if (p >= kNP) break;
            float s = 0.0f;
            if (p < DEG_MLP_OFF_B1) {
                const int k = p / DEG_MLP_HID, m = p % DEG_MLP_HID;
                for (int rr = 0; rr < kTile; ++rr) s = __builtin_fmaf(sx[rr][k], sd1[rr][m], s);
            } else if (p < DEG_MLP_OFF_W2) {
                const int m = p - DEG_MLP_OFF_B1;
                for (int rr = 0; rr < kTile; ++rr) s += sd1[rr][m];
            } else if (p < DEG_MLP_OFF_B2) {
                const int qq = p - DEG_MLP_OFF_W2, k = qq / DEG_MLP_HID, m = qq % DEG_MLP_HID;
                for (int rr = 0; rr < kTile; ++rr) s = __builtin_fmaf(sh1[rr][k], sd2[rr][m], s);
            } else if (p < DEG_MLP_OFF_W3) {
                const int m = p - DEG_MLP_OFF_B2;
                for (int rr = 0; rr < kTile; ++rr) s += sd2[rr][m];
            } else if (p < DEG_MLP_OFF_B3) {
                const int m = p - DEG_MLP_OFF_W3;
                for (int rr = 0; rr < kTile; ++rr) s = __builtin_fmaf(sh2[rr][m], sdo[rr], s);
            } else {
                for (int rr = 0; rr < kTile; ++rr) s += sdo[rr];
            }
            acc[q] += s;
        }
        __syncthreads();
    }
    float* out = partial + (long long)blockIdx.x * kPartial;
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
        const int p = t + 64 * q;
        if (p < kNP) out[p] = acc[q];
    }
    // block sums of the loss and the row count (fixed order)
    sred[t] = loss;
    __syncthreads();
    if (t == 0) {
        float s = 0.0f;
        for (int j = 0; j < 64; ++j) s += sred[j];
        out[kNP] = s;
    }
    __syncthreads();
    sred[t] = count;
    __syncthreads();
    if (t == 0) {
        float s = 0.0f;
        for (int j = 0; j < 64; ++j) s += sred[j];
        out[kNP + 1] = s;
    }
}

// One block per 64 columns of the partial rows (parameters, then the loss sum and the row count):
// each of the block's four wavefronts sums every fourth partial row of its columns, then the four
// sums are added in a fixed order. Deterministic for a given nblocks, and the loads of one column
// are independent of each other (no serial chain of nblocks dependent loads).
constexpr int kAdamSeg = 4;
__global__ void __launch_bounds__(256) deg_mlp_adam_kernel(float* __restrict__ P, float* __restrict__ M,
                                                           float* __restrict__ V, const float* __restrict__ partial,
                                                           int nblocks, float lr, float beta1, float beta2, float eps,
                                                           float bc1, float bc2, float* __restrict__ loss_out) {
    __shared__ float sg[kAdamSeg][64];
    __shared__ float tot[64];
    const int pl = threadIdx.x & 63, seg = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + pl;
    float g = 0.0f;
    if (p < kPartial) {
        int b = seg;
        for (; b + 3 * kAdamSeg < nblocks; b += 4 * kAdamSeg) {
            const float a0 = partial[(long long)b * kPartial + p];
            const float a1 = partial[(long long)(b + kAdamSeg) * kPartial + p];
            const float a2 = partial[(long long)(b + 2 * kAdamSeg) * kPartial + p];
            const float a3 = partial[(long long)(b + 3 * kAdamSeg) * kPartial + p];
            g = (((g + a0) + a1) + a2) + a3;
        }
        for (; b < nblocks; b += kAdamSeg) g += partial[(long long)b * kPartial + p];
    }
    sg[seg][pl] = g;
    __syncthreads();
    if (seg == 0) {
        const float gs = ((sg[0][pl] + sg[1][pl]) + sg[2][pl]) + sg[3][pl];
        tot[pl] = gs;
        if (p < kNP) {
            const float m = beta1 * M[p] + (1.0f - beta1) * gs;
            const float v = beta2 * V[p] + (1.0f - beta2) * gs * gs;
            M[p] = m;
            V[p] = v;
            const float mh = m / bc1, vh = v / bc2;
            P[p] -= lr * mh / (__builtin_sqrtf(vh) + eps);
        }
    }
    __syncthreads();
    if (seg == 0 && p == kNP && loss_out) {
        const float s = tot[pl], c = tot[pl + 1];  // kNP and kNP + 1 fall in the same block
        loss_out[0] = c > 0.0f ? s / c : 0.0f;
    }
}

__global__ void __launch_bounds__(256) deg_mlp_forward_kernel(const float* __restrict__ P, TrainCfg cfg,
                                                              const float* __restrict__ X, long long ld, long long n,
                                                              float* __restrict__ out) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float x[DEG_MLP_IN], h1[DEG_MLP_HID];
#pragma unroll
    for (int k = 0; k < DEG_MLP_IN; ++k) x[k] = (X[r * ld + k] - cfg.mean[k]) * cfg.inv_std[k];
#pragma unroll
    for (int m = 0; m < DEG_MLP_HID; ++m) {
        float s = P[DEG_MLP_OFF_B1 + m];
#pragma unroll
        for (int k = 0; k < DEG_MLP_IN; ++k) s = __builtin_fmaf(x[k], P[DEG_MLP_OFF_W1 + k * DEG_MLP_HID + m], s);
        h1[m] = __builtin_fmaxf(s, 0.0f);
    }
    float o = P[DEG_MLP_OFF_B3];
    for (int m = 0; m < DEG_MLP_HID; ++m) {
        float s = P[DEG_MLP_OFF_B2 + m];
#pragma unroll
        for (int k = 0; k < DEG_MLP_HID; ++k) s = __builtin_fmaf(h1[k], P[DEG_MLP_OFF_W2 + k * DEG_MLP_HID + m], s);
        o = __builtin_fmaf(P[DEG_MLP_OFF_W3 + m], __builtin_fmaxf(s, 0.0f), o);
    }
    out[r] = o;
}

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    const deg_mlp_bf2 v = __builtin_convertvector((deg_mlp_f2){f, 0.0f}, deg_mlp_bf2);
    return (unsigned short)(__builtin_bit_cast(unsigned int, v) & 0xffffu);
}

__global__ void __launch_bounds__(64) deg_mlp_pack_kernel(const float* __restrict__ P, TrainCfg cfg, float zlo,
                                                          float zhi, deg_mlp_packed* __restrict__ W) {
    const int l = threadIdx.x;
    const int g = l >> 4;
    // hid2(8 g + i): the layer-1 hidden unit a lane of group g holds as the i-th K value of layer 2
    auto hid2 = [](int gg, int i) { return i < 4 ? 4 * gg + i : 12 + 4 * gg + i; };
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (l < 32) {
            float w = 0.0f;
            if (i < DEG_MLP_IN) w = P[DEG_MLP_OFF_W1 + i * DEG_MLP_HID + l];
            else if (i == DEG_MLP_IN) w = P[DEG_MLP_OFF_B1 + l];
            W->a1[l][i] = bf16_rne(w);
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
            W->a2[r][l][i] = bf16_rne(P[DEG_MLP_OFF_W2 + hid2(g, i) * DEG_MLP_HID + 16 * r + (l & 15)]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) W->a3[l][i] = (l & 15) == 0 ? bf16_rne(P[DEG_MLP_OFF_W3 + hid2(g, i)]) : (unsigned short)0;
    if (l < 32) W->b2c[l >> 4][l & 15] = P[DEG_MLP_OFF_B2 + l];
    if (l < DEG_MLP_IN) {
        W->xscale[l] = cfg.inv_std[l];
        W->xshift[l] = -cfg.mean[l] * cfg.inv_std[l];
    }
    if (l == 0) {
        W->b3 = P[DEG_MLP_OFF_B3];
        W->zlo = zlo;
        W->zhi = zhi;
        W->magic = DEG_MLP_MAGIC;
    }
}

__global__ void __launch_bounds__(256) deg_mlp_infer_kernel(const deg_mlp_packed* __restrict__ W,
                                                            const float* __restrict__ X, long long ld, long long n,
                                                            float* __restrict__ out) {
    __shared__ deg_mlp_packed w;
    deg_mlp_stage(w, W);
    __syncthreads();
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float x[DEG_MLP_IN];
#pragma unroll
    for (int k = 0; k < DEG_MLP_IN; ++k) x[k] = (r < n) ? X[r * ld + k] : 0.0f;
    // the whole wavefront evaluates (rows past n carry zeros)
    const float z = deg_mlp_eval(w, x);
    if (r < n) out[r] = z;
}

TrainCfg make_cfg(const float* norm, float expo, float log2_rho, float ycap) {
    TrainCfg c;
    for (int k = 0; k < DEG_MLP_IN; ++k) {
        c.mean[k] = norm ? norm[k] : 0.0f;
        c.inv_std[k] = norm ? norm[DEG_MLP_IN + k] : 1.0f;
    }
    c.expo = expo;
    c.log2_rho = log2_rho;
    c.ycap = ycap;
    c.pad = 0.0f;
    return c;
}

int launch_status(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return deg_rng_fail(DEG_E_HIP, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
    return DEG_OK;
}

}  // namespace

extern "C" {

uint32_t deg_mlp_packed_size(void) { return (uint32_t)sizeof(deg_mlp_packed); }

int32_t deg_mlp_partial_cols(void) { return kPartial; }

int deg_mlp_grad(const float* rows, const int32_t* idx, int64_t nrows, const float* params, const float* norm,
                 float expo, float log2_rho, float ycap, float* partial, int32_t nblocks, void* stream) {
    if (!rows || !params || !partial) return deg_rng_fail(DEG_E_INVALID, "rows, params and partial are required");
    if (nrows < 1 || nblocks < 1 || nblocks > 65535) return deg_rng_fail(DEG_E_INVALID, "need nrows >= 1 and 1 <= nblocks <= 65535");
    const long long per_block = ((nrows + nblocks - 1) / nblocks + kTile - 1) / kTile * kTile;
    hipLaunchKernelGGL(deg_mlp_grad_kernel, dim3((unsigned)nblocks), dim3(64), 0, (hipStream_t)stream, rows, idx,
                       (long long)nrows, per_block, params, make_cfg(norm, expo, log2_rho, ycap), partial);
    return launch_status("deg_mlp_grad launch");
}

int deg_mlp_adam(float* params, float* m, float* v, const float* partial, int32_t nblocks, int32_t step, float lr,
                 float beta1, float beta2, float eps, float* loss_out, void* stream) {
    if (!params || !m || !v || !partial) return deg_rng_fail(DEG_E_INVALID, "params, m, v and partial are required");
    if (step < 1 || nblocks < 1) return deg_rng_fail(DEG_E_INVALID, "step and nblocks must be >= 1");
    double b1t = 1.0, b2t = 1.0;
    for (int i = 0; i < step && i < 100000; ++i) {
        b1t *= beta1;
        b2t *= beta2;
    }
    hipLaunchKernelGGL(deg_mlp_adam_kernel, dim3((kPartial + 63) / 64), dim3(256), 0, (hipStream_t)stream, params, m, v,
                       partial, (int)nblocks, lr, beta1, beta2, eps, (float)(1.0 - b1t), (float)(1.0 - b2t), loss_out);
    return launch_status("deg_mlp_adam launch");
}

int deg_mlp_forward(const float* params, const float* norm, const float* x, int64_t ld, int64_t n, float* out,
                    void* stream) {
    if (!params || !x || !out || ld < DEG_MLP_IN || n < 0) return deg_rng_fail(DEG_E_INVALID, "bad deg_mlp_forward arguments");
    if (n == 0) return DEG_OK;
    hipLaunchKernelGGL(deg_mlp_forward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       params, make_cfg(norm, 0.0f, 0.0f, 0.0f), x, (long long)ld, (long long)n, out);
    return launch_status("deg_mlp_forward launch");
}

int deg_mlp_pack(const float* params, const float* norm, float zlo, float zhi, void* packed, void* stream) {
    if (!params || !packed) return deg_rng_fail(DEG_E_INVALID, "params and packed are required");
    if (!(zlo < zhi)) return deg_rng_fail(DEG_E_INVALID, "need zlo < zhi");
    hipLaunchKernelGGL(deg_mlp_pack_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, params,
                       make_cfg(norm, 0.0f, 0.0f, 0.0f), zlo, zhi, (deg_mlp_packed*)packed);
    return launch_status("deg_mlp_pack launch");
}

int deg_mlp_infer(const void* packed, const float* x, int64_t ld, int64_t n, float* out, void* stream) {
    if (!packed || !x || !out || ld < DEG_MLP_IN || n < 0) return deg_rng_fail(DEG_E_INVALID, "bad deg_mlp_infer arguments");
    if (n == 0) return DEG_OK;
    hipLaunchKernelGGL(deg_mlp_infer_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const deg_mlp_packed*)packed, x, (long long)ld, (long long)n, out);
    return launch_status("deg_mlp_infer launch");
}

}  // extern "C"
