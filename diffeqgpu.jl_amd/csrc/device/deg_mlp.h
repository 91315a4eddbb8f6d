// deg_mlp.h -- the learned step-size proposer: a 6 -> 32 -> 32 -> 1 ReLU MLP with BF16 weights
// and FP32 accumulation, evaluated on the matrix cores for the 64 trajectories of one
// wavefront per call. Shared by the JIT-compiled solver kernels and the library's own
// inference kernel (deg_mlp_infer), so the tested path and the solver path are the same code.
//
// Features (all base-2 logarithms; the solver computes them after each attempt):
//   x0 = log2 EEst_n, x1 = log2 EEst_{n-1}, x2 = log2 dt_n, x3 = log2 ||u||, x4 = log2 ||f(u)||,
//   x5 = log2(dt_n / dt_{n-1}).
// Output: a correction to the PI controller's proposal, clamped to [zlo, zhi]; the solver uses
// log2(dt_{n+1} / dt_n) = clamp(z_PI + z, log2 qmin, log2 qmax).
//
// Layout on the matrix cores: gfx950's v_mfma_f32_16x16x32_bf16 over four column tiles of 16
// trajectories, one tile after the other. The layers are evaluated transposed, H^T = W^T X^T, so the
// output fragment of one layer is already the B operand of the next. With g = l / 16 the lane group of
// lane l:
//   A (16 x 32):  lane l holds A[m = l % 16][k = 8 g + i], i = 0..7
//   B (32 x 16):  lane l holds B[k = 8 g + i][n = l % 16]
//   D (16 x 16):  lane l holds D[m = 4 g + v][n = l % 16], v = 0..3
// Column n of tile c is trajectory 16 c + n of the wavefront, i.e. lane 16 c + n.
// Layer 1 (K = 32, two row tiles of 16 hidden units): each lane puts its own normalised features, a
// constant 1 (carrying b1) and a zero at k = 8 g .. 8 g + 7 of B; the A operand of tile c holds W1 only
// in lane group c (zeros elsewhere), so tile c reads the features of lanes 16 c .. 16 c + 15 and no lane
// exchange is needed. Layer 2 (K = 32, two row tiles): lane l's B values are its layer-1 outputs after
// ReLU and BF16 rounding, the hidden units hid2(g, i) = 4 g + i (i < 4) and 12 + 4 g + i (i >= 4), and
// W2's rows are packed in that order; b2 is the accumulator's initial value. Layer 3 (32 -> 1) is one
// more matrix instruction whose A operand holds w3 in row 0 only: its B operand is layer 2's output
// packed the same way, and row 0 of the result (lanes 0..15, register 0) is the tile's 16 outputs.
// Only the lanes whose A values are not zero read them from LDS.
#pragma once

#define DEG_MLP_IN 6
#define DEG_MLP_HID 32
#define DEG_MLP_MAGIC 0x31504c4du
// FP32 master parameters (training): W1[k][m] (6 x 32), b1[32], W2[k][m] (32 x 32), b2[32],
// w3[32], b3
#define DEG_MLP_OFF_W1 0
#define DEG_MLP_OFF_B1 192
#define DEG_MLP_OFF_W2 224
#define DEG_MLP_OFF_B2 1248
#define DEG_MLP_OFF_W3 1280
#define DEG_MLP_OFF_B3 1312
#define DEG_MLP_NPARAM 1313
#ifndef DEG_STEPLOG_COLS
#define DEG_STEPLOG_COLS 12
#endif

typedef __bf16 deg_mlp_bf8 __attribute__((ext_vector_type(8)));
typedef float deg_mlp_f8 __attribute__((ext_vector_type(8)));
typedef float deg_mlp_f16 __attribute__((ext_vector_type(16)));

// Packed inference weights (built by deg_mlp_pack from the FP32 master parameters).
struct deg_mlp_packed {
    unsigned short a1[32][8];     // a1[m][i] = bf16 W1x[i][m]; W1x rows 0..5 = W1, 6 = b1, 7 = 0
    unsigned short a2[2][64][8];  // a2[r][l][i] = bf16 W2[hid2(l / 16, i)][16 r + l % 16]
    unsigned short a3[64][8];     // a3[l][i] = bf16 w3[hid2(l / 16, i)] for l % 16 == 0, else 0
    float b2c[2][16];             // b2c[r][j] = b2[16 r + j]
    float xshift[6];              // normalised feature = x * xscale + xshift
    float xscale[6];
    float b3;
    float zlo, zhi;               // clamp of the output
    unsigned int magic;
};

typedef float deg_mlp_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int deg_mlp_u4 __attribute__((ext_vector_type(4)));
typedef short deg_mlp_s2 __attribute__((ext_vector_type(2)));
typedef float deg_mlp_f2 __attribute__((ext_vector_type(2)));
typedef __bf16 deg_mlp_bf2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ deg_mlp_bf8 deg_mlp_load8(const unsigned short* p) {
    return __builtin_bit_cast(deg_mlp_bf8, *(const deg_mlp_u4*)p);
}

// two floats -> ReLU -> two bf16 (round to nearest even), packed low | high << 16. The ReLU acts on
// the rounded values as signed 16-bit integers (a bf16 with the sign bit set becomes +0), which is
// the same as rounding the ReLU's output.
__device__ __forceinline__ unsigned int deg_mlp_relu_pk(float lo, float hi) {
    const deg_mlp_bf2 v = __builtin_convertvector((deg_mlp_f2){lo, hi}, deg_mlp_bf2);
    const deg_mlp_s2 q = __builtin_elementwise_max(__builtin_bit_cast(deg_mlp_s2, v), (deg_mlp_s2){0, 0});
    return __builtin_bit_cast(unsigned int, q);
}

// the B operand of the next layer from two row tiles of a layer's output
__device__ __forceinline__ deg_mlp_bf8 deg_mlp_next_b(const deg_mlp_f4& t0, const deg_mlp_f4& t1) {
    deg_mlp_u4 q;
    q[0] = deg_mlp_relu_pk(t0[0], t0[1]);
    q[1] = deg_mlp_relu_pk(t0[2], t0[3]);
    q[2] = deg_mlp_relu_pk(t1[0], t1[1]);
    q[3] = deg_mlp_relu_pk(t1[2], t1[3]);
    return __builtin_bit_cast(deg_mlp_bf8, q);
}

// Copies the packed weights into LDS (every thread of the block takes part; a barrier follows).
__device__ __forceinline__ void deg_mlp_stage(deg_mlp_packed& dst, const deg_mlp_packed* __restrict__ src) {
    const unsigned int* s = (const unsigned int*)src;
    unsigned int* d = (unsigned int*)&dst;
    for (int j = threadIdx.x; j < (int)(sizeof(deg_mlp_packed) / 4); j += blockDim.x) d[j] = s[j];
}

// Evaluates the proposer for the calling lane's features x[0..5] and returns z. Every lane of
// the wavefront must make the call together (the matrix instructions read all 64 lanes);
// lanes without a trajectory pass finite placeholder features and ignore the result.
__device__ __forceinline__ float deg_mlp_eval(const deg_mlp_packed& W, const float* x) {
    const int l = (int)__lane_id();
    const int g = l >> 4;
    deg_mlp_f8 xf;
#pragma unroll
    for (int j = 0; j < DEG_MLP_IN; ++j) xf[j] = __builtin_fmaf(x[j], W.xscale[j], W.xshift[j]);
    xf[6] = 1.0f;
    xf[7] = 0.0f;
    const deg_mlp_bf8 B1 = __builtin_convertvector(xf, deg_mlp_bf8);
    // layer 2's weights are the same for every tile: read once
    const deg_mlp_bf8 A20 = deg_mlp_load8(W.a2[0][l]), A21 = deg_mlp_load8(W.a2[1][l]);
    const deg_mlp_f4 zero = {};
    float zt[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        // the scheduling barriers keep each tile's loads and temporaries inside the tile
        __builtin_amdgcn_sched_barrier(0);
        // layer 1's A operand is W1 in lane group c and zero elsewhere: only group c reads it
        deg_mlp_bf8 A10 = {}, A11 = {};
        if (g == c) {
            A10 = deg_mlp_load8(W.a1[l & 15]);
            A11 = deg_mlp_load8(W.a1[16 + (l & 15)]);
        }
        const deg_mlp_f4 C0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A10, B1, zero, 0, 0, 0);
        const deg_mlp_f4 C1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A11, B1, zero, 0, 0, 0);
        const deg_mlp_bf8 B2 = deg_mlp_next_b(C0, C1);
        const deg_mlp_f4 D0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A20, B2, *(const deg_mlp_f4*)&W.b2c[0][4 * g], 0, 0, 0);
        const deg_mlp_f4 D1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A21, B2, *(const deg_mlp_f4*)&W.b2c[1][4 * g], 0, 0, 0);
        // layer 3's A operand is w3 in row 0: only lanes 0, 16, 32 and 48 read it
        deg_mlp_bf8 A3 = {};
        if ((l & 15) == 0) A3 = deg_mlp_load8(W.a3[l]);
        const deg_mlp_f4 Z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A3, deg_mlp_next_b(D0, D1), zero, 0, 0, 0);
        zt[c] = Z[0];  // lanes 0..15: the outputs of trajectories 16 c .. 16 c + 15
    }
    __builtin_amdgcn_sched_barrier(0);
    // lane l's output is tile g's, in lane l % 16
    const int src = l & 15;
    const float v0 = __shfl(zt[0], src), v1 = __shfl(zt[1], src), v2 = __shfl(zt[2], src), v3 = __shfl(zt[3], src);
    const float z = (g == 0 ? v0 : g == 1 ? v1 : g == 2 ? v2 : v3) + W.b3;
    return __builtin_fminf(__builtin_fmaxf(z, W.zlo), W.zhi);
}
