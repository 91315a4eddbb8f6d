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
// Layout on the matrix cores. The layers are evaluated transposed, H^T = W^T X^T, so that the
// output fragment of one v_mfma_f32_32x32x8_bf16 is already the B-operand fragment of the
// next layer's K-slices:
//   A (32 x 8):  lane l holds A[m = l % 32][k = 4 (l / 32) + i], i = 0..3
//   B (8 x 32):  lane l holds B[k = 4 (l / 32) + i][n = l % 32]
//   D (32 x 32): lane l, register v holds D[m = 8 (v / 4) + 4 (l / 32) + v % 4][n = l % 32]
// Column n of block b is trajectory 32 b + n of the wavefront; K-slice s of the next layer
// is registers 4s..4s+3 of D. Layer 1 has K = 8: six normalised features, a constant 1
// (carrying b1) and a zero. Layer 3 (32 -> 1) is a VALU dot product over each lane's 16 rows
// plus one cross-lane add.
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

typedef short deg_mlp_s4 __attribute__((ext_vector_type(4)));
typedef float deg_mlp_f16 __attribute__((ext_vector_type(16)));
typedef float deg_mlp_f2 __attribute__((ext_vector_type(2)));
typedef __bf16 deg_mlp_bf2 __attribute__((ext_vector_type(2)));

// Packed inference weights (built by deg_mlp_pack from the FP32 master parameters).
struct deg_mlp_packed {
    unsigned short a1[64][4];     // a1[l][i] = bf16 W1x[k = 4 (l/32) + i][m = l % 32]; W1x rows 0..5 = W1, 6 = b1, 7 = 0
    unsigned short a2[4][64][4];  // a2[s][l][i] = bf16 W2[k = 8 s + 4 (l/32) + i][m = l % 32]
    float b2c[2][16];             // b2c[h][v] = b2[8 (v/4) + 4 h + v % 4]
    float w3c[2][16];             // w3c[h][v] = w3[8 (v/4) + 4 h + v % 4]
    float xshift[6];              // normalised feature = x * xscale + xshift
    float xscale[6];
    float b3;
    float zlo, zhi;               // clamp of the output
    unsigned int magic;
};

// two floats -> two bf16 (round to nearest even), packed low | high << 16
__device__ __forceinline__ unsigned int deg_mlp_pack2(float lo, float hi) {
    const deg_mlp_bf2 v = __builtin_convertvector((deg_mlp_f2){lo, hi}, deg_mlp_bf2);
    return __builtin_bit_cast(unsigned int, v);
}

__device__ __forceinline__ deg_mlp_s4 deg_mlp_frag(unsigned int lo, unsigned int hi) {
    return __builtin_bit_cast(deg_mlp_s4, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

__device__ __forceinline__ deg_mlp_s4 deg_mlp_relu_frag(const deg_mlp_f16& c, int s) {
    return deg_mlp_frag(deg_mlp_pack2(__builtin_fmaxf(c[4 * s + 0], 0.0f), __builtin_fmaxf(c[4 * s + 1], 0.0f)),
                        deg_mlp_pack2(__builtin_fmaxf(c[4 * s + 2], 0.0f), __builtin_fmaxf(c[4 * s + 3], 0.0f)));
}

// Evaluates the proposer for the calling lane's features x[0..5] and returns z. Every lane of
// the wavefront must make the call together (the matrix instructions read all 64 lanes);
// lanes without a trajectory pass finite placeholder features and ignore the result.
__device__ __forceinline__ float deg_mlp_eval(const deg_mlp_packed* __restrict__ W, const float* x) {
    const int l = (int)__lane_id();
    const int h = l >> 5;
    float xn[DEG_MLP_IN];
#pragma unroll
    for (int j = 0; j < DEG_MLP_IN; ++j) xn[j] = __builtin_fmaf(x[j], W->xscale[j], W->xshift[j]);
    // this lane supplies K-rows 4h..4h+3 of its own column in block h and of its partner's
    // (lane l ^ 32) column in block 1 - h; the partner's features arrive by one exchange
    const unsigned int f01 = deg_mlp_pack2(xn[0], xn[1]), f23 = deg_mlp_pack2(xn[2], xn[3]);
    const unsigned int f45 = deg_mlp_pack2(xn[4], xn[5]), f67 = deg_mlp_pack2(1.0f, 0.0f);
    const unsigned int own_lo = h ? f45 : f01, own_hi = h ? f67 : f23;
    const unsigned int snd_lo = h ? f01 : f45, snd_hi = h ? f23 : f67;
    const unsigned int rcv_lo = (unsigned int)__shfl_xor((int)snd_lo, 32);
    const unsigned int rcv_hi = (unsigned int)__shfl_xor((int)snd_hi, 32);
    const deg_mlp_s4 own = deg_mlp_frag(own_lo, own_hi), rcv = deg_mlp_frag(rcv_lo, rcv_hi);
    const deg_mlp_s4 B0 = h ? rcv : own;  // block 0: trajectories 0..31 of the wave
    const deg_mlp_s4 B1 = h ? own : rcv;  // block 1: trajectories 32..63

    // the two column blocks one after the other (fewer live accumulators)
    float part[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const deg_mlp_f16 zero = {};
        const deg_mlp_s4 A1 = __builtin_bit_cast(deg_mlp_s4, *(const unsigned long long*)W->a1[l]);
        const deg_mlp_f16 C = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A1, b ? B1 : B0, zero, 0, 0, 0);
        deg_mlp_s4 Bn[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) Bn[s] = deg_mlp_relu_frag(C, s);
        deg_mlp_f16 D;
#pragma unroll
        for (int v = 0; v < 16; ++v) D[v] = W->b2c[h][v];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const deg_mlp_s4 A2 = __builtin_bit_cast(deg_mlp_s4, *(const unsigned long long*)W->a2[s][l]);
            D = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A2, Bn[s], D, 0, 0, 0);
        }
        float acc = 0.0f;
#pragma unroll
        for (int v = 0; v < 16; ++v) acc = __builtin_fmaf(W->w3c[h][v], __builtin_fmaxf(D[v], 0.0f), acc);
        part[b] = acc;
    }
    const float p0 = part[0], p1 = part[1];
    // this lane's trajectory is column l % 32 of block h: its own half-sum plus the partner's
    const float mine = h ? p1 : p0, other = h ? p0 : p1;
    const float z = mine + __shfl_xor(other, 32) + W->b3;
    return __builtin_fminf(__builtin_fmaxf(z, W->zlo), W->zhi);
}
