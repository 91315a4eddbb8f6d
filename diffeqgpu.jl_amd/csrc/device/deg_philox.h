// Counter-based normal draws for the SDE kernels: Philox4x32-10 and a deterministic Box-Muller.
//
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC11) maps a 128-bit counter and a 64-bit key to four 32-bit words; a draw is a pure function
// of (key, counter), so every trajectory and step reads its own numbers without carried state.
// The SDE kernels use key = the problem's 64-bit seed and counter = (step, block, trajectory
// id low, trajectory id high): normal c of a step is word pair (c % 4) / 2 of block c / 4,
// taking the cosine branch for even c and the sine branch for odd c.
//
// Box-Muller on a word pair (w0, w1): U1 = ((w0 >> 8) + 1) 2^-24 in (0, 1],
// R = sqrt(-2 ln U1), theta = 2 pi (w1 >> 8) 2^-24, normals (R cos theta, R sin theta).
// ln and sincos are evaluated from IEEE +, -, *, / and bit manipulation only (no library
// calls, no contraction), so a CPU build of the same expressions gives the same bits:
//   ln: U1 = 2^e m with m in [sqrt(1/2), sqrt(2)], ln m = 2 atanh(s), s = (m - 1) / (m + 1),
//       by its odd series to s^11; e ln 2 with ln 2 split into a 15-bit head and a tail.
//   sincos: the quadrant and the reflection about pi/4 are taken from the integer angle, so
//       the reduced argument x = r 2 pi 2^-24 (r <= 2^21) lies in [0, pi/4] with one rounding;
//       sin and cos by their Taylor series to x^9 and x^10.
// Relative error of the normals: a few Float32 ulps. The tail is cut at sqrt(2 * 24 ln 2) = 5.77.
//
// DEG_SDE_FASTNORM (contraction-allowed SDE kernels) evaluates the same transform of the same words
// with the hardware log2, square root and sine/cosine instead: the same normals to a few ulps, not
// bit for bit. deg_draw_normal always uses the exact form.
#pragma once

#ifndef DEG_SDE_FASTNORM
#define DEG_SDE_FASTNORM 0
#endif

#define DEG_PHILOX_M0 0xD2511F53u
#define DEG_PHILOX_M1 0xCD9E8D57u
#define DEG_PHILOX_W0 0x9E3779B9u
#define DEG_PHILOX_W1 0xBB67AE85u

struct deg_u4 {
    unsigned int x, y, z, w;
};

__device__ __forceinline__ deg_u4 deg_philox4x32_10(deg_u4 c, unsigned int k0, unsigned int k1) {
    _Pragma("unroll") for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += DEG_PHILOX_W0;
            k1 += DEG_PHILOX_W1;
        }
        const unsigned long long p0 = (unsigned long long)DEG_PHILOX_M0 * c.x;
        const unsigned long long p1 = (unsigned long long)DEG_PHILOX_M1 * c.z;
        const unsigned int hi0 = (unsigned int)(p0 >> 32), lo0 = (unsigned int)p0;
        const unsigned int hi1 = (unsigned int)(p1 >> 32), lo1 = (unsigned int)p1;
        c = deg_u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

// -2 ln U for U = ((w >> 8) + 1) 2^-24
__device__ __forceinline__ float deg_bm_m2log(unsigned int w) {
#pragma clang fp contract(off)
    const float u = (float)((w >> 8) + 1u) * 5.9604644775390625e-08f;
    const unsigned int b = __builtin_bit_cast(unsigned int, u);
    int e = (int)(b >> 23) - 127;
    float m = __builtin_bit_cast(float, (b & 0x007fffffu) | 0x3f800000u);
    if ((b & 0x007fffffu) > 0x003504f3u) {  // m > sqrt(2)
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float lnm = s * (2.0f + z * (0.666666687f + z * (0.400000006f + z * (0.285714298f + z * (0.222222224f + z * 0.181818187f)))));
    const float fe = (float)e;
    const float ln = fe * 0.693145752f + (fe * 1.42860677e-06f + lnm);
    return -2.0f * ln;
}

// (cos, sin) of 2 pi (w >> 8) 2^-24
__device__ __forceinline__ void deg_bm_sincos(unsigned int w, float& co, float& si) {
#pragma clang fp contract(off)
    const unsigned int k = w >> 8;
    const unsigned int q = k >> 22;
    unsigned int r = k & 0x3fffffu;
    const bool refl = r > 0x200000u;
    if (refl) r = 0x400000u - r;
    const float x = (float)r * 3.74507028e-07f;  // fl(2 pi) 2^-24
    const float x2 = x * x;
    const float sx = x * (1.0f + x2 * (-0.166666672f + x2 * (0.00833333377f + x2 * (-0.000198412701f + x2 * 2.75573188e-06f))));
    const float cx = 1.0f + x2 * (-0.5f + x2 * (0.0416666679f + x2 * (-0.00138888892f + x2 * (2.48015876e-05f + x2 * -2.75573188e-07f))));
    const float cp = refl ? sx : cx;
    const float sp = refl ? cx : sx;
    switch (q) {
        case 0: co = cp; si = sp; break;
        case 1: co = -sp; si = cp; break;
        case 2: co = -cp; si = -sp; break;
        default: co = sp; si = -cp; break;
    }
}

// Normals 2j and 2j+1 of a Philox block from its word pair (w0, w1).
__device__ __forceinline__ void deg_bm_pair(unsigned int w0, unsigned int w1, float& z0, float& z1) {
#pragma clang fp contract(off)
    const float rr = __builtin_sqrtf(deg_bm_m2log(w0));
    float co, si;
    deg_bm_sincos(w1, co, si);
    z0 = rr * co;
    z1 = rr * si;
}

// Hardware form of deg_bm_pair: v_log_f32 (log2: -2 ln U = -2 ln 2 log2 U), v_sqrt_f32, and v_cos_f32 /
// v_sin_f32, which take their argument in revolutions (sin(2 pi x)), so the angle needs no scaling.
__device__ __forceinline__ void deg_bm_pair_fast(unsigned int w0, unsigned int w1, float& z0, float& z1) {
    const float u = (float)((w0 >> 8) + 1u) * 5.9604644775390625e-08f;
    const float rr = __builtin_amdgcn_sqrtf(-1.38629436f * __builtin_amdgcn_logf(u));
    const float x = (float)(w1 >> 8) * 5.9604644775390625e-08f;
    z0 = rr * __builtin_amdgcn_cosf(x);
    z1 = rr * __builtin_amdgcn_sinf(x);
}

#if DEG_SDE_FASTNORM
#define DEG_BM_PAIR deg_bm_pair_fast
#else
#define DEG_BM_PAIR deg_bm_pair
#endif

// The M standard normals of one step of trajectory `id` under `seed`.
template <int M>
__device__ __forceinline__ void deg_step_normals(unsigned long long seed, unsigned long long id, unsigned int step,
                                                 float (&w)[M]) {
    _Pragma("unroll") for (int b = 0; b < (M + 3) / 4; ++b) {
        const deg_u4 r = deg_philox4x32_10(deg_u4{step, (unsigned int)b, (unsigned int)id, (unsigned int)(id >> 32)},
                                           (unsigned int)seed, (unsigned int)(seed >> 32));
        float z[4];
        DEG_BM_PAIR(r.x, r.y, z[0], z[1]);
        if (4 * b + 2 < M) DEG_BM_PAIR(r.z, r.w, z[2], z[3]);
        _Pragma("unroll") for (int j = 0; j < 4; ++j) if (4 * b + j < M) w[4 * b + j] = z[j];
    }
}
