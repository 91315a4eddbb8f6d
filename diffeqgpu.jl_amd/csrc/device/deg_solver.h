// Per-lane ensemble ODE solver for gfx950, compiled at run time by hiprtc together with
// a traced right-hand side. One trajectory per lane; state, stages and controller in
// VGPRs; tableau coefficients are compile-time literals.
//
// Reference behaviour restated here (DiffEqGPU src/ensemblegpukernel/):
//   kernels.jl:1-72 / :74-152          per-trajectory loop, initial save, endpoint rules
//   perform_step/gpu_tsit5_perform_step.jl:1-175, gpu_vern7_perform_step.jl:1-210,
//   perform_step/gpu_rosenbrock23_perform_step.jl:1-201, gpu_vern9_perform_step.jl:1-319   one step / one attempt
//   integrators/integrator_utils.jl:1-50   PI constants, savevalues!
//   integrators/nonstiff/interpolants.jl:28-370, :372-387, stiff/interpolants.jl   dense output
//
// Configuration macros supplied by the host before this file:
//   DEG_N, DEG_NP                state and parameter counts
//   DEG_METHOD                   0 Tsit5, 1 Vern7, 2 Rosenbrock23, 3 Vern9, 4 Rodas4, 5 Rodas5P,
//                                6 Kvaerno3, 7 Kvaerno5
//   DEG_ADAPTIVE                 0 fixed step, 1 adaptive (PI controller)
//   DEG_DETPOW                   controller powers via det_pow (bitwise-reproducible on CPU)
//   DEG_FASTPOW                  controller powers via hardware log2/exp2
//   DEG_KAHAN                    compensated accumulation of t
//   DEG_PERSISTENT               lanes refill from a device work counter
//   DEG_LEARNED                  MLP step-size proposal
//   DEG_SAVEAT_GLOBAL            read saveat times from global memory (more than fit the LDS)
//   DEG_WAVES_MIN                register cap: at least this many waves per SIMD (0: none)
//   DEG_PARITY                   1: the reference's exact Float32 operations (no contraction, IEEE
//                                division in the dense-output abscissa)
//   DEG_KERNEL                   kernel symbol name
//   DEG_BLOCK_MAX                launch-bounds block size
// and the functions deg_rhs(u, p, t, du) (+ deg_jac, deg_tgrad for Rosenbrock23, Rodas4, Rodas5P;
// deg_jac for Kvaerno3, Kvaerno5).
#pragma once

#define DEG_RET_SUCCESS 0
#define DEG_RET_DTMIN 1
#define DEG_RET_MAXITERS 2
#define DEG_RET_NONFINITE 3
#define DEG_RET_TERMINATED 4  // terminate! from a callback's affect
#define DEG_RET_FAILURE 5     // an event time before tprev (the reference raises an error there)

#define DEG_FOR(c) _Pragma("unroll") for (int c = 0; c < DEG_N; ++c)

// Stopping times and callbacks (integrators/integrator_utils.jl:52-479). DEG_TSTOPS comes
// with the launch flags; the traced source that defines the callbacks also defines
//   DEG_NDCB discrete callbacks:   bool  deg_dcb_cond(j, u, p, t),  int deg_dcb_affect(j, u, p, t)
//   DEG_NCCB continuous callbacks: float deg_ccb_cond(j, u, p, t),  int deg_ccb_affect(j, neg, u, p, t)
//   and for each continuous callback deg_ccb_mode(j) (bit 0 affect!, bit 1 affect_neg!,
//   bits 2-3 root finding: 0 none, 1 left, 2 right), deg_ccb_abstol(j), deg_ccb_nudge(j)
//   (repeat_nudge) and deg_ccb_dtrelax(j).
// An affect returns DEG_AFF_* bits.
#ifndef DEG_TSTOPS
#define DEG_TSTOPS 0
#endif
#ifndef DEG_NDCB
#define DEG_NDCB 0
#endif
#ifndef DEG_NCCB
#define DEG_NCCB 0
#endif
#define DEG_CALLBACKS ((DEG_NDCB + DEG_NCCB) > 0)
#define DEG_AFF_TERMINATE 1  // terminate!(integrator)
#define DEG_AFF_KEEP_FSAL 2  // u_modified!(integrator, false)

// ------------------------------------------------------------------------------------
// Julia Float32 semantics
// ------------------------------------------------------------------------------------
// Base.min / Base.max on IEEE floats propagate NaN.
__device__ __forceinline__ float jl_min(float a, float b) {
    return (a != a || b != b) ? a + b : ((b < a) ? b : a);
}
__device__ __forceinline__ float jl_max(float a, float b) {
    return (a != a || b != b) ? a + b : ((b > a) ? b : a);
}

// Deterministic x^y (x >= 0) from IEEE +,-,*,/ and exponent bit manipulation: log2 by the
// atanh series on a mantissa in [sqrt(1/2), sqrt(2)], exp2 by a degree-7 Taylor
// polynomial on [-ln2/2, ln2/2]. Relative error about 1e-6. Evaluated without contraction
// so that a CPU build of the same expression gives the same bits.
__device__ __forceinline__ float det_exp2i(int k) {
    return __builtin_bit_cast(float, (unsigned int)(k + 127) << 23);
}
__device__ __noinline__ float det_pow(float x, float y) {
#pragma clang fp contract(off)
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return 0.0f;
    if (x == __builtin_inff()) return __builtin_inff();
    unsigned int b = __builtin_bit_cast(unsigned int, x);
    int e = (int)((b >> 23) & 0xffu);
    if (e == 0) {
        x = x * 16777216.0f;
        b = __builtin_bit_cast(unsigned int, x);
        e = (int)((b >> 23) & 0xffu) - 24;
    }
    e -= 127;
    float m = __builtin_bit_cast(float, (b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float lnm = (2.0f * s) * (1.0f + s2 * (0.333333343f + s2 * (0.2f + s2 * (0.142857149f + s2 * 0.111111112f))));
    float l2 = (float)e + lnm * 1.44269502f;
    float z = y * l2;
    if (z > 128.0f) return __builtin_inff();
    if (z < -149.0f) return 0.0f;
    float n = __builtin_rintf(z);
    float g = (z - n) * 0.693147182f;
    float r = 1.0f + g * (1.0f + g * (0.5f + g * (0.166666672f + g * (0.0416666679f + g * (0.00833333377f + g * (0.00138888892f + g * 0.000198412701f))))));
    int ni = (int)n;
    int h1 = ni / 2;
    int h2 = ni - h1;
    return (r * det_exp2i(h1)) * det_exp2i(h2);
}

__device__ __forceinline__ float ctl_pow(float x, float y) {
#if DEG_DETPOW
    return det_pow(x, y);
#elif DEG_FASTPOW
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
#else
    return powf(x, y);
#endif
}

// DiffEqBase.ODE_DEFAULT_NORM for an SVector: sqrt(((x1^2 + x2^2) + ...) / n)
__device__ __forceinline__ float rms_norm(const float (&x)[DEG_N]) {
    float acc = x[0] * x[0];
    _Pragma("unroll") for (int c = 1; c < DEG_N; ++c) acc = acc + x[c] * x[c];
    return __builtin_sqrtf(acc / (float)DEG_N);
}

#ifndef DEG_DUAL_P
#define DEG_DUAL_P 0
#endif
// columns of the Jacobian deg_jac writes (J[r * DEG_JC + c], DEG_N rows): the values of a dual-number
// state (rows DEG_JC + r P + j then hold partial j of row r), else every state component
#if DEG_DUAL_P
#define DEG_JC DEG_DUAL_NV
#else
#define DEG_JC DEG_N
#endif

#if DEG_DUAL_P
// Dual-number state (DEG_DUAL_NV components, each with DEG_DUAL_P partials; the values are
// entries 0..NV-1 and component i's partials are NV + i P .. NV + i P + P - 1). The error
// scaling e ./ (abstol .+ max.(abs.(uprev), abs.(u)) * reltol) is formed in dual arithmetic with
// DiffRules' derivative rules (ForwardDiff's source of them): abs negates the partials where the
// value's sign bit is set (-0.0 included), max(x, y) takes x's partials only when x > y and y's
// otherwise (a tie or a NaN takes |u|'s), x * reltol scales the partials, e / s takes partials
// e' (1 / s) + s' (-(e / s^2)).
// The squared norm sums value^2 + (partial_1^2 + ... + partial_P^2) over the components, over
// NV (1 + P) entries in all (ODE_DEFAULT_NORM of a dual array). rcp: reciprocal multiply instead
// of the IEEE division of the value.
static_assert(DEG_N == DEG_DUAL_NV * (1 + DEG_DUAL_P), "a dual state holds NV values and NV * P partials");
template <bool rcp>
__device__ __forceinline__ float dual_scaled_error2(const float (&e)[DEG_N], const float (&uprev)[DEG_N],
                                                    const float (&u)[DEG_N], float abstol, float reltol) {
    float acc = 0.0f;
    _Pragma("unroll") for (int i = 0; i < DEG_DUAL_NV; ++i) {
        const float av = uprev[i], bv = u[i];
        const float aa = __builtin_fabsf(av), ba = __builtin_fabsf(bv);
        const bool nan = (aa != aa) || (ba != ba);
        const bool pick_u = !(aa > ba);
        const float mv = nan ? aa + ba : (pick_u ? ba : aa);
        const bool neg = pick_u ? __builtin_signbit(bv) : __builtin_signbit(av);
        const float sv = abstol + mv * reltol;
        const float ev = e[i];
        const float qv = rcp ? ev * __builtin_amdgcn_rcpf(sv) : ev / sv;
        const float inv = rcp ? __builtin_amdgcn_rcpf(sv) : 1.0f / sv;
        const float c = -(rcp ? ev * __builtin_amdgcn_rcpf(sv * sv) : ev / (sv * sv));
        float ps = 0.0f;
        _Pragma("unroll") for (int j = 0; j < DEG_DUAL_P; ++j) {
            const int q = DEG_DUAL_NV + i * DEG_DUAL_P + j;
            const float md = pick_u ? u[q] : uprev[q];
            const float sd = (neg ? -md : md) * reltol;
            const float qd = e[q] * inv + sd * c;
            ps = (j == 0) ? qd * qd : ps + qd * qd;
        }
        const float term = qv * qv + ps;
        acc = (i == 0) ? term : acc + term;
    }
    return acc;
}
#endif

// error scaling: tmp ./ (abstol .+ max.(abs.(uprev), abs.(u)) * reltol)
__device__ __forceinline__ float scaled_error(float (&e)[DEG_N], const float (&uprev)[DEG_N],
                                              const float (&u)[DEG_N], float abstol, float reltol) {
#if DEG_DUAL_P
    return __builtin_sqrtf(dual_scaled_error2<false>(e, uprev, u, abstol, reltol) / (float)DEG_N);
#else
    DEG_FOR(c) {
        float sc = abstol + jl_max(__builtin_fabsf(uprev[c]), __builtin_fabsf(u[c])) * reltol;
        e[c] = e[c] / sc;
    }
    return rms_norm(e);
#endif
}

// Fast error norm (DEG_FASTPOW): the squared norm E2 = EEst^2 with reciprocal-multiply
// scaling. The controller then works on log2(E2) (see deg_fast_control).
__device__ __forceinline__ float scaled_error2_fast(float (&e)[DEG_N], const float (&uprev)[DEG_N],
                                                    const float (&u)[DEG_N], float abstol, float reltol) {
#if DEG_DUAL_P
    return dual_scaled_error2<true>(e, uprev, u, abstol, reltol) * (1.0f / (float)DEG_N);
#endif
    float acc = 0.0f;
    DEG_FOR(c) {
        const float sc = abstol + __builtin_fmaxf(__builtin_fabsf(uprev[c]), __builtin_fabsf(u[c])) * reltol;
        const float r = e[c] * __builtin_amdgcn_rcpf(sc);
        acc = acc + r * r;
    }
    return acc * (1.0f / (float)DEG_N);
}

// ------------------------------------------------------------------------------------
// PI controller (integrator_utils.jl:1-11; gpu_tsit5_perform_step.jl:124-137)
// ------------------------------------------------------------------------------------
#if DEG_METHOD == 0
#define DEG_ORDER 5
#elif DEG_METHOD == 1
#define DEG_ORDER 7
#elif DEG_METHOD == 3
#define DEG_ORDER 9
#elif DEG_METHOD == 4
#define DEG_ORDER 4
#elif DEG_METHOD == 5
#define DEG_ORDER 5
#elif DEG_METHOD == 6
#define DEG_ORDER 3
#elif DEG_METHOD == 7
#define DEG_ORDER 5
#else
#define DEG_ORDER 2
#endif
#define DEG_BETA1 ((float)(7.0 / (10.0 * DEG_ORDER)))
#define DEG_BETA2 ((float)(2.0 / (5.0 * DEG_ORDER)))
#define DEG_QMAX 10.0f
#define DEG_QMIN ((float)(1.0 / 5.0))
#define DEG_GAMMA ((float)(9.0 / 10.0))
#define DEG_QOLDINIT ((float)1.0e-4)
#define DEG_DTMIN 1.0e-14f
// a stopping time within 100 eps(Float32) past the end of the step is taken by it
#define DEG_TSTOP_SLACK (100.0f * 1.1920928955078125e-7f)

// ------------------------------------------------------------------------------------
// Tsit5 (gpu_tsit5_perform_step.jl:36-49, :104-122; interpolants.jl:372-387)
// ------------------------------------------------------------------------------------
#if DEG_METHOD == 0
#define DEG_NK 7
__device__ __forceinline__ void deg_attempt(const float (&uprev)[DEG_N], const float* p, float t, float dt,
                                            float (&k)[DEG_NK + 1][DEG_N], float (&unew)[DEG_N]) {
    float tmp[DEG_N];
    const float a = dt * T5_a21;
    DEG_FOR(c) tmp[c] = uprev[c] + a * k[1][c];
    deg_rhs(tmp, p, t + T5_c1 * dt, k[2]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (T5_a31 * k[1][c] + T5_a32 * k[2][c]);
    deg_rhs(tmp, p, t + T5_c2 * dt, k[3]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (T5_a41 * k[1][c] + T5_a42 * k[2][c] + T5_a43 * k[3][c]);
    deg_rhs(tmp, p, t + T5_c3 * dt, k[4]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (T5_a51 * k[1][c] + T5_a52 * k[2][c] + T5_a53 * k[3][c] + T5_a54 * k[4][c]);
    deg_rhs(tmp, p, t + T5_c4 * dt, k[5]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (T5_a61 * k[1][c] + T5_a62 * k[2][c] + T5_a63 * k[3][c] + T5_a64 * k[4][c] +
                                         T5_a65 * k[5][c]);
    deg_rhs(tmp, p, t + dt, k[6]);
    DEG_FOR(c) unew[c] = uprev[c] + dt * (T5_a71 * k[1][c] + T5_a72 * k[2][c] + T5_a73 * k[3][c] + T5_a74 * k[4][c] +
                                          T5_a75 * k[5][c] + T5_a76 * k[6][c]);
    deg_rhs(unew, p, t + dt, k[7]);
}
__device__ __forceinline__ float deg_error(float dt, const float (&k)[DEG_NK + 1][DEG_N], const float (&uprev)[DEG_N],
                                           const float (&unew)[DEG_N], float abstol, float reltol) {
    float e[DEG_N];
    DEG_FOR(c) e[c] = dt * (T5_bt1 * k[1][c] + T5_bt2 * k[2][c] + T5_bt3 * k[3][c] + T5_bt4 * k[4][c] +
                            T5_bt5 * k[5][c] + T5_bt6 * k[6][c] + T5_bt7 * k[7][c]);
#if DEG_FASTPOW
    return scaled_error2_fast(e, uprev, unew, abstol, reltol);
#else
    return scaled_error(e, uprev, unew, abstol, reltol);
#endif
}
// y0 + dt * (b1(th) k1 + ... + b7(th) k7) with b_i evaluated by fused Horner steps
__device__ __forceinline__ void deg_interp(float th, float dt, float tprev, const float* p, const float (&y0)[DEG_N],
                                           const float (&y1)[DEG_N],
                                           const float (&k)[DEG_NK + 1][DEG_N], float (&out)[DEG_N]) {
    (void)tprev;
    (void)p;
    (void)y1;
    const float b1 = th * __builtin_fmaf(th, __builtin_fmaf(th, __builtin_fmaf(th, T5_r14, T5_r13), T5_r12), T5_r11);
#if !DEG_PARITY
    // b2..b7 are th^2 times a quadratic: apply th^2 (and dt) once per component instead of twice per
    // weight (12 fewer instructions per save point; rounding differs from the reference's form)
    const float q2 = __builtin_fmaf(th, __builtin_fmaf(th, T5_r24, T5_r23), T5_r22);
    const float q3 = __builtin_fmaf(th, __builtin_fmaf(th, T5_r34, T5_r33), T5_r32);
    const float q4 = __builtin_fmaf(th, __builtin_fmaf(th, T5_r44, T5_r43), T5_r42);
    const float q5 = __builtin_fmaf(th, __builtin_fmaf(th, T5_r54, T5_r53), T5_r52);
    const float q6 = __builtin_fmaf(th, __builtin_fmaf(th, T5_r64, T5_r63), T5_r62);
    const float q7 = __builtin_fmaf(th, __builtin_fmaf(th, T5_r74, T5_r73), T5_r72);
    const float db1 = dt * b1;
    const float dth2 = dt * (th * th);
    DEG_FOR(c) {
        float s = q2 * k[2][c];
        s = __builtin_fmaf(q3, k[3][c], s);
        s = __builtin_fmaf(q4, k[4][c], s);
        s = __builtin_fmaf(q5, k[5][c], s);
        s = __builtin_fmaf(q6, k[6][c], s);
        s = __builtin_fmaf(q7, k[7][c], s);
        out[c] = __builtin_fmaf(dth2, s, __builtin_fmaf(db1, k[1][c], y0[c]));
    }
    return;
#endif
    const float b2 = th * (th * __builtin_fmaf(th, __builtin_fmaf(th, T5_r24, T5_r23), T5_r22));
    const float b3 = th * (th * __builtin_fmaf(th, __builtin_fmaf(th, T5_r34, T5_r33), T5_r32));
    const float b4 = th * (th * __builtin_fmaf(th, __builtin_fmaf(th, T5_r44, T5_r43), T5_r42));
    const float b5 = th * (th * __builtin_fmaf(th, __builtin_fmaf(th, T5_r54, T5_r53), T5_r52));
    const float b6 = th * (th * __builtin_fmaf(th, __builtin_fmaf(th, T5_r64, T5_r63), T5_r62));
    const float b7 = th * (th * __builtin_fmaf(th, __builtin_fmaf(th, T5_r74, T5_r73), T5_r72));
    DEG_FOR(c) out[c] = y0[c] + dt * (b1 * k[1][c] + b2 * k[2][c] + b3 * k[3][c] + b4 * k[4][c] + b5 * k[5][c] +
                                      b6 * k[6][c] + b7 * k[7][c]);
}
#endif

// ------------------------------------------------------------------------------------
// Vern7 (gpu_vern7_perform_step.jl:30-56, :117-155); not FSAL: k1 recomputed per attempt
// ------------------------------------------------------------------------------------
#if DEG_METHOD == 1
#define DEG_NK 10
__device__ __forceinline__ void deg_attempt(const float (&uprev)[DEG_N], const float* p, float t, float dt,
                                            float (&k)[DEG_NK + 1][DEG_N], float (&unew)[DEG_N]) {
    float tmp[DEG_N];
    const float a = dt * V7_a021;
    DEG_FOR(c) tmp[c] = uprev[c] + a * k[1][c];
    deg_rhs(tmp, p, t + V7_c2 * dt, k[2]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V7_a031 * k[1][c] + V7_a032 * k[2][c]);
    deg_rhs(tmp, p, t + V7_c3 * dt, k[3]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V7_a041 * k[1][c] + V7_a043 * k[3][c]);
    deg_rhs(tmp, p, t + V7_c4 * dt, k[4]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V7_a051 * k[1][c] + V7_a053 * k[3][c] + V7_a054 * k[4][c]);
    deg_rhs(tmp, p, t + V7_c5 * dt, k[5]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V7_a061 * k[1][c] + V7_a063 * k[3][c] + V7_a064 * k[4][c] + V7_a065 * k[5][c]);
    deg_rhs(tmp, p, t + V7_c6 * dt, k[6]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V7_a071 * k[1][c] + V7_a073 * k[3][c] + V7_a074 * k[4][c] + V7_a075 * k[5][c] +
                                         V7_a076 * k[6][c]);
    deg_rhs(tmp, p, t + V7_c7 * dt, k[7]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V7_a081 * k[1][c] + V7_a083 * k[3][c] + V7_a084 * k[4][c] + V7_a085 * k[5][c] +
                                         V7_a086 * k[6][c] + V7_a087 * k[7][c]);
    deg_rhs(tmp, p, t + V7_c8 * dt, k[8]);
    float g10[DEG_N];
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V7_a091 * k[1][c] + V7_a093 * k[3][c] + V7_a094 * k[4][c] + V7_a095 * k[5][c] +
                                         V7_a096 * k[6][c] + V7_a097 * k[7][c] + V7_a098 * k[8][c]);
    DEG_FOR(c) g10[c] = uprev[c] + dt * (V7_a101 * k[1][c] + V7_a103 * k[3][c] + V7_a104 * k[4][c] + V7_a105 * k[5][c] +
                                         V7_a106 * k[6][c] + V7_a107 * k[7][c]);
    deg_rhs(tmp, p, t + dt, k[9]);
    deg_rhs(g10, p, t + dt, k[10]);
    DEG_FOR(c) unew[c] = uprev[c] + dt * (V7_b1 * k[1][c] + V7_b4 * k[4][c] + V7_b5 * k[5][c] + V7_b6 * k[6][c] +
                                          V7_b7 * k[7][c] + V7_b8 * k[8][c] + V7_b9 * k[9][c]);
}
__device__ __forceinline__ float deg_error(float dt, const float (&k)[DEG_NK + 1][DEG_N], const float (&uprev)[DEG_N],
                                           const float (&unew)[DEG_N], float abstol, float reltol) {
    float e[DEG_N];
    DEG_FOR(c) e[c] = dt * (V7_btilde1 * k[1][c] + V7_btilde4 * k[4][c] + V7_btilde5 * k[5][c] + V7_btilde6 * k[6][c] +
                            V7_btilde7 * k[7][c] + V7_btilde8 * k[8][c] + V7_btilde9 * k[9][c] +
                            V7_btilde10 * k[10][c]);
#if DEG_FASTPOW
    return scaled_error2_fast(e, uprev, unew, abstol, reltol);
#else
    return scaled_error(e, uprev, unew, abstol, reltol);
#endif
}
// Vern7 dense output (interpolants.jl:28-190): extra stages k11..k16 from uprev, then
// y0 + th*step_sum + th*(th-1)*correction with q_i(th) from the synthetic division of the
// r coefficients. Scalar muladds (stage times, Horner steps) are fused.
__device__ __forceinline__ float v7q(float th, float r2, float r3, float r4, float r5, float r6, float r7) {
    const float s5 = r7;
    const float s4 = r6 + s5;
    const float s3 = r5 + s4;
    const float s2 = r4 + s3;
    const float s1 = r3 + s2;
    const float c0 = r2 + s1;
    return __builtin_fmaf(th, __builtin_fmaf(th, __builtin_fmaf(th, __builtin_fmaf(th, __builtin_fmaf(th, s5, s4), s3), s2), s1), c0);
}
#define V7Q(i) v7q(th, V7_r##i##2, V7_r##i##3, V7_r##i##4, V7_r##i##5, V7_r##i##6, V7_r##i##7)
__device__ __noinline__ void deg_interp_ool(float th, float dt, float tprev, const float* p, const float (&y0)[DEG_N],
                                               const float (&y1)[DEG_N],
                                               const float (&k)[DEG_NK + 1][DEG_N], float (&out)[DEG_N]) {
    (void)y1;
    float tmp[DEG_N], k11[DEG_N], k12[DEG_N], k13[DEG_N], k14[DEG_N], k15[DEG_N], k16[DEG_N];
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V7_a1101 * k[1][c] + V7_a1104 * k[4][c] + V7_a1105 * k[5][c] + V7_a1106 * k[6][c] +
                                      V7_a1107 * k[7][c] + V7_a1108 * k[8][c] + V7_a1109 * k[9][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V7_c11, dt, tprev), k11);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V7_a1201 * k[1][c] + V7_a1204 * k[4][c] + V7_a1205 * k[5][c] + V7_a1206 * k[6][c] +
                                      V7_a1207 * k[7][c] + V7_a1208 * k[8][c] + V7_a1209 * k[9][c] + V7_a1211 * k11[c]);
    deg_rhs(tmp, p, __builtin_fmaf(V7_c12, dt, tprev), k12);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V7_a1301 * k[1][c] + V7_a1304 * k[4][c] + V7_a1305 * k[5][c] + V7_a1306 * k[6][c] +
                                      V7_a1307 * k[7][c] + V7_a1308 * k[8][c] + V7_a1309 * k[9][c] + V7_a1311 * k11[c] +
                                      V7_a1312 * k12[c]);
    deg_rhs(tmp, p, __builtin_fmaf(V7_c13, dt, tprev), k13);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V7_a1401 * k[1][c] + V7_a1404 * k[4][c] + V7_a1405 * k[5][c] + V7_a1406 * k[6][c] +
                                      V7_a1407 * k[7][c] + V7_a1408 * k[8][c] + V7_a1409 * k[9][c] + V7_a1411 * k11[c] +
                                      V7_a1412 * k12[c] + V7_a1413 * k13[c]);
    deg_rhs(tmp, p, __builtin_fmaf(V7_c14, dt, tprev), k14);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V7_a1501 * k[1][c] + V7_a1504 * k[4][c] + V7_a1505 * k[5][c] + V7_a1506 * k[6][c] +
                                      V7_a1507 * k[7][c] + V7_a1508 * k[8][c] + V7_a1509 * k[9][c] + V7_a1511 * k11[c] +
                                      V7_a1512 * k12[c] + V7_a1513 * k13[c]);
    deg_rhs(tmp, p, __builtin_fmaf(V7_c15, dt, tprev), k15);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V7_a1601 * k[1][c] + V7_a1604 * k[4][c] + V7_a1605 * k[5][c] + V7_a1606 * k[6][c] +
                                      V7_a1607 * k[7][c] + V7_a1608 * k[8][c] + V7_a1609 * k[9][c] + V7_a1611 * k11[c] +
                                      V7_a1612 * k12[c] + V7_a1613 * k13[c]);
    deg_rhs(tmp, p, __builtin_fmaf(V7_c16, dt, tprev), k16);
    const float q1 = V7Q(01), q4 = V7Q(04), q5 = V7Q(05), q6 = V7Q(06), q7 = V7Q(07), q8 = V7Q(08), q9 = V7Q(09);
    const float q11 = V7Q(11), q12 = V7Q(12), q13 = V7Q(13), q14 = V7Q(14), q15 = V7Q(15), q16 = V7Q(16);
    const float thm = th * (th - 1.0f);
    DEG_FOR(c) {
        const float step_sum = dt * (V7_b1 * k[1][c] + V7_b4 * k[4][c] + V7_b5 * k[5][c] + V7_b6 * k[6][c] +
                                     V7_b7 * k[7][c] + V7_b8 * k[8][c] + V7_b9 * k[9][c]);
        const float corr = dt * (k[1][c] * q1 + k[4][c] * q4 + k[5][c] * q5 + k[6][c] * q6 + k[7][c] * q7 +
                                 k[8][c] * q8 + k[9][c] * q9 + k11[c] * q11 + k12[c] * q12 + k13[c] * q13 +
                                 k14[c] * q14 + k15[c] * q15 + k16[c] * q16);
        out[c] = (y0[c] + th * step_sum) + thm * corr;
    }
}
#endif

// ------------------------------------------------------------------------------------
// Vern9 (gpu_vern9_perform_step.jl:36-105, :166-240); not FSAL: k1 recomputed per attempt
// ------------------------------------------------------------------------------------
#if DEG_METHOD == 3
#define DEG_NK 16
__device__ __forceinline__ void deg_attempt(const float (&uprev)[DEG_N], const float* p, float t, float dt,
                                            float (&k)[DEG_NK + 1][DEG_N], float (&unew)[DEG_N]) {
    float tmp[DEG_N];
    const float a = dt * V9_a0201;
    DEG_FOR(c) tmp[c] = uprev[c] + a * k[1][c];
    deg_rhs(tmp, p, t + V9_c1 * dt, k[2]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a0301 * k[1][c] + V9_a0302 * k[2][c]);
    deg_rhs(tmp, p, t + V9_c2 * dt, k[3]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a0401 * k[1][c] + V9_a0403 * k[3][c]);
    deg_rhs(tmp, p, t + V9_c3 * dt, k[4]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a0501 * k[1][c] + V9_a0503 * k[3][c] + V9_a0504 * k[4][c]);
    deg_rhs(tmp, p, t + V9_c4 * dt, k[5]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a0601 * k[1][c] + V9_a0604 * k[4][c] + V9_a0605 * k[5][c]);
    deg_rhs(tmp, p, t + V9_c5 * dt, k[6]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a0701 * k[1][c] + V9_a0704 * k[4][c] + V9_a0705 * k[5][c] + V9_a0706 * k[6][c]);
    deg_rhs(tmp, p, t + V9_c6 * dt, k[7]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a0801 * k[1][c] + V9_a0806 * k[6][c] + V9_a0807 * k[7][c]);
    deg_rhs(tmp, p, t + V9_c7 * dt, k[8]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a0901 * k[1][c] + V9_a0906 * k[6][c] + V9_a0907 * k[7][c] + V9_a0908 * k[8][c]);
    deg_rhs(tmp, p, t + V9_c8 * dt, k[9]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a1001 * k[1][c] + V9_a1006 * k[6][c] + V9_a1007 * k[7][c] + V9_a1008 * k[8][c] + V9_a1009 * k[9][c]);
    deg_rhs(tmp, p, t + V9_c9 * dt, k[10]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a1101 * k[1][c] + V9_a1106 * k[6][c] + V9_a1107 * k[7][c] + V9_a1108 * k[8][c] + V9_a1109 * k[9][c] + V9_a1110 * k[10][c]);
    deg_rhs(tmp, p, t + V9_c10 * dt, k[11]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a1201 * k[1][c] + V9_a1206 * k[6][c] + V9_a1207 * k[7][c] + V9_a1208 * k[8][c] + V9_a1209 * k[9][c] + V9_a1210 * k[10][c] + V9_a1211 * k[11][c]);
    deg_rhs(tmp, p, t + V9_c11 * dt, k[12]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a1301 * k[1][c] + V9_a1306 * k[6][c] + V9_a1307 * k[7][c] + V9_a1308 * k[8][c] + V9_a1309 * k[9][c] + V9_a1310 * k[10][c] + V9_a1311 * k[11][c] + V9_a1312 * k[12][c]);
    deg_rhs(tmp, p, t + V9_c12 * dt, k[13]);
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a1401 * k[1][c] + V9_a1406 * k[6][c] + V9_a1407 * k[7][c] + V9_a1408 * k[8][c] + V9_a1409 * k[9][c] + V9_a1410 * k[10][c] + V9_a1411 * k[11][c] + V9_a1412 * k[12][c] + V9_a1413 * k[13][c]);
    deg_rhs(tmp, p, t + V9_c13 * dt, k[14]);
    float g16[DEG_N];
    DEG_FOR(c) tmp[c] = uprev[c] + dt * (V9_a1501 * k[1][c] + V9_a1506 * k[6][c] + V9_a1507 * k[7][c] + V9_a1508 * k[8][c] + V9_a1509 * k[9][c] + V9_a1510 * k[10][c] + V9_a1511 * k[11][c] + V9_a1512 * k[12][c] + V9_a1513 * k[13][c] + V9_a1514 * k[14][c]);
    DEG_FOR(c) g16[c] = uprev[c] + dt * (V9_a1601 * k[1][c] + V9_a1606 * k[6][c] + V9_a1607 * k[7][c] + V9_a1608 * k[8][c] + V9_a1609 * k[9][c] + V9_a1610 * k[10][c] + V9_a1611 * k[11][c] + V9_a1612 * k[12][c] + V9_a1613 * k[13][c]);
    deg_rhs(tmp, p, t + dt, k[15]);
    deg_rhs(g16, p, t + dt, k[16]);
    DEG_FOR(c) unew[c] = uprev[c] + dt * (V9_b1 * k[1][c] + V9_b8 * k[8][c] + V9_b9 * k[9][c] + V9_b10 * k[10][c] + V9_b11 * k[11][c] + V9_b12 * k[12][c] + V9_b13 * k[13][c] + V9_b14 * k[14][c] + V9_b15 * k[15][c]);
}
__device__ __forceinline__ float deg_error(float dt, const float (&k)[DEG_NK + 1][DEG_N], const float (&uprev)[DEG_N],
                                           const float (&unew)[DEG_N], float abstol, float reltol) {
    float e[DEG_N];
    DEG_FOR(c) e[c] = dt * (V9_btilde1 * k[1][c] + V9_btilde8 * k[8][c] + V9_btilde9 * k[9][c] + V9_btilde10 * k[10][c] + V9_btilde11 * k[11][c] + V9_btilde12 * k[12][c] + V9_btilde13 * k[13][c] + V9_btilde14 * k[14][c] + V9_btilde15 * k[15][c] + V9_btilde16 * k[16][c]);
#if DEG_FASTPOW
    return scaled_error2_fast(e, uprev, unew, abstol, reltol);
#else
    return scaled_error(e, uprev, unew, abstol, reltol);
#endif
}
// b_i(th) by fused Horner steps: th * H(r_i1..r_i9) for stage 1, th^2 * H(r_i2..r_i9) otherwise
#define V9B1(th) ((th) * V9H8(th, V9_r011, V9_r012, V9_r013, V9_r014, V9_r015, V9_r016, V9_r017, V9_r018, V9_r019))
#define V9B(i, th) ((th) * ((th) * V9H7(th, V9_r##i##2, V9_r##i##3, V9_r##i##4, V9_r##i##5, V9_r##i##6, V9_r##i##7, V9_r##i##8, V9_r##i##9)))
__device__ __forceinline__ float V9H7(float th, float r2, float r3, float r4, float r5, float r6, float r7, float r8, float r9) {
    float h = __builtin_fmaf(th, r9, r8);
    h = __builtin_fmaf(th, h, r7);
    h = __builtin_fmaf(th, h, r6);
    h = __builtin_fmaf(th, h, r5);
    h = __builtin_fmaf(th, h, r4);
    h = __builtin_fmaf(th, h, r3);
    return __builtin_fmaf(th, h, r2);
}
__device__ __forceinline__ float V9H8(float th, float r1, float r2, float r3, float r4, float r5, float r6, float r7, float r8, float r9) {
    return __builtin_fmaf(th, V9H7(th, r2, r3, r4, r5, r6, r7, r8, r9), r1);
}
// Vern9 dense output (interpolants.jl:192-370): extra stages 17..26 from uprev and the step's
// stages 1, 8..15, then y0 + dt * (sum of b_i(th) k_i). Scalar muladds (stage times, Horner
// steps) are fused; vector sums are unfused left folds.
__device__ __noinline__ void deg_interp_ool(float th, float dt, float tprev, const float* p, const float (&y0)[DEG_N],
                                               const float (&y1)[DEG_N],
                                               const float (&k)[DEG_NK + 1][DEG_N], float (&out)[DEG_N]) {
    (void)y1;
    float tmp[DEG_N], x[10][DEG_N];  // extra stages 17..26
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a1701 * k[1][c] + V9_a1708 * k[8][c] + V9_a1709 * k[9][c] + V9_a1710 * k[10][c] + V9_a1711 * k[11][c] + V9_a1712 * k[12][c] + V9_a1713 * k[13][c] + V9_a1714 * k[14][c] + V9_a1715 * k[15][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c17, dt, tprev), x[0]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a1801 * k[1][c] + V9_a1808 * k[8][c] + V9_a1809 * k[9][c] + V9_a1810 * k[10][c] + V9_a1811 * k[11][c] + V9_a1812 * k[12][c] + V9_a1813 * k[13][c] + V9_a1814 * k[14][c] + V9_a1815 * k[15][c] + V9_a1817 * x[0][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c18, dt, tprev), x[1]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a1901 * k[1][c] + V9_a1908 * k[8][c] + V9_a1909 * k[9][c] + V9_a1910 * k[10][c] + V9_a1911 * k[11][c] + V9_a1912 * k[12][c] + V9_a1913 * k[13][c] + V9_a1914 * k[14][c] + V9_a1915 * k[15][c] + V9_a1917 * x[0][c] + V9_a1918 * x[1][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c19, dt, tprev), x[2]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a2001 * k[1][c] + V9_a2008 * k[8][c] + V9_a2009 * k[9][c] + V9_a2010 * k[10][c] + V9_a2011 * k[11][c] + V9_a2012 * k[12][c] + V9_a2013 * k[13][c] + V9_a2014 * k[14][c] + V9_a2015 * k[15][c] + V9_a2017 * x[0][c] + V9_a2018 * x[1][c] + V9_a2019 * x[2][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c20, dt, tprev), x[3]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a2101 * k[1][c] + V9_a2108 * k[8][c] + V9_a2109 * k[9][c] + V9_a2110 * k[10][c] + V9_a2111 * k[11][c] + V9_a2112 * k[12][c] + V9_a2113 * k[13][c] + V9_a2114 * k[14][c] + V9_a2115 * k[15][c] + V9_a2117 * x[0][c] + V9_a2118 * x[1][c] + V9_a2119 * x[2][c] + V9_a2120 * x[3][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c21, dt, tprev), x[4]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a2201 * k[1][c] + V9_a2208 * k[8][c] + V9_a2209 * k[9][c] + V9_a2210 * k[10][c] + V9_a2211 * k[11][c] + V9_a2212 * k[12][c] + V9_a2213 * k[13][c] + V9_a2214 * k[14][c] + V9_a2215 * k[15][c] + V9_a2217 * x[0][c] + V9_a2218 * x[1][c] + V9_a2219 * x[2][c] + V9_a2220 * x[3][c] + V9_a2221 * x[4][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c22, dt, tprev), x[5]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a2301 * k[1][c] + V9_a2308 * k[8][c] + V9_a2309 * k[9][c] + V9_a2310 * k[10][c] + V9_a2311 * k[11][c] + V9_a2312 * k[12][c] + V9_a2313 * k[13][c] + V9_a2314 * k[14][c] + V9_a2315 * k[15][c] + V9_a2317 * x[0][c] + V9_a2318 * x[1][c] + V9_a2319 * x[2][c] + V9_a2320 * x[3][c] + V9_a2321 * x[4][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c23, dt, tprev), x[6]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a2401 * k[1][c] + V9_a2408 * k[8][c] + V9_a2409 * k[9][c] + V9_a2410 * k[10][c] + V9_a2411 * k[11][c] + V9_a2412 * k[12][c] + V9_a2413 * k[13][c] + V9_a2414 * k[14][c] + V9_a2415 * k[15][c] + V9_a2417 * x[0][c] + V9_a2418 * x[1][c] + V9_a2419 * x[2][c] + V9_a2420 * x[3][c] + V9_a2421 * x[4][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c24, dt, tprev), x[7]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a2501 * k[1][c] + V9_a2508 * k[8][c] + V9_a2509 * k[9][c] + V9_a2510 * k[10][c] + V9_a2511 * k[11][c] + V9_a2512 * k[12][c] + V9_a2513 * k[13][c] + V9_a2514 * k[14][c] + V9_a2515 * k[15][c] + V9_a2517 * x[0][c] + V9_a2518 * x[1][c] + V9_a2519 * x[2][c] + V9_a2520 * x[3][c] + V9_a2521 * x[4][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c25, dt, tprev), x[8]);
    DEG_FOR(c) tmp[c] = y0[c] + dt * (V9_a2601 * k[1][c] + V9_a2608 * k[8][c] + V9_a2609 * k[9][c] + V9_a2610 * k[10][c] + V9_a2611 * k[11][c] + V9_a2612 * k[12][c] + V9_a2613 * k[13][c] + V9_a2614 * k[14][c] + V9_a2615 * k[15][c] + V9_a2617 * x[0][c] + V9_a2618 * x[1][c] + V9_a2619 * x[2][c] + V9_a2620 * x[3][c] + V9_a2621 * x[4][c]);
    deg_rhs(tmp, p, __builtin_fmaf(V9_c26, dt, tprev), x[9]);
    const float b1 = V9B1(th);
    const float b8 = V9B(08, th);
    const float b9 = V9B(09, th);
    const float b10 = V9B(10, th);
    const float b11 = V9B(11, th);
    const float b12 = V9B(12, th);
    const float b13 = V9B(13, th);
    const float b14 = V9B(14, th);
    const float b15 = V9B(15, th);
    const float b17 = V9B(17, th);
    const float b18 = V9B(18, th);
    const float b19 = V9B(19, th);
    const float b20 = V9B(20, th);
    const float b21 = V9B(21, th);
    const float b22 = V9B(22, th);
    const float b23 = V9B(23, th);
    const float b24 = V9B(24, th);
    const float b25 = V9B(25, th);
    const float b26 = V9B(26, th);
    DEG_FOR(c) out[c] = y0[c] + dt * (k[1][c] * b1 + k[8][c] * b8 + k[9][c] * b9 + k[10][c] * b10 + k[11][c] * b11 + k[12][c] * b12 + k[13][c] * b13 + k[14][c] * b14 + k[15][c] * b15 + x[0][c] * b17 + x[1][c] * b18 + x[2][c] * b19 + x[3][c] * b20 + x[4][c] * b21 + x[5][c] * b22 + x[6][c] * b23 + x[7][c] * b24 + x[8][c] * b25 + x[9][c] * b26);
}
#endif

// ------------------------------------------------------------------------------------
// Rosenbrock23 (gpu_rosenbrock23_perform_step.jl:33-60, :109-153) with an FP64 W = M - gamma J
// ------------------------------------------------------------------------------------
#ifndef DEG_MASS
#define DEG_MASS 0
#endif
static_assert(!(DEG_MASS && DEG_DUAL_P), "dual-number states take no mass matrix");
#if DEG_METHOD == 2 || DEG_METHOD == 4 || DEG_METHOD == 5 || DEG_METHOD == 6 || DEG_METHOD == 7
#if DEG_MASS
// The constant mass matrix M of the traced source (DEG_MASS_ENTRIES, row-major): out = M v, each
// row summed left to right (utils.jl:42-57 turns it into an SMatrix; the implicit methods use
// it in W and in their right-hand sides).
__device__ __forceinline__ float deg_mass_at(int r, int c) {
    constexpr float M[DEG_N * DEG_N] = {DEG_MASS_ENTRIES};
    return M[r * DEG_N + c];
}
__device__ __forceinline__ void deg_mass_mul(const float (&v)[DEG_N], float (&out)[DEG_N]) {
    DEG_FOR(r) {
        float acc = deg_mass_at(r, 0) * v[0];
        _Pragma("unroll") for (int c = 1; c < DEG_N; ++c) acc = acc + deg_mass_at(r, c) * v[c];
        out[r] = acc;
    }
}
#endif
// The factor of W (FP64), over a dual-number state of W's values (DEG_JC x DEG_JC) together with what
// the partials of W need: W = alpha M + beta J, so partial j of W is beta times partial j of J, the
// rows DEG_JC + r P + j of the dual Jacobian deg_jac writes. For DEG_JC <= 3 the factor is Cramer's
// rule, the reference's own solve for these sizes (linalg/linsolve.jl:8-57): W holds the cofactors
// (x_i = sum_j W[i][j] b_j / det, each row of cofactors in the reference's formula); above that, LU
// with partial pivoting.
#define DEG_CRAMER (DEG_JC <= 3)
struct deg_wf {
    double W[DEG_JC][DEG_JC];
    int piv[DEG_JC];
    double det;  // Cramer: the determinant (its reciprocal outside parity mode)
#if DEG_DUAL_P
    double beta;
    float jp[DEG_N - DEG_JC][DEG_JC];
#endif
};
// Record the dual part of a factor from the dual Jacobian J (DEG_N rows of DEG_JC columns)
__device__ __forceinline__ void w_dual(deg_wf& F, double beta, const float (&J)[DEG_N * DEG_JC]) {
#if DEG_DUAL_P
    F.beta = beta;
    _Pragma("unroll") for (int r = DEG_JC; r < DEG_N; ++r)
        _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) F.jp[r - DEG_JC][c] = J[r * DEG_JC + c];
#else
    (void)F;
    (void)beta;
    (void)J;
#endif
}
// Factor W in place. Cramer (DEG_JC <= 3): the cofactor rows and det as StaticArrays forms them
// (det of a 3 x 3 expanded along the first row). LU: row swaps are applied with compile-time indices
// so W stays in VGPRs. With DEG_PARITY every division is an IEEE double division, as the oracle does;
// otherwise the factor keeps reciprocals (1 / det, or 1 / U_kk on the diagonal), so a solve divides
// nowhere.
// 1 / d in FP64 for the contraction-allowed mode: the hardware reciprocal estimate refined by two
// Newton steps (r += r (1 - d r), each doubling the correct bits), in place of the IEEE division
// sequence; within a few units in the last place of the division (parity mode divides).
__device__ __forceinline__ double deg_recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    return __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
}
__device__ __forceinline__ void w_lu(deg_wf& F) {
    auto& W = F.W;
#if DEG_CRAMER
#if DEG_JC == 1
    F.det = W[0][0];
    W[0][0] = 1.0;
#elif DEG_JC == 2
    const double a00 = W[0][0], a01 = W[0][1], a10 = W[1][0], a11 = W[1][1];
    F.det = a00 * a11 - a01 * a10;
    W[0][0] = a11;
    W[0][1] = -a01;
    W[1][0] = -a10;
    W[1][1] = a00;
#else
    const double a00 = W[0][0], a01 = W[0][1], a02 = W[0][2];
    const double a10 = W[1][0], a11 = W[1][1], a12 = W[1][2];
    const double a20 = W[2][0], a21 = W[2][1], a22 = W[2][2];
    F.det = (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20)) + a02 * (a10 * a21 - a11 * a20);
    W[0][0] = a11 * a22 - a12 * a21;
    W[0][1] = a02 * a21 - a01 * a22;
    W[0][2] = a01 * a12 - a02 * a11;
    W[1][0] = a12 * a20 - a10 * a22;
    W[1][1] = a00 * a22 - a02 * a20;
    W[1][2] = a02 * a10 - a00 * a12;
    W[2][0] = a10 * a21 - a11 * a20;
    W[2][1] = a01 * a20 - a00 * a21;
    W[2][2] = a00 * a11 - a01 * a10;
#endif
#if !DEG_PARITY
    F.det = deg_recip(F.det);
#endif
#else
    _Pragma("unroll") for (int kk = 0; kk < DEG_JC; ++kk) {
        int pk = kk;
        double mx = __builtin_fabs(W[kk][kk]);
        _Pragma("unroll") for (int i = kk + 1; i < DEG_JC; ++i) {
            double v = __builtin_fabs(W[i][kk]);
            if (v > mx) {
                mx = v;
                pk = i;
            }
        }
        F.piv[kk] = pk;
        _Pragma("unroll") for (int i = kk + 1; i < DEG_JC; ++i) {
            if (pk == i) {
                _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) {
                    double tv = W[kk][c];
                    W[kk][c] = W[i][c];
                    W[i][c] = tv;
                }
            }
        }
#if !DEG_PARITY
        const double rp = deg_recip(W[kk][kk]);
#endif
        _Pragma("unroll") for (int i = kk + 1; i < DEG_JC; ++i) {
#if DEG_PARITY
            double l = W[i][kk] / W[kk][kk];
#else
            double l = W[i][kk] * rp;
#endif
            W[i][kk] = l;
            _Pragma("unroll") for (int c = kk + 1; c < DEG_JC; ++c) W[i][c] = W[i][c] - l * W[kk][c];
        }
#if !DEG_PARITY
        W[kk][kk] = rp;
#endif
    }
#endif
}
// y = W^-1 y in FP64
__device__ __forceinline__ void lu_solve_d(const deg_wf& F, double (&y)[DEG_JC]) {
    const auto& W = F.W;
#if DEG_CRAMER
    double x[DEG_JC];
    _Pragma("unroll") for (int i = 0; i < DEG_JC; ++i) {
        double v = W[i][0] * y[0];
        _Pragma("unroll") for (int j = 1; j < DEG_JC; ++j) v = v + W[i][j] * y[j];
#if DEG_PARITY
        x[i] = v / F.det;
#else
        x[i] = v * F.det;
#endif
    }
    _Pragma("unroll") for (int i = 0; i < DEG_JC; ++i) y[i] = x[i];
#else
    _Pragma("unroll") for (int kk = 0; kk < DEG_JC; ++kk) {
        _Pragma("unroll") for (int i = kk + 1; i < DEG_JC; ++i) {
            if (F.piv[kk] == i) {
                double tv = y[kk];
                y[kk] = y[i];
                y[i] = tv;
            }
        }
    }
    _Pragma("unroll") for (int i = 0; i < DEG_JC; ++i) {
        double v = y[i];
        _Pragma("unroll") for (int j = 0; j < i; ++j) v = v - W[i][j] * y[j];
        y[i] = v;
    }
    _Pragma("unroll") for (int i = DEG_JC - 1; i >= 0; --i) {
        double v = y[i];
        _Pragma("unroll") for (int j = i + 1; j < DEG_JC; ++j) v = v - W[i][j] * y[j];
#if DEG_PARITY
        y[i] = v / W[i][i];
#else
        y[i] = v * W[i][i];  // the factor keeps 1 / U_ii
#endif
    }
#endif
}
// x = W \ b. Over a dual-number state the solve is the dual one, the value LU plus one solve per
// partial: x_v = W_v \ b_v, x_j = W_v \ (b_j - W_j x_v) with W_j = beta J_j and x_v kept in FP64
// for the product (the derivative of W^-1 b is W^-1 (b' - W' W^-1 b)).
__device__ __forceinline__ void w_solve(const deg_wf& F, const float (&b)[DEG_N], float (&x)[DEG_N]) {
    double y[DEG_JC];
    _Pragma("unroll") for (int i = 0; i < DEG_JC; ++i) y[i] = (double)b[i];
    lu_solve_d(F, y);
    _Pragma("unroll") for (int i = 0; i < DEG_JC; ++i) x[i] = (float)y[i];
#if DEG_DUAL_P
    _Pragma("unroll") for (int j = 0; j < DEG_DUAL_P; ++j) {
        double z[DEG_JC];
        _Pragma("unroll") for (int r = 0; r < DEG_JC; ++r) {
            const int q = DEG_JC + r * DEG_DUAL_P + j;
            double v = (double)b[q];
            _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) v = v - (F.beta * (double)F.jp[q - DEG_JC][c]) * y[c];
            z[r] = v;
        }
        lu_solve_d(F, z);
        _Pragma("unroll") for (int r = 0; r < DEG_JC; ++r) x[DEG_JC + r * DEG_DUAL_P + j] = (float)z[r];
    }
#endif
}
#endif

#if DEG_METHOD == 1 || DEG_METHOD == 3
// The Vern7 / Vern9 dense output is out of line (it recomputes 6 / 10 extra stages) and takes the
// stages by reference. Handing it a copy made here keeps the step's own stage array out of
// memory: the copy is written to scratch only on the (rare) path that interpolates, instead of
// every attempt storing its stages for a call that may never come.
__device__ __forceinline__ void deg_interp(float th, float dt, float tprev, const float* p, const float (&y0)[DEG_N],
                                           const float (&y1)[DEG_N], const float (&k)[DEG_NK + 1][DEG_N],
                                           float (&out)[DEG_N]) {
    float y0c[DEG_N], y1c[DEG_N], kc[DEG_NK + 1][DEG_N];
    DEG_FOR(c) {
        y0c[c] = y0[c];
        y1c[c] = y1[c];
    }
    _Pragma("unroll") for (int i = 0; i <= DEG_NK; ++i) DEG_FOR(c) kc[i][c] = k[i][c];
    deg_interp_ool(th, dt, tprev, p, y0c, y1c, kc, out);
}
#endif

#if DEG_METHOD == 2
#define DEG_NK 2
// W = M - gamma J (M = I without a mass matrix)
__device__ __forceinline__ void w_factor(deg_wf& F, float gam, const float (&J)[DEG_N * DEG_JC]) {
    const double g = (double)gam;
    _Pragma("unroll") for (int r = 0; r < DEG_JC; ++r) {
#if DEG_MASS
        _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) F.W[r][c] = (double)deg_mass_at(r, c) - g * (double)J[r * DEG_JC + c];
#else
        _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) F.W[r][c] = (r == c ? 1.0 : 0.0) - g * (double)J[r * DEG_JC + c];
#endif
    }
    w_lu(F);
    w_dual(F, -g, J);
}
#define RB23_D (1.0f / (2.0f + __builtin_sqrtf(2.0f)))
#define RB23_E32 (6.0f + __builtin_sqrtf(2.0f))
// One Rosenbrock23 step from (uprev, t) with step dt; k[1], k[2] receive the slopes used by
// the dense output. Returns EEst when want_error is set.
template <bool want_error>
__device__ __forceinline__ float deg_rb23(const float (&uprev)[DEG_N], const float* p, float t, float dt,
                                          float (&k)[DEG_NK + 1][DEG_N], float (&unew)[DEG_N], float abstol,
                                          float reltol) {
    const float d = RB23_D;
    const float gam = dt * d;
    const float dto2 = dt / 2.0f;
    float Jm[DEG_N * DEG_JC], dT[DEG_N], F0[DEG_N], F1[DEG_N], rhs[DEG_N], tmp[DEG_N], x[DEG_N];
    deg_jac(uprev, p, t, Jm);
    deg_tgrad(uprev, p, t, dT);
    deg_wf W;
    w_factor(W, gam, Jm);
    deg_rhs(uprev, p, t, F0);
    DEG_FOR(c) rhs[c] = F0[c] + gam * dT[c];
    w_solve(W, rhs, k[1]);
    DEG_FOR(c) tmp[c] = uprev[c] + dto2 * k[1][c];
    deg_rhs(tmp, p, t + dto2, F1);
#if DEG_MASS
    deg_mass_mul(k[1], x);  // F1 - M k1 (gpu_rosenbrock23_perform_step.jl:54-58)
    DEG_FOR(c) rhs[c] = F1[c] - x[c];
#else
    DEG_FOR(c) rhs[c] = F1[c] - k[1][c];
#endif
    w_solve(W, rhs, x);
    DEG_FOR(c) k[2][c] = x[c] + k[1][c];
    DEG_FOR(c) unew[c] = uprev[c] + dt * k[2][c];
    if (!want_error) return 0.0f;
    const float dto6 = dt / 6.0f;
    const float e32 = RB23_E32;
    float F2[DEG_N], k3[DEG_N], e[DEG_N];
    deg_rhs(unew, p, t + dt, F2);
#if DEG_MASS
    {
        // F2 - M (e32 k2 + 2 k1) + e32 F1 + 2 F0 + dt dT (:144-148)
        float v[DEG_N], mv[DEG_N];
        DEG_FOR(c) v[c] = e32 * k[2][c] + 2.0f * k[1][c];
        deg_mass_mul(v, mv);
        DEG_FOR(c) rhs[c] = (((F2[c] - mv[c]) + e32 * F1[c]) + 2.0f * F0[c]) + dt * dT[c];
    }
#else
    DEG_FOR(c) rhs[c] = ((F2[c] - e32 * (k[2][c] - F1[c])) - 2.0f * (k[1][c] - F0[c])) + dt * dT[c];
#endif
    w_solve(W, rhs, k3);
    DEG_FOR(c) e[c] = dto6 * ((k[1][c] - 2.0f * k[2][c]) + k3[c]);
#if DEG_FASTPOW
    return scaled_error2_fast(e, uprev, unew, abstol, reltol);
#else
    return scaled_error(e, uprev, unew, abstol, reltol);
#endif
}
// y0 + dt * (c1 k1 + c2 k2), c1 = th(1-th)/(1-2d), c2 = th(th-2d)/(1-2d)
__device__ __forceinline__ void deg_interp(float th, float dt, float tprev, const float* p, const float (&y0)[DEG_N],
                                           const float (&y1)[DEG_N],
                                           const float (&k)[DEG_NK + 1][DEG_N], float (&out)[DEG_N]) {
    (void)tprev;
    (void)p;
    (void)y1;
    const float d = RB23_D;
    const float den = __builtin_fmaf(-2.0f, d, 1.0f);
#if DEG_PARITY
    const float c1 = (th * (1.0f - th)) / den;
    const float c2 = (th * __builtin_fmaf(-2.0f, d, th)) / den;
#else
    const float rden = 1.0f / den;  // a compile-time constant
    const float c1 = (th * (1.0f - th)) * rden;
    const float c2 = (th * __builtin_fmaf(-2.0f, d, th)) * rden;
#endif
    DEG_FOR(c) out[c] = y0[c] + dt * (c1 * k[1][c] + c2 * k[2][c]);
}
#endif

// ------------------------------------------------------------------------------------
// Rodas4 / Rodas5P (gpu_rodas4_perform_step.jl:1-277, gpu_rodas5P_perform_step.jl:1-351) with
// an FP64 W = J - I / (dt gamma); dense output y0 (1-th) + th (y1 + (1-th)(k1 + th k2 [+ th k3]))
// from the step's end state y1 (stiff/interpolants.jl:1-23)
// ------------------------------------------------------------------------------------
#if DEG_METHOD == 4 || DEG_METHOD == 5
// lt = -(A + C), with C the stage's combination of earlier stages (times M with a mass matrix:
// linsolve_tmp = du + dtd dT + M (dtC k), gpu_rodas4_perform_step.jl:77-101)
#if DEG_MASS
#define DEG_RODAS_LT(A, COMBO)                  \
    {                                           \
        float cm_[DEG_N], mc_[DEG_N];           \
        DEG_FOR(c) cm_[c] = (COMBO);            \
        deg_mass_mul(cm_, mc_);                 \
        DEG_FOR(c) lt[c] = -((A) + mc_[c]);     \
    }
#else
#define DEG_RODAS_LT(A, COMBO) DEG_FOR(c) lt[c] = -((A) + (COMBO));
#endif
#if DEG_METHOD == 4
#define DEG_NK 2
#define RO(x) RO4_##x
#else
#define DEG_NK 3
#define RO(x) RO5_##x
#endif
// One Rodas step from (uprev, t) with step dt. k[1..DEG_NK] receive the dense-output
// combinations (h rows); returns EEst from the last stage when want_error is set.
template <bool want_error>
__device__ __forceinline__ float deg_rodas(const float (&uprev)[DEG_N], const float* p, float t, float dt,
                                           float (&k)[DEG_NK + 1][DEG_N], float (&unew)[DEG_N], float abstol,
                                           float reltol) {
    float Jm[DEG_N * DEG_JC], dT[DEG_N], du[DEG_N], lt[DEG_N], u[DEG_N];
    deg_jac(uprev, p, t, Jm);
    deg_tgrad(uprev, p, t, dT);
    const float dtC21 = RO(C21) / dt, dtC31 = RO(C31) / dt, dtC32 = RO(C32) / dt;
    const float dtC41 = RO(C41) / dt, dtC42 = RO(C42) / dt, dtC43 = RO(C43) / dt;
    const float dtC51 = RO(C51) / dt, dtC52 = RO(C52) / dt, dtC53 = RO(C53) / dt, dtC54 = RO(C54) / dt;
    const float dtC61 = RO(C61) / dt, dtC62 = RO(C62) / dt, dtC63 = RO(C63) / dt, dtC64 = RO(C64) / dt;
    const float dtC65 = RO(C65) / dt;
    const float dtd1 = dt * RO(d1), dtd2 = dt * RO(d2), dtd3 = dt * RO(d3), dtd4 = dt * RO(d4);
    const float dtgamma = dt * RO(gamma);
    deg_wf W;
    {
        // W = J - M inv(dt gamma) (gpu_rodas4_perform_step.jl:66-67), the product with M in Float32
        const float igf = 1.0f / dtgamma;
        const double ig = (double)igf;
        _Pragma("unroll") for (int r = 0; r < DEG_JC; ++r) {
#if DEG_MASS
            _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) W.W[r][c] = (double)Jm[r * DEG_JC + c] - (double)(deg_mass_at(r, c) * igf);
#else
            _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) W.W[r][c] = (double)Jm[r * DEG_JC + c] - (r == c ? ig : 0.0);
#endif
        }
        w_lu(W);
        w_dual(W, 1.0, Jm);
    }
    float K1[DEG_N], K2[DEG_N], K3[DEG_N], K4[DEG_N], K5[DEG_N], K6[DEG_N];
    deg_rhs(uprev, p, t, du);
    // stage 1
    DEG_FOR(c) lt[c] = -(du[c] + dtd1 * dT[c]);
    w_solve(W, lt, K1);
    DEG_FOR(c) u[c] = uprev[c] + RO(a21) * K1[c];
    deg_rhs(u, p, t + RO(c2) * dt, du);
    // stage 2
    DEG_RODAS_LT((du[c] + dtd2 * dT[c]), dtC21 * K1[c])
    w_solve(W, lt, K2);
    DEG_FOR(c) u[c] = (uprev[c] + RO(a31) * K1[c]) + RO(a32) * K2[c];
    deg_rhs(u, p, t + RO(c3) * dt, du);
    // stage 3
    DEG_RODAS_LT((du[c] + dtd3 * dT[c]), (dtC31 * K1[c] + dtC32 * K2[c]))
    w_solve(W, lt, K3);
    DEG_FOR(c) u[c] = ((uprev[c] + RO(a41) * K1[c]) + RO(a42) * K2[c]) + RO(a43) * K3[c];
    deg_rhs(u, p, t + RO(c4) * dt, du);
    // stage 4
    DEG_RODAS_LT((du[c] + dtd4 * dT[c]), ((dtC41 * K1[c] + dtC42 * K2[c]) + dtC43 * K3[c]))
    w_solve(W, lt, K4);
    DEG_FOR(c) u[c] = (((uprev[c] + RO(a51) * K1[c]) + RO(a52) * K2[c]) + RO(a53) * K3[c]) + RO(a54) * K4[c];
#if DEG_METHOD == 4
    deg_rhs(u, p, t + dt, du);
    // stage 5 (its C terms in the reference's order 2, 4, 1, 3)
    DEG_RODAS_LT(du[c], (((dtC52 * K2[c] + dtC54 * K4[c]) + dtC51 * K1[c]) + dtC53 * K3[c]))
    w_solve(W, lt, K5);
    DEG_FOR(c) u[c] = u[c] + K5[c];
    deg_rhs(u, p, t + dt, du);
    // stage 6 (order 1, 2, 5, 4, 3)
    DEG_RODAS_LT(du[c], ((((dtC61 * K1[c] + dtC62 * K2[c]) + dtC65 * K5[c]) + dtC64 * K4[c]) + dtC63 * K3[c]))
    w_solve(W, lt, K6);
    DEG_FOR(c) unew[c] = u[c] + K6[c];
    DEG_FOR(c) {
        k[1][c] = (((RO(h21) * K1[c] + RO(h22) * K2[c]) + RO(h23) * K3[c]) + RO(h24) * K4[c]) + RO(h25) * K5[c];
        k[2][c] = (((RO(h31) * K1[c] + RO(h32) * K2[c]) + RO(h33) * K3[c]) + RO(h34) * K4[c]) + RO(h35) * K5[c];
    }
    float* const Kerr = K6;
#else
    const float dtd5 = dt * RO(d5);
    const float dtC71 = RO(C71) / dt, dtC72 = RO(C72) / dt, dtC73 = RO(C73) / dt, dtC74 = RO(C74) / dt;
    const float dtC75 = RO(C75) / dt, dtC76 = RO(C76) / dt;
    const float dtC81 = RO(C81) / dt, dtC82 = RO(C82) / dt, dtC83 = RO(C83) / dt, dtC84 = RO(C84) / dt;
    const float dtC85 = RO(C85) / dt, dtC86 = RO(C86) / dt, dtC87 = RO(C87) / dt;
    float K7[DEG_N], K8[DEG_N];
    deg_rhs(u, p, t + RO(c5) * dt, du);
    // stage 5 (its C terms in the reference's order 2, 4, 1, 3)
    DEG_RODAS_LT((du[c] + dtd5 * dT[c]), (((dtC52 * K2[c] + dtC54 * K4[c]) + dtC51 * K1[c]) + dtC53 * K3[c]))
    w_solve(W, lt, K5);
    DEG_FOR(c) u[c] = ((((uprev[c] + RO(a61) * K1[c]) + RO(a62) * K2[c]) + RO(a63) * K3[c]) + RO(a64) * K4[c]) + RO(a65) * K5[c];
    deg_rhs(u, p, t + dt, du);
    // stage 6
    DEG_RODAS_LT(du[c], ((((dtC61 * K1[c] + dtC62 * K2[c]) + dtC63 * K3[c]) + dtC64 * K4[c]) + dtC65 * K5[c]))
    w_solve(W, lt, K6);
    DEG_FOR(c) u[c] = u[c] + K6[c];
    deg_rhs(u, p, t + dt, du);
    // stage 7
    DEG_RODAS_LT(du[c], (((((dtC71 * K1[c] + dtC72 * K2[c]) + dtC73 * K3[c]) + dtC74 * K4[c]) + dtC75 * K5[c]) + dtC76 * K6[c]))
    w_solve(W, lt, K7);
    DEG_FOR(c) u[c] = u[c] + K7[c];
    deg_rhs(u, p, t + dt, du);
    // stage 8
    DEG_RODAS_LT(du[c], ((((((dtC81 * K1[c] + dtC82 * K2[c]) + dtC83 * K3[c]) + dtC84 * K4[c]) + dtC85 * K5[c]) + dtC86 * K6[c]) + dtC87 * K7[c]))
    w_solve(W, lt, K8);
    DEG_FOR(c) unew[c] = u[c] + K8[c];
    DEG_FOR(c) {
        k[1][c] = ((((((RO(h21) * K1[c] + RO(h22) * K2[c]) + RO(h23) * K3[c]) + RO(h24) * K4[c]) + RO(h25) * K5[c]) + RO(h26) * K6[c]) + RO(h27) * K7[c]) + RO(h28) * K8[c];
        k[2][c] = ((((((RO(h31) * K1[c] + RO(h32) * K2[c]) + RO(h33) * K3[c]) + RO(h34) * K4[c]) + RO(h35) * K5[c]) + RO(h36) * K6[c]) + RO(h37) * K7[c]) + RO(h38) * K8[c];
        k[3][c] = ((((((RO(h41) * K1[c] + RO(h42) * K2[c]) + RO(h43) * K3[c]) + RO(h44) * K4[c]) + RO(h45) * K5[c]) + RO(h46) * K6[c]) + RO(h47) * K7[c]) + RO(h48) * K8[c];
    }
    float* const Kerr = K8;
#endif
    if (!want_error) return 0.0f;
    float e[DEG_N];
    DEG_FOR(c) e[c] = Kerr[c];
#if DEG_FASTPOW
    return scaled_error2_fast(e, uprev, unew, abstol, reltol);
#else
    return scaled_error(e, uprev, unew, abstol, reltol);
#endif
}
__device__ __forceinline__ void deg_interp(float th, float dt, float tprev, const float* p, const float (&y0)[DEG_N],
                                           const float (&y1)[DEG_N],
                                           const float (&k)[DEG_NK + 1][DEG_N], float (&out)[DEG_N]) {
    (void)dt;
    (void)tprev;
    (void)p;
    const float th1 = 1.0f - th;
#if DEG_METHOD == 4
    DEG_FOR(c) out[c] = th1 * y0[c] + th * (y1[c] + th1 * (k[1][c] + th * k[2][c]));
#else
    DEG_FOR(c) out[c] = th1 * y0[c] + th * (y1[c] + th1 * (k[1][c] + th * (k[2][c] + th * k[3][c])));
#endif
}
#endif

// ------------------------------------------------------------------------------------
// Kvaerno3 / Kvaerno5 (gpu_kvaerno3_perform_step.jl:1-232, gpu_kvaerno5_perform_step.jl:1-300):
// ESDIRK stages solved by Newton's method (nlsolve/utils.jl:1-27, nlsolve/type.jl:129-185) with
// W = -I + gamma dt J re-evaluated each iteration and factored in FP64; Hermite dense output
// (nonstiff/interpolants.jl:1-20) from uprev, u, k1 and k2 = z_last / dt
// ------------------------------------------------------------------------------------
#if DEG_METHOD == 6 || DEG_METHOD == 7
#define DEG_NK 2
#if DEG_METHOD == 6
#define KV(x) KV3_##x
#else
#define KV(x) KV5_##x
#endif
#define DEG_KV_NEWTON_MAX 30
#define DEG_KV_NEWTON_TOL (100.0f * 1.1920928955078125e-7f)  // 100 eps(Float32)
// W = -M + (gamma dt) J(x, tc) (M = I without a mass matrix), LU-factored in FP64
__device__ __forceinline__ void kv_w(deg_wf& F, float gdt, const float (&x)[DEG_N], const float* p, float tc) {
    float Jm[DEG_N * DEG_JC];
    deg_jac(x, p, tc, Jm);
    const double g = (double)gdt;
    _Pragma("unroll") for (int r = 0; r < DEG_JC; ++r) {
#if DEG_MASS
        _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) F.W[r][c] = -(double)deg_mass_at(r, c) + g * (double)Jm[r * DEG_JC + c];
#else
        _Pragma("unroll") for (int c = 0; c < DEG_JC; ++c) F.W[r][c] = (r == c ? -1.0 : 0.0) + g * (double)Jm[r * DEG_JC + c];
#endif
    }
    w_lu(F);
    w_dual(F, g, Jm);
}
// Newton on z for the stage x = tmp + gamma z at tc = tnl + c dtnl: z -= W \ (dtnl f(x) - z), at
// most 30 iterations, until ||dtnl f(tmp + gamma z) - z|| < 100 eps(Float32)
__device__ __forceinline__ void kv_newton(float (&z)[DEG_N], const float (&tmp)[DEG_N], const float* p, float tnl,
                                          float dtnl, float c) {
    const float tc = tnl + c * dtnl;
    const float gdt = KV(gamma) * dtnl;
    for (int it = 0; it < DEG_KV_NEWTON_MAX; ++it) {
        float x[DEG_N], fe[DEG_N], r[DEG_N], dz[DEG_N];
        DEG_FOR(q) x[q] = tmp[q] + KV(gamma) * z[q];
        deg_wf W;
        kv_w(W, gdt, x, p, tc);
        deg_rhs(x, p, tc, fe);
#if DEG_MASS
        float mz[DEG_N];
        deg_mass_mul(z, mz);  // dt f - M z (nlsolve/utils.jl:17, :21)
        DEG_FOR(q) r[q] = dtnl * fe[q] - mz[q];
#else
        DEG_FOR(q) r[q] = dtnl * fe[q] - z[q];
#endif
        w_solve(W, r, dz);
        DEG_FOR(q) z[q] = z[q] - dz[q];
        DEG_FOR(q) x[q] = tmp[q] + KV(gamma) * z[q];
        deg_rhs(x, p, tc, fe);
#if DEG_MASS
        deg_mass_mul(z, mz);
        DEG_FOR(q) r[q] = dtnl * fe[q] - mz[q];
#else
        DEG_FOR(q) r[q] = dtnl * fe[q] - z[q];
#endif
#if DEG_DUAL_P
        // diffeqgpunorm (utils.jl:1-4): over a dual array, the values alone
        float acc = r[0] * r[0];
        _Pragma("unroll") for (int q = 1; q < DEG_JC; ++q) acc = acc + r[q] * r[q];
        if (__builtin_sqrtf(acc / (float)DEG_JC) < DEG_KV_NEWTON_TOL) break;
#else
        if (rms_norm(r) < DEG_KV_NEWTON_TOL) break;
#endif
    }
}
// One Kvaerno step from (uprev, t) with step dt and first stage slope k1 (z1 = dt k1); the
// implicit stages use the nonlinear solver's time tnl and step dtnl. zl receives the last stage.
// Returns EEst (W(u, t + dt) \ (btilde . z), scaled) when want_error is set.
template <bool want_error>
__device__ __forceinline__ float deg_kvaerno(const float (&uprev)[DEG_N], const float* p, float t, float dt, float tnl,
                                             float dtnl, const float (&k1)[DEG_N], float (&zl)[DEG_N],
                                             float (&unew)[DEG_N], float abstol, float reltol) {
    float z1[DEG_N], z2[DEG_N], z3[DEG_N], tmp[DEG_N], ev[DEG_N];
    DEG_FOR(q) z1[q] = dt * k1[q];
    DEG_FOR(q) {
        z2[q] = z1[q];
        tmp[q] = uprev[q] + KV(gamma) * z1[q];
    }
    kv_newton(z2, tmp, p, tnl, dtnl, KV(gamma));
    DEG_FOR(q) {
        z3[q] = KV(alpha31) * z1[q] + KV(alpha32) * z2[q];
        tmp[q] = (uprev[q] + KV(a31) * z1[q]) + KV(a32) * z2[q];
    }
    kv_newton(z3, tmp, p, tnl, dtnl, KV(c3));
#if DEG_METHOD == 6
    DEG_FOR(q) {
        zl[q] = (KV(a31) * z1[q] + KV(a32) * z2[q]) + KV(gamma) * z3[q];
        tmp[q] = ((uprev[q] + KV(a41) * z1[q]) + KV(a42) * z2[q]) + KV(a43) * z3[q];
    }
    kv_newton(zl, tmp, p, tnl, dtnl, 1.0f);
    DEG_FOR(q) unew[q] = tmp[q] + KV(gamma) * zl[q];
    if (!want_error) return 0.0f;
    DEG_FOR(q) ev[q] = ((KV(btilde1) * z1[q] + KV(btilde2) * z2[q]) + KV(btilde3) * z3[q]) + KV(btilde4) * zl[q];
#else
    float z4[DEG_N], z5[DEG_N], z6[DEG_N];
    DEG_FOR(q) {
        z4[q] = (KV(alpha41) * z1[q] + KV(alpha42) * z2[q]) + KV(alpha43) * z3[q];
        tmp[q] = ((uprev[q] + KV(a41) * z1[q]) + KV(a42) * z2[q]) + KV(a43) * z3[q];
    }
    kv_newton(z4, tmp, p, tnl, dtnl, KV(c4));
    DEG_FOR(q) {
        z5[q] = (KV(alpha51) * z1[q] + KV(alpha52) * z2[q]) + KV(alpha53) * z3[q];
        tmp[q] = (((uprev[q] + KV(a51) * z1[q]) + KV(a52) * z2[q]) + KV(a53) * z3[q]) + KV(a54) * z4[q];
    }
    kv_newton(z5, tmp, p, tnl, dtnl, KV(c5));
    DEG_FOR(q) {
        z6[q] = (KV(alpha61) * z1[q] + KV(alpha62) * z2[q]) + KV(alpha63) * z3[q];
        tmp[q] = (((uprev[q] + KV(a61) * z1[q]) + KV(a63) * z3[q]) + KV(a64) * z4[q]) + KV(a65) * z5[q];
    }
    kv_newton(z6, tmp, p, tnl, dtnl, KV(c6));
    DEG_FOR(q) {
        zl[q] = (((KV(a61) * z1[q] + KV(a63) * z3[q]) + KV(a64) * z4[q]) + KV(a65) * z5[q]) + KV(gamma) * z6[q];
        tmp[q] = ((((uprev[q] + KV(a71) * z1[q]) + KV(a73) * z3[q]) + KV(a74) * z4[q]) + KV(a75) * z5[q]) + KV(a76) * z6[q];
    }
    kv_newton(zl, tmp, p, tnl, dtnl, 1.0f);
    DEG_FOR(q) unew[q] = tmp[q] + KV(gamma) * zl[q];
    if (!want_error) return 0.0f;
    DEG_FOR(q) ev[q] = ((((KV(btilde1) * z1[q] + KV(btilde3) * z3[q]) + KV(btilde4) * z4[q]) + KV(btilde5) * z5[q]) +
                        KV(btilde6) * z6[q]) + KV(btilde7) * zl[q];
#endif
    deg_wf W;
    float e[DEG_N];
    kv_w(W, KV(gamma) * dt, unew, p, t + 1.0f * dt);
    w_solve(W, ev, e);
#if DEG_FASTPOW
    return scaled_error2_fast(e, uprev, unew, abstol, reltol);
#else
    return scaled_error(e, uprev, unew, abstol, reltol);
#endif
}
// Hermite: (1-th) y0 + th y1 + th (th-1) ((1-2th)(y1-y0) + (th-1) dt k1 + th dt k2)
__device__ __forceinline__ void deg_interp(float th, float dt, float tprev, const float* p, const float (&y0)[DEG_N],
                                           const float (&y1)[DEG_N],
                                           const float (&k)[DEG_NK + 1][DEG_N], float (&out)[DEG_N]) {
    (void)tprev;
    (void)p;
    const float a0 = 1.0f - th;
    const float s = th * (th - 1.0f);
    const float w = 1.0f - 2.0f * th;
    const float d1 = (th - 1.0f) * dt;
    const float d2 = th * dt;
    DEG_FOR(c) out[c] = (a0 * y0[c] + th * y1[c]) + s * ((w * (y1[c] - y0[c]) + d1 * k[1][c]) + d2 * k[2][c]);
}
#endif

#define DEG_HAS_INTERP 1

// ------------------------------------------------------------------------------------
// Persistent work queue: lanes that need a trajectory take ids from one device counter,
// one atomic per wave (ballot + popcount + prefix rank).
// ------------------------------------------------------------------------------------
// Persistent scheduling: DEG_NQUEUE counters, each owning a slab of trajectory ids, spaced
// DEG_QSTRIDE words apart (separate memory channels). A wave claims chunks of ids from
// its counter with one atomic (issued by one lane) and hands them to its idle lanes by
// ballot rank; when its slab is exhausted it moves on to the next counter.
#define DEG_NQUEUE 8
#define DEG_QSTRIDE 64
__device__ __forceinline__ unsigned int deg_claim(unsigned int* q, int leader, int lane, unsigned int amount) {
    unsigned int r = 0;
    if (lane == leader) r = atomicAdd(q, amount);
    return __shfl(r, leader);
}

// the next float above x (x finite)
__device__ __forceinline__ float deg_next_up(float x) {
    if (x == 0.0f) return 1.4e-45f;
    const unsigned int b = __builtin_bit_cast(unsigned int, x);
    return __builtin_bit_cast(float, (x > 0.0f) ? b + 1u : b - 1u);
}

// PI controller on the squared error norm in log2 space (DEG_FASTPOW): the reference's
// q11 = EEst^beta1, q = q11 / qold^beta2, reject dt /= min(1/qmin, q11/gamma), accept
// dt /= clamp(q/gamma, 1/qmax, 1/qmin) become one log2 and one exp2:
//   reject: dt * max(qmin, gamma * 2^(-beta1/2 * L))
//   accept: dt * clamp(gamma * 2^(beta2 * Lq - beta1/2 * L), qmin, qmax), Lq = log2 qold
// with L = log2(EEst^2); EEst == 0 gives the reference's factor 1/((1/qmax)/gamma).
// Returns the step factor; lq carries log2(qold) and is updated on acceptance.
__device__ __forceinline__ float deg_fast_control(float E2, bool accept, float& lq) {
    const float L = __builtin_amdgcn_logf(E2);
    const float x = (accept ? DEG_BETA2 * lq : 0.0f) - (0.5f * DEG_BETA1) * L;
    const float f = DEG_GAMMA * __builtin_amdgcn_exp2f(x);
    float factor;
    if (accept) {
        factor = __builtin_fminf(__builtin_fmaxf(f, DEG_QMIN), DEG_QMAX);
        if (E2 == 0.0f) factor = 1.0f / __builtin_fmaxf(1.0f / DEG_QMAX, __builtin_fminf(1.0f / DEG_QMIN, (1.0f / DEG_QMAX) / DEG_GAMMA));
        lq = __builtin_fmaxf(0.5f * L, -13.2877123795f);  // log2(max(EEst, 1e-4))
    } else {
        factor = __builtin_fmaxf(f, DEG_QMIN);
    }
    return factor;
}

// Julia's sign(x) for Float32: -1, +1, or x itself for zeros and NaN
__device__ __forceinline__ float jl_sign(float x) { return (x < 0.0f) ? -1.0f : ((x > 0.0f) ? 1.0f : x); }

// Julia's eps(x) for Float32: the spacing 2^(exponent(x) - 23) for normal x, the smallest
// subnormal below floatmin, NaN for Inf and NaN
__device__ __forceinline__ float jl_eps(float x) {
    const float ax = __builtin_fabsf(x);
    if (!(ax <= 3.40282347e38f)) return __builtin_nanf("");
    if (ax < 1.17549435e-38f) return 1.4e-45f;
    const int e = (int)((__builtin_bit_cast(unsigned int, ax) >> 23) & 0xffu) - 127 - 23;
    if (e >= -126) return __builtin_bit_cast(float, (unsigned int)(e + 127) << 23);
    return __builtin_bit_cast(float, 1u << (e + 149));
}

// Julia's nextfloat(x) for Float32 (NaN stays NaN, Inf stays Inf)
__device__ __forceinline__ float jl_nextfloat(float x) {
    if (x != x || x == __builtin_inff()) return x;
    return deg_next_up(x);
}

// t <- t + dt, optionally compensated
__device__ __forceinline__ float deg_advance(float t, float dt, float& comp) {
#if DEG_KAHAN
    float y = dt - comp;
    float tn = t + y;
    comp = (tn - t) - y;
    return tn;
#else
    (void)comp;
    return t + dt;
#endif
}

#ifndef DEG_BLOCK_MAX
#define DEG_BLOCK_MAX 256
#endif

// ------------------------------------------------------------------------------------
// Tail compaction inside a workgroup (DEG_COMPACT). Once a wave can claim no more ids,
// a wave with few trajectories left hands their complete state to the workgroup's LDS
// mailbox and leaves; waves with idle lanes adopt mailbox states. One 64-bit LDS word
// tracks the mailbox: live waves (bits 0-7), donated states (8-15), states written (16-23),
// states claimed (24-31), states read (32-39). A donation needs another live wave that
// has not left; a wave leaves only while the mailbox is empty, so no state is stranded.
// No wave waits on another: every step is one bounded CAS loop by the wave's leader lane.
// ------------------------------------------------------------------------------------
#ifndef DEG_COMPACT
#define DEG_COMPACT 0
#endif
#ifndef DEG_COMPACT_LANES
#define DEG_COMPACT_LANES 24
#endif
// the start order's inputs laid out in claim order (deg_solve_args.inputs_ordered)
#ifndef DEG_ORDERED_INPUTS
#define DEG_ORDERED_INPUTS 0
#endif
// Batched dense output (DEG_BATCH_SAVES, adaptive Tsit5): a lane whose accepted step crosses
// save points copies the step's dense-output data to its LDS stash; the wave interpolates
// for all stashed lanes together once DEG_SAVE_BATCH lanes wait (or a lane needs its stash
// again or ends), instead of running the interpolation branch in nearly every iteration.
#ifndef DEG_BATCH_SAVES
#define DEG_BATCH_SAVES 0
#endif
#define DEG_STASH (((8 * DEG_N + 2) + 3) & ~3)
#ifndef DEG_SAVE_BATCH
#define DEG_SAVE_BATCH 16
#endif
#define DEG_MB_SLOTS 64
#define DEG_MB_WORDS (16 + 2 * DEG_N)

__device__ __forceinline__ unsigned long long deg_mb_read(unsigned long long* mb) {
    return *(volatile unsigned long long*)mb;
}

// the leader lane of a wave with k (1..64) trajectories left reserves the mailbox for them
__device__ __forceinline__ int deg_mb_donate(unsigned long long* mb, int k) {
    unsigned long long w = deg_mb_read(mb);
    for (int it = 0; it < 64; ++it) {
        const unsigned long long live = w & 0xffull, pend = (w >> 8) & 0xffull;
        if (pend != 0ull || live < 2ull) return 0;
        const unsigned long long nw = (live - 1ull) | ((unsigned long long)k << 8);  // the donor leaves
        const unsigned long long prev = atomicCAS(mb, w, nw);
        if (prev == w) return 1;
        w = prev;
    }
    return 0;
}

// the leader lane of a wave without trajectories leaves, unless donated states still wait
__device__ __forceinline__ int deg_mb_retire(unsigned long long* mb) {
    unsigned long long w = deg_mb_read(mb);
    for (int it = 0; it < 64; ++it) {
        if (((w >> 8) & 0xffull) != 0ull) return 0;
        const unsigned long long prev = atomicCAS(mb, w, w - 1ull);
        if (prev == w) return 1;
        w = prev;
    }
    return 0;
}

// the leader lane claims up to `room` written states; returns the count, first slot in *base
__device__ __forceinline__ int deg_mb_take(unsigned long long* mb, int room, int* base) {
    unsigned long long w = deg_mb_read(mb);
    for (int it = 0; it < 64; ++it) {
        const int ready = (int)((w >> 16) & 0xffull), taken = (int)((w >> 24) & 0xffull);
        if (ready <= taken) return 0;
        const int m = min(room, ready - taken);
        const unsigned long long prev = atomicCAS(mb, w, w + ((unsigned long long)m << 24));
        if (prev == w) {
            *base = taken;
            return m;
        }
        w = prev;
    }
    return 0;
}

// after reading m states: the last reader frees the mailbox (keeps the live count)
__device__ __forceinline__ void deg_mb_consumed(unsigned long long* mb, int m) {
    const unsigned long long old = atomicAdd(mb, (unsigned long long)m << 32);
    const unsigned long long pend = (old >> 8) & 0xffull, cons = ((old >> 32) & 0xffull) + (unsigned long long)m;
    if (cons == pend) atomicAnd(mb, 0xffull);
}

// ------------------------------------------------------------------------------------
// The ensemble kernel (kernels.jl:1-152), one trajectory per lane, one attempt per
// loop iteration so that lanes stay in step and a finished lane can be refilled.
// ------------------------------------------------------------------------------------
#ifndef DEG_WAVES_MIN
#define DEG_WAVES_MIN 0
#endif
#if DEG_WAVES_MIN > 0
// register cap: at least DEG_WAVES_MIN resident waves per SIMD
#define DEG_ATTR_WAVES __attribute__((amdgpu_waves_per_eu(DEG_WAVES_MIN)))
#else
#define DEG_ATTR_WAVES
#endif
#ifndef DEG_SAVEAT_GLOBAL
#define DEG_SAVEAT_GLOBAL 0
#endif

extern "C" __global__ void __launch_bounds__(DEG_BLOCK_MAX) DEG_ATTR_WAVES DEG_KERNEL(const deg_kargs a) {
    const long long S = a.stride;
#if DEG_COMPACT
    __shared__ unsigned long long mb_word;
    __shared__ float mb_state[DEG_MB_SLOTS][DEG_MB_WORDS];
    if (threadIdx.x == 0) mb_word = (unsigned long long)(blockDim.x / 64);  // every wave starts live
#if DEG_SAVEAT_GLOBAL
    __syncthreads();
#endif
#endif
#if DEG_LEARNED
    // the proposer's packed weights in LDS (read by every evaluation without a vector-memory wait)
    __shared__ deg_mlp_packed deg_mlpw;
    deg_mlp_stage(deg_mlpw, (const deg_mlp_packed*)a.mlp);
#if DEG_SAVEAT_GLOBAL
    __syncthreads();
#endif
#endif
    // saveat times staged in LDS: reading them then needs no vector-memory wait (which on
    // CDNA would also wait for every store still in flight)
    extern __shared__ float deg_dyn[];  // sized at launch: saveat times, then the save stash
    const float finf_lds = __builtin_inff();
#if DEG_SAVEAT_GLOBAL
    const float* __restrict__ svp = a.saveat;  // more save points than the LDS cache holds
#else
    float* svp = deg_dyn;  // nsaveat floats and a +inf sentinel
    if (a.saveat != nullptr) {
        for (int j = threadIdx.x; j <= a.nsaveat; j += blockDim.x) svp[j] = j < a.nsaveat ? a.saveat[j] : finf_lds;
    }
    __syncthreads();
#endif
#if DEG_BATCH_SAVES
    // this lane's stash of one accepted step's dense-output data (uprev, k1..k7, dt, tprev)
#if DEG_SAVEAT_GLOBAL
    float* const stash = deg_dyn + (long long)threadIdx.x * DEG_STASH;
#else
    float* const stash = deg_dyn + ((a.nsaveat + 4) & ~3) + threadIdx.x * DEG_STASH;
#endif
    bool stashed = false;  // the stash holds saves [plo, phi) not yet written
    int plo = 0, phi = 0;
#endif
#if DEG_PERSISTENT
    long long gid = 0;
#else
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= a.n_traj) return;
#endif
    const bool use_saveat = a.saveat != nullptr;
    const float finf = __builtin_inff();

    long long i = 0;
    float p[DEG_NP > 0 ? DEG_NP : 1];
    float u[DEG_N], uprev[DEG_N], unew[DEG_N];
    float k[DEG_NK + 1][DEG_N];
    float t = 0.0f, tf = 0.0f, tprev = 0.0f, dt = 0.0f, dtacc = 0.0f, tc = 0.0f, tnext = finf;
    float qold = DEG_QOLDINIT, qpow = 1.0f;
    int cur = 0, step_idx = 1, nacc = 0, nrej = 0, status = DEG_RET_SUCCESS;
    float* ts_w = nullptr;  // this lane's ts/us entries of save row `cur`
    float* us_w = nullptr;
    long long attempts = 0;
    bool have_k = false;
    bool fsal_pending = false;
#if DEG_TSTOPS
    int tsi = 0;     // tstops_idx (0-based)
#endif
#if DEG_NCCB > 0
    int evlast = 1;  // event_last_time: 1-based index of the last continuous event, 0 none
#endif
#if DEG_LEARNED
    float zml_keep = 0.0f;  // the proposer's last correction for this lane
    int mlp_tick = 0;       // wave-uniform iteration counter modulo mlp_every
    // the last attempt's features and the step its controller left pending for the proposer:
    // accepted?, log2 of the PI factor, the attempt's dt, and (tf - t) - dt before t advanced
    float lfeat[DEG_MLP_IN] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    bool lpend = false, lacc = false;
    float lzpi = 0.0f, ldt0 = 0.0f, lrem = 0.0f;
#endif
#if DEG_LEARNED || DEG_STEPLOG
    float lx_prev = -13.2877123795f;  // log2 EEst of the previous attempt (log2 qoldinit at first)
    float ldt_prev = 0.0f;            // log2 dt of the previous attempt
    bool first_attempt = true;
#endif
#if DEG_STEPLOG
    float fprev[DEG_MLP_IN] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float zpi_prev = 0.0f;
    bool acc_prev = false;
    bool have_prev = false;
#endif
#if DEG_PERSISTENT
    bool need_init = false;
    bool have = false;  // the lane holds a trajectory
    bool want = true;   // the lane waits for one
    // wave-uniform scheduler state
    const long long slab = (a.n_traj + DEG_NQUEUE - 1) / DEG_NQUEUE;
    int qi = blockIdx.x % DEG_NQUEUE;   // blocks are dispatched round-robin over the XCDs
    int qtries = 0;
    long long c_next = 0, c_end = 0;    // the wave's claimed chunk [c_next, c_end)
    const int lane = (int)__lane_id();
#else
    bool need_init = true;
#endif

    for (;;) {
#if DEG_PERSISTENT
        // Scheduling block, executed by every live lane of the wave. It is written without
        // continue/break inside so that the wave-uniform chunk state (c_next, c_end, qi,
        // qtries) is updated by all live lanes together.
        bool retire = false;
        {
            // Finished lanes wait until `refill` lanes of the wave are idle (or none is
            // running) and then take new trajectories together: one queue atomic and one
            // load latency per batch instead of per trajectory.
            const unsigned long long wmask = __ballot(want);
            const unsigned long long hmask = __ballot(have);
            if (wmask != 0ull && (hmask == 0ull || __popcll(wmask) >= (unsigned long long)a.refill)) {
                // hand ids of the wave's chunk to the waiting lanes in rank order, claiming
                // further chunks (and moving on to other counters) as needed
                const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
                unsigned long long pending = wmask;
                while (pending != 0ull && qtries < DEG_NQUEUE) {
                    if (c_next >= c_end) {
                        // with a start order, queue q serves its entries q, q + 8, q + 16, ... (see the
                        // lookup at init), so every XCD starts on its share of the first entries
                        const long long lim = a.order ? (long long)qi * slab + (a.n_traj - qi + DEG_NQUEUE - 1) / DEG_NQUEUE
                                                      : min((long long)(qi + 1) * slab, a.n_traj);
                        const long long base = (long long)qi * slab +
                            (long long)deg_claim(a.queue + qi * DEG_QSTRIDE, leader, lane, (unsigned int)a.chunk);
                        if (base >= lim) {
                            qi = (qi + 1) % DEG_NQUEUE;
                            ++qtries;
                        } else {
                            c_next = base;
                            c_end = min(base + (long long)a.chunk, lim);
                        }
                    } else {
                        const long long avail = c_end - c_next;
                        const bool mine = (pending >> lane) & 1ull;
                        const long long rank = (long long)__popcll(pending & ((1ull << lane) - 1ull));
                        if (mine && rank < avail) {
                            // with a start order, position j of queue qi holds order entry j * DEG_NQUEUE + qi
                            // (< n_traj by the queue's limit): the lane keeps that entry's index
                            gid = a.order ? (c_next + rank - (long long)qi * slab) * DEG_NQUEUE + qi : c_next + rank;
                            want = false;
                            need_init = true;
                        }
                        const long long np_ = (long long)__popcll(pending);
                        c_next += (np_ < avail) ? np_ : avail;
                        pending = __ballot(want) & pending;
                    }
                }
                retire = want;  // every slab is exhausted
            }
        }
#if DEG_COMPACT
        (void)retire;
        {
            // tail compaction: only once this wave can claim no more ids
            bool wave_exit = false;
            // (lanes that were just handed an id still have to initialise it: not this iteration)
            if (qtries >= DEG_NQUEUE && __ballot(need_init) == 0ull) {
                const unsigned long long hm = __ballot(have);
                const unsigned long long wm = __ballot(want);
                const int ldr = __ffsll((unsigned long long)__ballot(1)) - 1;
                const int nact = __popcll(hm);
                const unsigned long long below = (1ull << lane) - 1ull;
                // one plain read of the mailbox word decides whether any atomic is worth trying
                const unsigned long long mw = deg_mb_read(&mb_word);
                const bool mb_free = ((mw >> 8) & 0xffull) == 0ull;
                const bool mb_has = ((mw >> 16) & 0xffull) > ((mw >> 24) & 0xffull);
                if (nact == 0 && !mb_free) {
                    // donated states are on their way: take them below
                } else if (nact == 0) {
                    int ok = 0;
                    if (lane == ldr) ok = deg_mb_retire(&mb_word);
                    wave_exit = __shfl(ok, ldr) != 0;
                } else if (nact <= DEG_COMPACT_LANES && mb_free && (mw & 0xffull) >= 2ull) {
                    int ok = 0;
                    if (lane == ldr) ok = deg_mb_donate(&mb_word, nact);
                    if (__shfl(ok, ldr) != 0) {
                        if (have) {
                            float* st = mb_state[__popcll(hm & below)];
                            st[0] = __builtin_bit_cast(float, (int)(i & 0xffffffffll));
                            st[1] = __builtin_bit_cast(float, (int)(i >> 32));
                            st[2] = t;
                            st[3] = tf;
                            st[4] = dt;
                            st[5] = tc;
                            st[6] = qold;
                            st[7] = qpow;
                            st[8] = __builtin_bit_cast(float, cur);
                            st[9] = __builtin_bit_cast(float, nacc);
                            st[10] = __builtin_bit_cast(float, nrej);
                            st[11] = __builtin_bit_cast(float, (int)(attempts & 0xffffffffll));
                            st[12] = __builtin_bit_cast(float, (int)(attempts >> 32));
                            st[13] = __builtin_bit_cast(float, have_k ? 1 : 0);
                            st[14] = tprev;
                            st[15] = dtacc;
                            DEG_FOR(c) st[16 + c] = u[c];
#if DEG_METHOD == 0
                            // the k1 of the next attempt: the FSAL stage after an accepted step,
                            // the unchanged k1 after a rejected one
                            DEG_FOR(c) st[16 + DEG_N + c] = fsal_pending ? k[7][c] : k[1][c];
#endif
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                        if (lane == ldr) atomicAdd(&mb_word, (unsigned long long)nact << 16);  // written
                        have = false;
                        wave_exit = true;
                    }
                }
                if (!wave_exit && wm != 0ull && mb_has) {
                    int base = 0, m = 0;
                    if (lane == ldr) m = deg_mb_take(&mb_word, __popcll(wm), &base);
                    m = __shfl(m, ldr);
                    base = __shfl(base, ldr);
                    if (m > 0) {
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                        const int rk = __popcll(wm & below);
                        if (want && rk < m) {
                            const float* st = mb_state[base + rk];
                            i = (long long)(unsigned int)__builtin_bit_cast(int, st[0]) |
                                ((long long)__builtin_bit_cast(int, st[1]) << 32);
                            t = st[2];
                            tf = st[3];
                            dt = st[4];
                            tc = st[5];
                            qold = st[6];
                            qpow = st[7];
                            cur = __builtin_bit_cast(int, st[8]);
                            nacc = __builtin_bit_cast(int, st[9]);
                            nrej = __builtin_bit_cast(int, st[10]);
                            attempts = (long long)(unsigned int)__builtin_bit_cast(int, st[11]) |
                                       ((long long)__builtin_bit_cast(int, st[12]) << 32);
                            const int fl = __builtin_bit_cast(int, st[13]);
                            have_k = (fl & 1) != 0;
                            fsal_pending = false;  // k1 arrives ready to use
                            tprev = st[14];
                            dtacc = st[15];
                            DEG_FOR(c) u[c] = st[16 + c];
#if DEG_METHOD == 0
                            DEG_FOR(c) k[1][c] = st[16 + DEG_N + c];
#endif
                            _Pragma("unroll") for (int j = 0; j < DEG_NP; ++j) p[j] = a.p[(long long)j * S + i];
                            status = DEG_RET_SUCCESS;
                            if (use_saveat) {
                                tnext = (cur < a.nsaveat) ? svp[cur] : finf;
                                ts_w = a.ts + (long long)cur * S + i;
                                us_w = a.us + (long long)cur * DEG_N * S + i;
                            }
                            have = true;
                            want = false;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // reads done
                        if (lane == ldr) deg_mb_consumed(&mb_word, m);
                    }
                }
            }
            if (wave_exit) break;  // wave-uniform
        }
#elif DEG_LEARNED
        // the matrix-core proposer needs every lane: the wave leaves together
        if (__ballot(!retire) == 0ull) break;
#else
        if (retire) break;
#endif
#endif
        long long oi = -1;
        if (need_init) {
            need_init = false;
            oi = a.order ? (long long)a.order[gid] : gid;
            if (oi < 0 || oi >= a.n_traj) {
                // an entry of a caller's start order outside [0, n_traj): nothing to solve
                oi = -1;
#if DEG_PERSISTENT
                want = true;  // the lane takes the next id instead
#else
                break;
#endif
            }
        }
        if (oi >= 0) {
            // ---- init (nonstiff/types.jl:305-402, stiff/types.jl:32-177; kernels.jl:35-45, :116-126)
            i = a.id_offset + oi;
#if DEG_PERSISTENT && DEG_ORDERED_INPUTS
            if (a.inputs_ordered != nullptr) {
                // the inputs of start-order position gid, laid out queue-major by the ordering pass: the
                // lanes that take consecutive entries of one queue read consecutive words
                const long long slab8 = (a.n_traj + DEG_NQUEUE - 1) / DEG_NQUEUE;
                const float* q = a.inputs_ordered + (gid % DEG_NQUEUE) * slab8 + gid / DEG_NQUEUE;
                DEG_FOR(c) u[c] = q[(long long)c * DEG_NQUEUE * slab8];
                _Pragma("unroll") for (int j = 0; j < DEG_NP; ++j) p[j] = q[(long long)(DEG_N + j) * DEG_NQUEUE * slab8];
            } else
#endif
            {
                DEG_FOR(c) u[c] = a.u0[(long long)c * S + i];
                _Pragma("unroll") for (int j = 0; j < DEG_NP; ++j) p[j] = a.p[(long long)j * S + i];
            }
            t = a.t0s ? a.t0s[i] : a.t0;
            tf = a.tfs ? a.tfs[i] : a.tf;
            tprev = t;
            dt = a.dt;
            dtacc = dt;
            tc = 0.0f;
            qold = DEG_QOLDINIT;
#if DEG_ADAPTIVE && DEG_FASTPOW
            qpow = -13.2877123795f;  // log2(qoldinit)
#elif DEG_ADAPTIVE
            qpow = ctl_pow(qold, DEG_BETA2);
#endif
            nacc = 0;
            nrej = 0;
            attempts = 0;
            status = DEG_RET_SUCCESS;
            have_k = false;
            fsal_pending = false;
#if DEG_TSTOPS
            tsi = 0;
#endif
#if DEG_NCCB > 0
            evlast = 1;  // nonstiff/types.jl:318, stiff/types.jl:44
#endif
#if DEG_LEARNED || DEG_STEPLOG
            lx_prev = -13.2877123795f;
            ldt_prev = 0.0f;
            first_attempt = true;
#endif
#if DEG_LEARNED
            zml_keep = 0.0f;  // a new trajectory starts from the plain PI proposal
            lpend = false;
#endif
#if DEG_STEPLOG
            have_prev = false;
#endif
            DEG_FOR(c) uprev[c] = u[c];
            if (use_saveat) {
                cur = 0;
                if (a.nsaveat > 0 && t == svp[0]) {
                    if (a.nsave > 0) {
                        DEG_FOR(c) a.us[(long long)c * S + i] = u[c];
                        a.ts[i] = a.t0;
                    }
                    cur = 1;
                }
                tnext = (cur < a.nsaveat) ? svp[cur] : finf;
                ts_w = a.ts + (long long)cur * S + i;
                us_w = a.us + (long long)cur * DEG_N * S + i;
            } else {
                a.ts[i] = t;
                DEG_FOR(c) a.us[(long long)c * S + i] = u[c];
                step_idx = 1;
            }
#if DEG_PERSISTENT
            have = true;
#endif
        }

        // Each iteration is one attempt (or one fixed step) per lane holding a trajectory:
        // (A) the stages and the error estimate, (B) the step controller, the saves and the
        // endpoint rules. In the learned variant the lane block is closed between A and B so
        // that the proposer runs on the matrix cores with all lanes of the wavefront.
#if DEG_BATCH_SAVES
        // interpolate every save point of the stashed step: rows [plo, phi)
        auto flush_stash = [&]() {
            float su[DEG_N], sk[DEG_NK + 1][DEG_N];
            DEG_FOR(c) su[c] = stash[c];
            _Pragma("unroll") for (int j = 1; j <= 7; ++j) DEG_FOR(c) sk[j][c] = stash[DEG_N * j + c];
            const float sdt = stash[8 * DEG_N], stp = stash[8 * DEG_N + 1];
            for (int r = plo; r < phi; ++r) {
                const float s_r = svp[r];
                float out[DEG_N];
#if DEG_PARITY
                const float th = (s_r - stp) / sdt;
#else
                const float th = (s_r - stp) * __builtin_amdgcn_rcpf(sdt);
#endif
                deg_interp(th, sdt, stp, p, su, su, sk, out);
                if (r < a.nsave) {
                    DEG_FOR(c) a.us[((long long)r * DEG_N + c) * S + i] = out[c];
                    a.ts[(long long)r * S + i] = s_r;
                }
            }
            stashed = false;
        };
        // batched dense output: once enough lanes of the wave hold a stash, write them all
        // (a lane that ended while holding a stash finalises only after its stash is written)
        if (__ballot(stashed && (status != DEG_RET_SUCCESS || !(t < tf))) != 0ull ||
            __popcll(__ballot(stashed)) >= DEG_SAVE_BATCH) {
            if (stashed) flush_stash();
        }
#endif
        bool done = false;
        bool accepted = false;
#if DEG_ADAPTIVE
        bool attempted = false;
        float EEst = 0.0f;
#endif
#if DEG_LEARNED
#if DEG_METHOD == 0
        // the FSAL stage moves into k1 now rather than at the attempt, so k7 is free during the proposer
        if (fsal_pending) {
            DEG_FOR(c) k[1][c] = k[7][c];
            fsal_pending = false;
        }
#endif
        {
            // the proposer, on the matrix cores with every lane of the wavefront (the loop exit is
            // wave-uniform in this variant), on the features of each lane's last attempt. It runs
            // here, before the next attempt, where the stage registers are free; between evaluations
            // each lane keeps its last correction, and lanes without a pending step get none
            if (mlp_tick == 0) {
                const float zo = deg_mlp_eval(deg_mlpw, lfeat);
                zml_keep = lpend ? zo : 0.0f;
            }
            if (++mlp_tick >= a.mlp_every) mlp_tick = 0;
            if (lpend) {
                // the pending step: the PI factor corrected in log2 space within the PI range
                // [qmin, qmax]; after a rejection never more than 0.9 dt; the PI step stands when the
                // proposal is not finite
                float zz = __builtin_fminf(__builtin_fmaxf(lzpi + zml_keep, -2.32192809f), 3.32192809f);
                if (!lacc) zz = __builtin_fminf(zz, -0.152003093f);
                const float fml = __builtin_amdgcn_exp2f(zz);
                if (fml == fml)
                    dt = lacc ? __builtin_fminf(__builtin_fabsf(ldt0 * fml), __builtin_fabsf(lrem)) : ldt0 * fml;
                lpend = false;
            }
        }
        float (&feat)[DEG_MLP_IN] = lfeat;
#elif DEG_STEPLOG
        float feat[DEG_MLP_IN] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#endif
#if DEG_ADAPTIVE
        // ---- step-size control and acceptance (gpu_*_perform_step.jl adaptive versions); run
        // right after the error estimate, or after the proposer in the learned variant
        auto control = [&]() {
            // ---- step-size control and acceptance (gpu_*_perform_step.jl adaptive versions)
#if DEG_FASTPOW
            // EEst holds EEst^2 here (scaled_error2_fast)
            if (EEst != EEst) status = DEG_RET_NONFINITE;
            const bool acc_step = !(EEst > 1.0f);
            float factor = deg_fast_control(EEst, acc_step, qpow);
#if DEG_LEARNED || DEG_STEPLOG
            const float zpi = __builtin_amdgcn_logf(factor);  // log2 of the PI factor
#endif
#if DEG_LEARNED
            // the next step is set by the proposer before the next attempt (the PI step stands until
            // then; the no-progress rule below decides with it)
            lpend = true;
            lacc = acc_step;
            lzpi = zpi;
            ldt0 = dt;
            lrem = (tf - t) - dt;
#endif
#if DEG_STEPLOG
            {
                // the previous attempt's row, completed by this attempt's outcome; slots are
                // claimed with one atomic per wavefront
                const bool wr = have_prev;
                const unsigned long long m = __ballot(wr);
                if (m != 0ull) {
                    const int ln = (int)__lane_id();
                    const int leader = __ffsll(m) - 1;
                    unsigned int base = 0u;
                    if (ln == leader) base = atomicAdd(a.steplog_count, (unsigned int)__popcll(m));
                    base = (unsigned int)__shfl((int)base, leader);
                    const long long slot = (long long)base + __popcll(m & ((1ull << ln) - 1ull));
                    if (wr && slot < a.steplog_cap) {
                        float4* row = (float4*)(a.steplog + slot * DEG_STEPLOG_COLS);
                        row[0] = make_float4(fprev[0], fprev[1], fprev[2], fprev[3]);
                        row[1] = make_float4(fprev[4], fprev[5], feat[2] - fprev[2], feat[0]);
                        row[2] = make_float4(zpi_prev, acc_prev ? 1.0f : 0.0f, 0.0f, 0.0f);
                    }
                }
                _Pragma("unroll") for (int j = 0; j < DEG_MLP_IN; ++j) fprev[j] = feat[j];
                zpi_prev = zpi;
                acc_prev = acc_step;
                have_prev = true;
            }
#endif
#if DEG_LEARNED || DEG_STEPLOG
            lx_prev = feat[0];
            ldt_prev = feat[2];
            first_attempt = false;
#endif
            if (!acc_step) {
                dt = dt * factor;
                ++nrej;
            } else {
                float dtnew = dt * factor;
                dtnew = __builtin_fminf(__builtin_fabsf(dtnew), __builtin_fabsf((tf - t) - dt));
#else
            float q11 = 0.0f, q;
            if (EEst == 0.0f) {
                q = 1.0f / DEG_QMAX;
            } else {
                q11 = ctl_pow(EEst, DEG_BETA1);
                q = q11 / qpow;
            }
            if (EEst > 1.0f) {
                dt = dt / jl_min(1.0f / DEG_QMIN, q11 / DEG_GAMMA);
                ++nrej;
            } else {
                q = jl_max(1.0f / DEG_QMAX, jl_min(1.0f / DEG_QMIN, q / DEG_GAMMA));
                qold = jl_max(EEst, DEG_QOLDINIT);
                qpow = ctl_pow(qold, DEG_BETA2);
                float dtnew = dt / q;
                dtnew = jl_min(__builtin_fabsf(dtnew), __builtin_fabsf((tf - t) - dt));
#endif
                dtacc = dt;
                tprev = t;
                DEG_FOR(c) {
                    uprev[c] = u[c];
                    u[c] = unew[c];
                }
                if ((tf - t) - dt < DEG_DTMIN) {
                    t = tf;
                    tc = 0.0f;
                }
#if DEG_TSTOPS
                else if (tsi < a.ntstops && ((a.tstops[tsi] - t) - dt) - DEG_TSTOP_SLACK < 0.0f) {
                    // gpu_tsit5_perform_step.jl:158-164: the step ends at the stopping time, with
                    // the state taken from this step's dense output there
                    t = a.tstops[tsi];
                    tc = 0.0f;
                    ++tsi;
                    float ut[DEG_N];
                    deg_interp((t - tprev) / dtacc, dtacc, tprev, p, uprev, u, k, ut);
                    DEG_FOR(c) u[c] = ut[c];
                }
#endif
                else {
                    float tn = deg_advance(t, dt, tc);
#if !DEG_KAHAN
                    // A step below the resolution of t whose successor is no larger would
                    // repeat forever: snap to tf when it is within two ulps, otherwise stop
                    // with DtLessThanMin. A growing dt goes on with t unchanged.
                    if (tn == t && !(dtnew > dt)) {
                        if (tf - t <= 2.0f * (deg_next_up(t) - t)) tn = tf;
                        else status = DEG_RET_DTMIN;
                    }
#endif
                    t = tn;
                }
                dt = dtnew;
                ++nacc;
                accepted = true;
            }
        };
#endif
#if DEG_PERSISTENT
        if (have) {
#else
        {
#endif
        done = !(t < tf);
#if DEG_BATCH_SAVES
        if (status != DEG_RET_SUCCESS) done = true;  // finalisation deferred for a stash
#endif
        if (!done) {
            if (attempts >= a.max_attempts) {
                status = DEG_RET_MAXITERS;
            }
#if DEG_ADAPTIVE
            else if (dt < DEG_DTMIN) {
                status = DEG_RET_DTMIN;
            }
#endif
            else {
                ++attempts;
#if DEG_ADAPTIVE
                // ---- one attempt of the `while EEst > 1` loop: stages and error estimate
#if DEG_METHOD == 0
                if (!have_k) {
                    deg_rhs(u, p, t, k[1]);
                    have_k = true;
                } else if (fsal_pending) {
                    DEG_FOR(c) k[1][c] = k[7][c];
                }
                fsal_pending = false;
                deg_attempt(u, p, t, dt, k, unew);
                EEst = deg_error(dt, k, u, unew, a.abstol, a.reltol);
#elif DEG_METHOD == 1 || DEG_METHOD == 3
                deg_rhs(u, p, t, k[1]);
                deg_attempt(u, p, t, dt, k, unew);
                EEst = deg_error(dt, k, u, unew, a.abstol, a.reltol);
#elif DEG_METHOD == 2
                EEst = deg_rb23<true>(u, p, t, dt, k, unew, a.abstol, a.reltol);
#elif DEG_METHOD == 6 || DEG_METHOD == 7
                {
                    // gpu_kvaerno*_perform_step.jl adaptive: k1 = f(uprev, t) in every attempt; the
                    // nonlinear solver runs at (t, dt); dense output keeps k1 and z_last / dt
                    float zl[DEG_N];
                    deg_rhs(u, p, t, k[1]);
                    EEst = deg_kvaerno<true>(u, p, t, dt, t, dt, k[1], zl, unew, a.abstol, a.reltol);
                    DEG_FOR(c) k[2][c] = zl[c] / dt;
                }
#else
                EEst = deg_rodas<true>(u, p, t, dt, k, unew, a.abstol, a.reltol);
#endif
                attempted = true;
#if DEG_LEARNED || DEG_STEPLOG
                {
                    // proposer features (base-2 logs; EEst holds EEst^2 on this path)
                    float su = 0.0f, sf = 0.0f;
                    DEG_FOR(c) {
                        su = __builtin_fmaf(u[c], u[c], su);
                        sf = __builtin_fmaf(k[1][c], k[1][c], sf);
                    }
                    feat[0] = 0.5f * __builtin_amdgcn_logf(__builtin_fmaxf(EEst, 1e-30f));
                    feat[1] = lx_prev;
                    feat[2] = __builtin_amdgcn_logf(dt);
                    feat[3] = 0.5f * __builtin_amdgcn_logf(__builtin_fmaxf(su * (1.0f / DEG_N), 1e-30f));
                    feat[4] = 0.5f * __builtin_amdgcn_logf(__builtin_fmaxf(sf * (1.0f / DEG_N), 1e-30f));
                    feat[5] = first_attempt ? 0.0f : feat[2] - ldt_prev;
                }
#endif
                control();
#else
                // ---- one fixed step (gpu_*_perform_step.jl fixed versions)
                const float told = t;
                const float tcomp_old = tc;
                float dts = dt;  // this step's length: shorter when it ends at a stopping time
                bool at_tstop = false;
#if DEG_TSTOPS
                // gpu_tsit5_perform_step.jl:18-27: the step ends at the next stopping time when that
                // lies within dt (+ 100 eps); the integrator's dt (dense output) stays dt
                at_tstop = tsi < a.ntstops && ((a.tstops[tsi] - t) - dt) - DEG_TSTOP_SLACK < 0.0f;
#endif
                float tn;
#if DEG_TSTOPS
                if (at_tstop) {
                    tn = a.tstops[tsi];
                    dts = tn - told;
                    ++tsi;
                    tc = 0.0f;
                } else
#endif
                {
                    tn = deg_advance(t, dt, tc);
                }
#if !DEG_KAHAN
                if (!at_tstop && tn == told) {
                    // the step is below the resolution of t: stop instead of looping forever
                    status = DEG_RET_DTMIN;
                    tc = tcomp_old;
                    --attempts;
                } else
#endif
                {
                DEG_FOR(c) uprev[c] = u[c];
                tprev = told;
                dtacc = dt;
                t = tn;
#if DEG_METHOD == 0
                if (!have_k) {
                    deg_rhs(uprev, p, told, k[1]);
                    have_k = true;
                } else {
                    DEG_FOR(c) k[1][c] = k[7][c];
                }
                deg_attempt(uprev, p, told, dts, k, unew);
#elif DEG_METHOD == 1 || DEG_METHOD == 3
                deg_rhs(uprev, p, told, k[1]);
                deg_attempt(uprev, p, told, dts, k, unew);
#elif DEG_METHOD == 2
                deg_rb23<false>(uprev, p, told, dts, k, unew, 0.0f, 0.0f);
#elif DEG_METHOD == 6 || DEG_METHOD == 7
                {
                    // gpu_kvaerno*_perform_step.jl fixed: k1 = f(uprev, t) after u_modified, else the
                    // last step's k1; the nonlinear solver is built with the advanced t and the
                    // integrator's dt; the stored k1 is f(u, t) at the step's start time t
                    float zl[DEG_N];
                    if (!have_k) {
                        deg_rhs(uprev, p, told, k[1]);
                        have_k = true;
                    }
                    deg_kvaerno<false>(uprev, p, told, dts, tn, dt, k[1], zl, unew, 0.0f, 0.0f);
                    DEG_FOR(c) k[2][c] = zl[c] / dts;
                    deg_rhs(unew, p, told, k[1]);
                }
#else
                deg_rodas<false>(uprev, p, told, dts, k, unew, 0.0f, 0.0f);
#endif
                DEG_FOR(c) u[c] = unew[c];
                ++nacc;
                accepted = true;
                }
#endif
            }
        }
#if DEG_BATCH_SAVES
        if (accepted && use_saveat) {
            const bool crossing = tnext <= t;
            const bool fin = (status != DEG_RET_SUCCESS) || !(t < tf);
            if (crossing) {
                if (stashed || fin) {
                    // the stash is taken or the lane ends this iteration: interpolate this step
                    // from the registers now (the stashed step is written below)
                    while (tnext <= t) {
                        float out[DEG_N];
#if DEG_PARITY
                        const float th = (tnext - tprev) / dtacc;
#else
                        const float th = (tnext - tprev) * __builtin_amdgcn_rcpf(dtacc);
#endif
                        deg_interp(th, dtacc, tprev, p, uprev, u, k, out);
                        if (cur < a.nsave) {
                            DEG_FOR(c) us_w[(long long)c * S] = out[c];
                            *ts_w = tnext;
                        }
                        ++cur;
                        ts_w += S;
                        us_w += (long long)DEG_N * S;
                        tnext = (cur < a.nsaveat) ? svp[cur] : finf;
                    }
                } else {
                    DEG_FOR(c) stash[c] = uprev[c];
                    _Pragma("unroll") for (int j = 1; j <= 7; ++j) DEG_FOR(c) stash[DEG_N * j + c] = k[j][c];
                    stash[8 * DEG_N] = dtacc;
                    stash[8 * DEG_N + 1] = tprev;
                    plo = cur;
                    while (tnext <= t) {
                        ++cur;
                        tnext = (cur < a.nsaveat) ? svp[cur] : finf;
                    }
                    phi = cur;
                    ts_w = a.ts + (long long)cur * S + i;
                    us_w = a.us + (long long)cur * DEG_N * S + i;
                    stashed = true;
                }
            }
        }
#endif
        if (accepted) {
            // ---- savevalues! (integrator_utils.jl:13-50)
            auto savevalues = [&]() {
                if (use_saveat) {
#if DEG_HAS_INTERP && !DEG_BATCH_SAVES
                    while (tnext <= t) {
                        // the following save time is read first, so that its load latency overlaps
                        // the interpolation (the LDS copy ends with a +inf sentinel at nsaveat)
#if DEG_SAVEAT_GLOBAL
                        const float tafter = (cur + 1 < a.nsaveat) ? svp[cur + 1] : finf;
#else
                        const float tafter = svp[min(cur + 1, a.nsaveat)];
#endif
                        float out[DEG_N];
#if DEG_PARITY
                        const float th = (tnext - tprev) / dtacc;
#else
                        const float th = (tnext - tprev) * __builtin_amdgcn_rcpf(dtacc);
#endif
                        deg_interp(th, dtacc, tprev, p, uprev, u, k, out);
                        if (cur < a.nsave) {
                            DEG_FOR(c) us_w[(long long)c * S] = out[c];
                            *ts_w = tnext;
                        }
                        ++cur;
                        ts_w += S;
                        us_w += (long long)DEG_N * S;
                        tnext = tafter;
                    }
#endif
                } else if (a.save_everystep) {
#if !DEG_ADAPTIVE
                    if (step_idx >= 0 && step_idx < a.nsave) {
                        DEG_FOR(c) a.us[((long long)step_idx * DEG_N + c) * S + i] = u[c];
                        a.ts[(long long)step_idx * S + i] = t;
                    }
                    ++step_idx;
#endif
                }
            };
            bool saved_in_cb = false;
#if DEG_CALLBACKS
            if (status == DEG_RET_SUCCESS) {
                // ---- handle_callbacks! (integrator_utils.jl:274-325), after the accepted step
                auto apply_affect = [&](int fl) {
                    if (fl & DEG_AFF_TERMINATE) status = DEG_RET_TERMINATED;
                    if (fl & DEG_AFF_KEEP_FSAL) have_k = true;  // u_modified!(integrator, false)
                };
                bool csaved = false;
#if DEG_NCCB > 0
                {
                    float tmin = t, upc = 0.0f;
                    bool occ = false;
                    int idx = 1;
                    _Pragma("unroll") for (int j = 0; j < DEG_NCCB; ++j) {
                        // DiffEqBase.get_condition (integrator_utils.jl:460-478): the step's end
                        // state, its start state, or the step's dense output
                        auto cond_at = [&](float s) -> float {
                            if (s == t) return deg_ccb_cond(j, u, p, s);
                            if (s == tprev) return deg_ccb_cond(j, uprev, p, s);
                            float us_[DEG_N];
                            deg_interp((s - tprev) / dtacc, dtacc, tprev, p, uprev, u, k, us_);
                            return deg_ccb_cond(j, us_, p, s);
                        };
                        // find_callback_time (integrator_utils.jl:386-446); last_event_error is 0
                        float bt = tprev;
                        float bc = deg_ccb_cond(j, uprev, p, tprev);
                        if (evlast == j + 1 && __builtin_fabsf(bc - 0.0f) <= deg_ccb_abstol(j)) {
                            bt = tprev + dtacc * deg_ccb_nudge(j);
                            bc = cond_at(bt);
                        }
                        const float bs = jl_sign(bc);
                        const float tsg = jl_sign(cond_at(t));
                        const int mode = deg_ccb_mode(j);
                        const bool oc = ((bs < 0.0f && (mode & 1)) || (bs > 0.0f && (mode & 2))) && bs * tsg <= 0.0f;
                        const int rf = (mode >> 2) & 3;
                        float tj = t;
                        if (oc && rf != 0 && tsg != 0.0f) {
                            // gpu_find_root (integrator_utils.jl:327-384): ITP, k1 = 0.2/span0, k2 = 2, n0 = 10
                            float left = bt, right = t;
                            float fl = cond_at(left), fr = cond_at(right);
                            const float span0 = right - left;
                            const float k1 = 0.2f / span0;
                            float eps_s = span0 * 512.0f;
                            for (int it = 0; it < 100; ++it) {
                                const float span = right - left;
                                const float mid = (left + right) / 2.0f;
                                const float r = eps_s - span / 2.0f;
                                const float xf = left + (span * fl) / (fl - fr);
                                const float delta = jl_max((k1 * span) * span, jl_eps(xf));
                                const float diff = mid - xf;
                                const float xt = (delta <= __builtin_fabsf(diff)) ? xf + __builtin_copysignf(delta, diff) : mid;
                                const float xp = (__builtin_fabsf(xt - mid) <= r) ? xt : mid - __builtin_copysignf(r, diff);
                                const float yp = cond_at(xp);
                                const float yps = yp * jl_sign(fr);
                                if (yps > 0.0f) {
                                    right = xp;
                                    fr = yp;
                                } else if (yps < 0.0f) {
                                    left = xp;
                                    fl = yp;
                                } else {
                                    left = xp;
                                    right = xp;
                                    break;
                                }
                                eps_s = eps_s / 2.0f;
                                if (jl_nextfloat(left) >= right) break;
                            }
                            tj = (rf == 1) ? left : right;
                        }
                        // DiffEqBase.find_first_continuous_callback: the earliest event, the first
                        // callback on ties
                        if (j == 0) {
                            tmin = tj;
                            upc = bs;
                            occ = oc;
                            idx = 1;
                        } else if (oc && (tj < tmin || !occ)) {
                            tmin = tj;
                            upc = bs;
                            occ = true;
                            idx = j + 1;
                        }
                    }
                    if (occ) {
                        evlast = idx;
                        // apply_callback! (integrator_utils.jl:232-272). The affect applied is the
                        // first continuous callback's, whichever callback found the event.
                        if (tmin < tprev) {
                            status = DEG_RET_FAILURE;  // change_t_via_interpolation! raises here
                        } else {
                            if (tmin != t) {
                                // _change_t_via_interpolation! (integrator_utils.jl:174-201)
                                float ut[DEG_N];
                                deg_interp((tmin - tprev) / dtacc, dtacc, tprev, p, uprev, u, k, ut);
                                DEG_FOR(c) u[c] = ut[c];
                                step_idx -= (int)__builtin_rintf((t - tmin) / dtacc);
                                t = tmin;
                                tc = 0.0f;
                            }
#if DEG_ADAPTIVE
                            if (dt < 1.0e-12f) dt = jl_min(deg_ccb_dtrelax(0) * dtacc, __builtin_fabsf(tf - t));
#endif
                            savevalues();
                            csaved = true;
                            have_k = false;  // u_modified = true
                            const int mode0 = deg_ccb_mode(0);
                            if (upc < 0.0f) {
                                if (mode0 & 1) apply_affect(deg_ccb_affect(0, 0, u, p, t));
                                else have_k = true;
                            } else if (upc > 0.0f) {
                                if (mode0 & 2) apply_affect(deg_ccb_affect(0, 1, u, p, t));
                                else have_k = true;
                            }
                        }
                    } else {
                        evlast = 0;
                    }
                }
#endif
#if DEG_NDCB > 0
                // apply_discrete_callback! (integrator_utils.jl:70-149), in order
                bool dsaved = false;
                _Pragma("unroll") for (int j = 0; j < DEG_NDCB; ++j) {
                    if (deg_dcb_cond(j, u, p, t)) {
                        savevalues();
                        dsaved = true;
                        have_k = false;  // u_modified = true
                        apply_affect(deg_dcb_affect(j, u, p, t));
                    }
                }
                saved_in_cb = dsaved;  // the discrete result replaces the continuous one
                (void)csaved;
#else
                saved_in_cb = csaved;
#endif
            }
#endif
            if (!saved_in_cb) savevalues();
            fsal_pending = true;
        }
        if (!done) done = (status != DEG_RET_SUCCESS) || !(t < tf);

#if DEG_BATCH_SAVES
        if (done && !stashed) {
#else
        if (done) {
#endif
            // ---- endpoint rules (kernels.jl:53-62 / :133-142) and unsaved rows
            if (use_saveat) {
                for (int r = cur; r < a.nsave; ++r) a.ts[(long long)r * S + i] = a.t0;
            } else if (a.save_everystep) {
                for (int r = step_idx; r < a.nsave; ++r) a.ts[(long long)r * S + i] = a.t0;
#if !DEG_ADAPTIVE && DEG_HAS_INTERP
                if (t > tf && (status == DEG_RET_SUCCESS || status == DEG_RET_TERMINATED) && a.nsave > 1) {
                    // k[1] holds k7 after the FSAL shift for Tsit5; undo it for the dense output
                    float out[DEG_N];
                    const float th = (tf - tprev) / dtacc;
                    deg_interp(th, dtacc, tprev, p, uprev, u, k, out);
                    const long long r = a.nsave - 1;
                    DEG_FOR(c) a.us[(r * DEG_N + c) * S + i] = out[c];
                    a.ts[r * S + i] = tf;
                }
#endif
            } else if (a.nsave > 1) {
                a.ts[S + i] = t;
                DEG_FOR(c) a.us[((long long)DEG_N + c) * S + i] = u[c];
            }
            if (status == DEG_RET_SUCCESS) {
                bool fin = __builtin_isfinite(t);
                DEG_FOR(c) fin = fin && __builtin_isfinite(u[c]);
                if (!fin) status = DEG_RET_NONFINITE;
            }
            if (a.retcode) a.retcode[i] = status;
            if (a.naccept) a.naccept[i] = nacc;
            if (a.nreject) a.nreject[i] = nrej;
#if DEG_PERSISTENT
            have = false;
            want = true;
#else
            break;
#endif
        }
        }  // lanes holding a trajectory
    }
}

// Start-order cost key of each trajectory: its span over the initial step estimate of Hairer,
// Norsett and Wanner (Solving Ordinary Differential Equations I, II.4): with sc = abstol + |u0| reltol,
// d0 = ||u0 / sc||, d1 = ||f(u0, p, t0) / sc|| (RMS norms), h0 = 0.01 d0 / d1, or 1e-6 when d0 or d1
// is below 1e-5. With a cost model (model != NULL) the key is instead the predicted attempt count
// 2^(w . x) over the features x = [1, log2(span / h0), log2(|u0_c| + 2^-20) for each c,
// log2(|p_j| + 2^-20) for each j, log2 span], DEG_COST_NFEAT = 3 + DEG_N + DEG_NP weights, summed in
// that order. Sorting by the key (descending) starts the trajectories that are likely to take the most
// steps first; a non-finite key becomes 0 (started last).
#define DEG_COST_NFEAT (3 + DEG_N + DEG_NP)
__device__ __forceinline__ float deg_cost_log2(float x) {
    return __builtin_amdgcn_logf(__builtin_fabsf(x) + 9.5367431640625e-07f);  // |x| + 2^-20
}
extern "C" __global__ void __launch_bounds__(256) DEG_COST_KEY(const deg_kargs a, float* __restrict__ key,
                                                               const float* __restrict__ model) {
    const long long li = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= a.n_traj) return;
    const long long i = a.id_offset + li;
    const long long S = a.stride;
    float u[DEG_N], f0[DEG_N];
    float p[DEG_NP > 0 ? DEG_NP : 1];
    DEG_FOR(c) u[c] = a.u0[(long long)c * S + i];
    _Pragma("unroll") for (int j = 0; j < DEG_NP; ++j) p[j] = a.p[(long long)j * S + i];
    const float t0 = a.t0s ? a.t0s[i] : a.t0;
    const float tf = a.tfs ? a.tfs[i] : a.tf;
    deg_rhs(u, p, t0, f0);
    float d0 = 0.0f, d1 = 0.0f;
    DEG_FOR(c) {
        const float sc = a.abstol + __builtin_fabsf(u[c]) * a.reltol;
        const float x = u[c] / sc, y = f0[c] / sc;
        d0 += x * x;
        d1 += y * y;
    }
    d0 = __builtin_sqrtf(d0 / (float)DEG_N);
    d1 = __builtin_sqrtf(d1 / (float)DEG_N);
    const float h0 = (d0 < 1e-5f || d1 < 1e-5f) ? 1e-6f : 0.01f * d0 / d1;
    const float span = __builtin_fabsf(tf - t0);
    float k = span / h0;
    if (model != nullptr) {
        float z = model[0] + model[1] * __builtin_amdgcn_logf(__builtin_fmaxf(k, 1e-30f));
        DEG_FOR(c) z = z + model[2 + c] * deg_cost_log2(u[c]);
        _Pragma("unroll") for (int j = 0; j < DEG_NP; ++j) z = z + model[2 + DEG_N + j] * deg_cost_log2(p[j]);
        z = z + model[2 + DEG_N + DEG_NP] * __builtin_amdgcn_logf(__builtin_fmaxf(span, 1e-30f));
        k = __builtin_amdgcn_exp2f(__builtin_fminf(z, 126.0f));
    }
    key[li] = (k == k && k < 3.0e38f) ? k : (k == k ? 3.0e38f : 0.0f);
}

// Clears the persistent work counters before a launch (a kernel, not a memset, so the
// sequence can be captured in a graph).
extern "C" __global__ void DEG_QUEUE_RESET(unsigned int* q) {
    if (blockIdx.x == 0 && threadIdx.x < DEG_NQUEUE) q[threadIdx.x * DEG_QSTRIDE] = 0u;
}
