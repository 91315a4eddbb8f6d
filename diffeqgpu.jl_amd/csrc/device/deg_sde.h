// Per-lane fixed-step ensemble SDE solver for gfx950 (GPUEM, GPUSIEA), compiled at run time by
// hiprtc together with the traced drift and diffusion. One trajectory per lane; state in VGPRs;
// the Wiener increments come from deg_philox.h, keyed by the problem seed and counted by
// (step, block, trajectory id), so no generator state is carried or stored.
//
// Reference behaviour restated here (DiffEqGPU src/ensemblegpukernel/):
//   perform_step/gpu_em_perform_step.jl:1-74     em_kernel: initial save, step count, EM step,
//                                                diagonal and non-diagonal noise, linear saveat
//   perform_step/gpu_siea_perform_step.jl:1-139  SIEA constants and siea_kernel (diagonal noise)
//
// Configuration macros supplied by the host before this file:
//   DEG_N, DEG_NP        state and parameter counts
//   DEG_METHOD           8 Euler-Maruyama, 9 SIEA
//   DEG_WAVES_MIN        register cap: at least this many waves per SIMD (0: none)
//   DEG_KERNEL           kernel symbol name
// and from the traced source:
//   DEG_NM               noise columns (diagonal noise: DEG_N)
//   DEG_DIAG_NOISE       1: g returns DEG_N entries multiplied elementwise with the increments;
//                        0: g returns the DEG_N x DEG_NM matrix, row-major g[r * DEG_NM + c]
//   deg_rhs(u, p, t, du) drift f, deg_noise(u, p, t, g) diffusion g
// Without DEG_F_PARITY the host compiles with FP contraction and DEG_SDE_FASTNORM, which draws the
// normals through the hardware transcendentals (deg_philox.h); with it every operation, the normals
// included, is the oracle's Float32 sequence.
#pragma once

#define DEG_RET_SUCCESS 0
#define DEG_RET_MAXITERS 2
#define DEG_RET_NONFINITE 3

#define DEG_EM 8
#define DEG_SIEA 9

#ifndef DEG_WAVES_MIN
#define DEG_WAVES_MIN 0
#endif
#if DEG_WAVES_MIN > 0
#define DEG_ATTR_WAVES __attribute__((amdgpu_waves_per_eu(DEG_WAVES_MIN)))
#else
#define DEG_ATTR_WAVES
#endif

#if DEG_METHOD == DEG_SIEA && !DEG_DIAG_NOISE
#error "GPUSIEA supports diagonal noise only"
#endif

#define DEG_SFOR(c) _Pragma("unroll") for (int c = 0; c < DEG_N; ++c)

extern "C" __global__ void __launch_bounds__(256) DEG_ATTR_WAVES DEG_KERNEL(const deg_kargs a) {
    const long long li = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= a.n_traj) return;
    const long long i = a.id_offset + li;
    const long long S = a.stride;
    float u[DEG_N];
    float p[DEG_NP > 0 ? DEG_NP : 1];
    DEG_SFOR(c) u[c] = a.u0[(long long)c * S + i];
    _Pragma("unroll") for (int j = 0; j < DEG_NP; ++j) p[j] = a.p[(long long)j * S + i];
    const float t0 = a.t0s ? a.t0s[i] : a.t0;
    const float tf = a.tfs ? a.tfs[i] : a.tf;
    const unsigned long long seed = a.seeds ? a.seeds[i] : a.seed;
    const unsigned long long id = (unsigned long long)(a.traj_base + i);
    const float dt = a.dt;
    const float sqdt = __builtin_sqrtf(dt);
    const bool use_saveat = a.saveat != nullptr;
    int status = DEG_RET_SUCCESS;

    // n = floor(Int, abs(tf - t0) / abs(dt)) + 1 points, n - 1 steps (gpu_em_perform_step.jl:43-44)
    const float ratio = __builtin_fabsf(tf - t0) / __builtin_fabsf(dt);
    long long nsteps = 0;
    if (ratio < 4294967296.0f)
        nsteps = (long long)__builtin_floorf(ratio);
    else
        status = DEG_RET_MAXITERS;  // more steps than the counter space (or a NaN span)

    // initial save (gpu_em_perform_step.jl:25-35)
    int cur = 0;
    if (use_saveat) {
        if (a.nsaveat > 0 && t0 == a.saveat[0]) {
            if (a.nsave > 0) {
                DEG_SFOR(c) a.us[(long long)c * S + i] = u[c];
                a.ts[i] = a.t0;
            }
            cur = 1;
        }
    } else {
        if (a.nsave > 0) {
            a.ts[i] = t0;
            DEG_SFOR(c) a.us[(long long)c * S + i] = u[c];
        }
    }

    float t = t0;
    long long row = 1;  // next every-step row
    for (long long j = 0; j < nsteps; ++j) {
        float uprev[DEG_N];
        DEG_SFOR(c) uprev[c] = u[c];
        float w[DEG_NM];
        deg_step_normals<DEG_NM>(seed, id, (unsigned int)j, w);
#if DEG_METHOD == DEG_EM
        // u = uprev + f(uprev, p, t) * dt + sqdt * g(uprev, p, t) .* randn   (diagonal)
        // u = uprev + f(uprev, p, t) * dt + sqdt * g(uprev, p, t) * randn    (n x m matrix)
        float fu[DEG_N];
        deg_rhs(uprev, p, t, fu);
#if DEG_DIAG_NOISE
        float gu[DEG_N];
        deg_noise(uprev, p, t, gu);
        DEG_SFOR(c) u[c] = (uprev[c] + fu[c] * dt) + (sqdt * gu[c]) * w[c];
#else
        float gu[DEG_N * DEG_NM];
        deg_noise(uprev, p, t, gu);
        DEG_SFOR(r) {
            float acc = (sqdt * gu[r * DEG_NM]) * w[0];
            _Pragma("unroll") for (int c = 1; c < DEG_NM; ++c) acc = acc + (sqdt * gu[r * DEG_NM + c]) * w[c];
            u[r] = (uprev[r] + fu[r] * dt) + acc;
        }
#endif
#else
        // SIEA (gpu_siea_perform_step.jl:98-119) with alpha1 = alpha2 = gamma1 = 1/2,
        // lambda = (1/4, -1/4, 1/4), mu = (1/4, 1/4, -1/4), mu0 = mubar0 = lambda0 = lambdabar0 = 1,
        // nu1 = 1, nu2 = 0, beta2 = 1, beta3 = 0, delta2 = -1, delta3 = 0
        const float al1 = 0.5f, al2 = 0.5f, ga1 = 0.5f;
        const float la1 = 0.25f, la2 = -0.25f, la3 = 0.25f;
        const float mu1 = 0.25f, mu2 = 0.25f, mu3 = -0.25f;
        const float mu0 = 1.0f, mub0 = 1.0f, la0 = 1.0f, lab0 = 1.0f;
        const float nu1 = 1.0f, nu2 = 0.0f, be2 = 1.0f, be3 = 0.0f, de2 = -1.0f, de3 = 0.0f;
        float k0[DEG_N], g0[DEG_N], dW[DEG_N], W2[DEG_N], tmp[DEG_N];
        deg_rhs(uprev, p, t, k0);
        deg_noise(uprev, p, t, g0);
        DEG_SFOR(c) {
            dW[c] = sqdt * w[c];
            W2[c] = (dW[c] * dW[c]) / sqdt;
            const float W3 = (nu2 * ((dW[c] * dW[c]) * dW[c])) / dt;
            tmp[c] = ((uprev[c] + (la0 * k0[c]) * dt) + (nu1 * g0[c]) * dW[c]) + g0[c] * W3;
        }
        float k1[DEG_N];
        deg_rhs(tmp, p, t + mu0 * dt, k1);
        DEG_SFOR(c) tmp[c] = ((uprev[c] + (lab0 * k0[c]) * dt) + (be2 * g0[c]) * sqdt) + (be3 * g0[c]) * W2[c];
        float g1[DEG_N];
        deg_noise(tmp, p, t + mub0 * dt, g1);
        DEG_SFOR(c) tmp[c] = ((uprev[c] + (lab0 * k0[c]) * dt) + (de2 * g0[c]) * sqdt) + (de3 * g0[c]) * W2[c];
        float g2[DEG_N];
        deg_noise(tmp, p, t + mub0 * dt, g2);
        DEG_SFOR(c) {
            const float ud = uprev[c] + (al1 * k0[c] + al2 * k1[c]) * dt;
            const float A = (ga1 * g0[c]) * dW[c];
            const float B = (((la1 * dW[c]) + la2 * sqdt) + la3 * W2[c]) * g1[c];
            const float C = (((mu1 * dW[c]) + mu2 * sqdt) + mu3 * W2[c]) * g2[c];
            u[c] = ud + ((A + B) + C);
        }
#endif
        t = t + dt;
        if (!use_saveat) {
            if (a.save_everystep) {
                if (row < a.nsave) {
                    DEG_SFOR(c) a.us[(row * DEG_N + c) * S + i] = u[c];
                    a.ts[row * S + i] = t;
                }
                ++row;
            }
        } else {
            // linear interpolation onto the save points passed by this step (gpu_em_perform_step.jl:58-66)
            while (cur < a.nsaveat && a.saveat[cur] <= t) {
                const float savet = a.saveat[cur];
                const float th = (savet - (t - dt)) / dt;
                if (cur < a.nsave) {
                    DEG_SFOR(c) a.us[((long long)cur * DEG_N + c) * S + i] = uprev[c] + (u[c] - uprev[c]) * th;
                    a.ts[(long long)cur * S + i] = savet;
                }
                ++cur;
            }
        }
    }

    // endpoint rule (gpu_em_perform_step.jl:69-72) and the t0 prefill of unsaved rows
    if (use_saveat) {
        for (long long r = cur; r < a.nsave; ++r) a.ts[r * S + i] = a.t0;
    } else if (a.save_everystep) {
        for (long long r = row; r < a.nsave; ++r) a.ts[r * S + i] = a.t0;
    } else if (a.nsave > 1) {
        a.ts[S + i] = t;
        DEG_SFOR(c) a.us[((long long)DEG_N + c) * S + i] = u[c];
    }
    if (status == DEG_RET_SUCCESS) {
        bool fin = __builtin_isfinite(t);
        DEG_SFOR(c) fin = fin && __builtin_isfinite(u[c]);
        if (!fin) status = DEG_RET_NONFINITE;
    }
    if (a.retcode) a.retcode[i] = status;
    if (a.naccept) a.naccept[i] = (int)(nsteps < 0x7fffffffll ? nsteps : 0x7fffffffll);
    if (a.nreject) a.nreject[i] = 0;
}
