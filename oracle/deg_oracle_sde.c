/* deg_oracle_sde.c -- CPU restatement of DiffEqGPU's EnsembleGPUKernel SDE kernels.
 *
 * TEST INFRASTRUCTURE. Only the test suite, the smoke check and the benchmark's CPU-baseline
 * leg load it, as the checker of the HIP SDE kernels. The product never calls it.
 *
 * Restated, operation for operation in IEEE Float32 (build with -ffp-contract=off):
 *   em_kernel: initial save, step count, Euler-Maruyama step with diagonal and non-diagonal
 *   noise, linear-interpolation saveat, endpoint rule   perform_step/gpu_em_perform_step.jl:1-74
 *   SIEA constants and siea_kernel (diagonal noise)     perform_step/gpu_siea_perform_step.jl:1-139
 * The Wiener increments are sqrt(dt) times standard normals from Philox4x32-10 (Salmon, Moraes,
 * Dror and Shaw, SC11; known-answer vectors of the Random123 distribution) keyed by the
 * problem seed, counter (step, block, trajectory id low, high), turned into normals by a
 * Box-Muller transform whose ln and sincos are written here from IEEE +, -, *, / only. The
 * reference draws from the GPU's device `randn`, whose streams are not reproducible outside
 * Julia: the oracle fixes the generator the kernels use, and the tests pin its distribution.
 *
 * Test systems (drift f, diffusion g; Float32, the reference test problems' association order):
 *   gbm       f = p0 u, g = p1 u                 test/gpu_kernel_de/gpu_sde_convergence.jl:8-10
 *   gbm1      f = u, g = u                       test/gpu_kernel_de/gpu_sde_regression.jl:13-15
 *   lorenz    Lorenz drift with parameters p, g = (3, 3, 3)   gpu_sde_regression.jl:48-58
 *   nondiag   f = 1.01 u, g the 2 x 4 matrix of gpu_sde_regression.jl:106-117 (SMatrix{2,4}
 *             filled column-major from its eight arguments)
 *   ou        f = p0 (p1 - u), g = p2            Ornstein-Uhlenbeck (closed-form moments)
 *   bm        f = 0, g = p0                      scaled Brownian motion
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define OS_MAXN 8
#define OS_MAXM 8

/* ---- Philox4x32-10 ---- */
void deg_oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline float bits_f(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}
static inline uint32_t f_bits(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}

/* -2 ln U, U = ((w >> 8) + 1) 2^-24: U = 2^e m, m in [sqrt(1/2), sqrt(2)], ln m = 2 atanh(s) */
float deg_oracle_bm_m2log(uint32_t w) {
    const float u = (float)((w >> 8) + 1u) * 5.9604644775390625e-08f;
    const uint32_t b = f_bits(u);
    int e = (int)(b >> 23) - 127;
    float m = bits_f((b & 0x007fffffu) | 0x3f800000u);
    if ((b & 0x007fffffu) > 0x003504f3u) {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    /* 2/3, 2/5, 2/7, 2/9, 2/11 rounded to Float32 */
    const float lnm = s * (2.0f + z * (0.666666687f + z * (0.400000006f + z * (0.285714298f + z * (0.222222224f + z * 0.181818187f)))));
    const float fe = (float)e;
    /* ln 2 = 0.693145752 (15 significant bits, so fe * head is exact) + 1.42860677e-06 */
    const float ln = fe * 0.693145752f + (fe * 1.42860677e-06f + lnm);
    return -2.0f * ln;
}

/* (cos, sin) of 2 pi (w >> 8) 2^-24: quadrant and reflection from the integer angle */
void deg_oracle_bm_sincos(uint32_t w, float *co, float *si) {
    const uint32_t k = w >> 8;
    const uint32_t q = k >> 22;
    uint32_t r = k & 0x3fffffu;
    const int refl = r > 0x200000u;
    if (refl) r = 0x400000u - r;
    const float x = (float)r * 3.74507028e-07f; /* fl(2 pi) 2^-24 */
    const float x2 = x * x;
    /* Taylor coefficients 1/k! rounded to Float32 */
    const float sx = x * (1.0f + x2 * (-0.166666672f + x2 * (0.00833333377f + x2 * (-0.000198412701f + x2 * 2.75573188e-06f))));
    const float cx = 1.0f + x2 * (-0.5f + x2 * (0.0416666679f + x2 * (-0.00138888892f + x2 * (2.48015876e-05f + x2 * -2.75573188e-07f))));
    const float cp = refl ? sx : cx;
    const float sp = refl ? cx : sx;
    switch (q) {
        case 0: *co = cp; *si = sp; break;
        case 1: *co = -sp; *si = cp; break;
        case 2: *co = -cp; *si = -sp; break;
        default: *co = sp; *si = -cp; break;
    }
}

static void bm_pair(uint32_t w0, uint32_t w1, float *z0, float *z1) {
    const float rr = sqrtf(deg_oracle_bm_m2log(w0));
    float co, si;
    deg_oracle_bm_sincos(w1, &co, &si);
    *z0 = rr * co;
    *z1 = rr * si;
}

/* the m normals of one step of stream `id` */
static void step_normals(uint64_t seed, uint64_t id, uint32_t step, int m, float *w) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int b = 0; 4 * b < m; ++b) {
        const uint32_t ctr[4] = {step, (uint32_t)b, (uint32_t)id, (uint32_t)(id >> 32)};
        uint32_t r[4];
        deg_oracle_philox(ctr, key, r);
        float z[4];
        bm_pair(r[0], r[1], &z[0], &z[1]);
        bm_pair(r[2], r[3], &z[2], &z[3]);
        for (int j = 0; j < 4 && 4 * b + j < m; ++j) w[4 * b + j] = z[j];
    }
}

/* out[c * n + i] = normal c of step `step`, stream id0 + i */
void deg_oracle_normals(uint64_t seed, int64_t id0, int64_t n, uint32_t step, int m, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float w[4];
        const int mm = m < 4096 ? m : 4096;
        const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        const uint64_t id = (uint64_t)(id0 + i);
        for (int b = 0; 4 * b < mm; ++b) {
            const uint32_t ctr[4] = {step, (uint32_t)b, (uint32_t)id, (uint32_t)(id >> 32)};
            uint32_t r[4];
            deg_oracle_philox(ctr, key, r);
            bm_pair(r[0], r[1], &w[0], &w[1]);
            bm_pair(r[2], r[3], &w[2], &w[3]);
            for (int j = 0; j < 4 && 4 * b + j < mm; ++j) out[(int64_t)(4 * b + j) * n + i] = w[j];
        }
    }
}

/* ---- test systems ---- */
enum { OSDE_GBM = 0, OSDE_GBM1 = 1, OSDE_LORENZ = 2, OSDE_NONDIAG = 3, OSDE_OU = 4, OSDE_BM = 5, OSDE_COUNT = 6 };

static void sde_f(int sys, const float *u, const float *p, float t, float *du) {
    (void)t;
    switch (sys) {
        case OSDE_GBM: du[0] = p[0] * u[0]; break;
        case OSDE_GBM1: du[0] = u[0]; break;
        case OSDE_LORENZ:
            du[0] = p[0] * (u[1] - u[0]);
            du[1] = u[0] * (p[1] - u[2]) - u[1];
            du[2] = u[0] * u[1] - p[2] * u[2];
            break;
        case OSDE_NONDIAG: du[0] = 1.01f * u[0]; du[1] = 1.01f * u[1]; break;
        case OSDE_OU: du[0] = p[0] * (p[1] - u[0]); break;
        default: du[0] = 0.0f; break;
    }
}

/* diagonal: n entries; non-diagonal: n x m row-major */
static void sde_g(int sys, const float *u, const float *p, float t, float *g) {
    (void)t;
    switch (sys) {
        case OSDE_GBM: g[0] = p[1] * u[0]; break;
        case OSDE_GBM1: g[0] = u[0]; break;
        case OSDE_LORENZ: g[0] = 3.0f; g[1] = 3.0f; g[2] = 3.0f; break;
        case OSDE_NONDIAG: {
            /* SMatrix{2,4}(a1, ..., a8) is column-major: row 1 = (a1, a3, a5, a7), row 2 = (a2, a4, a6, a8) */
            const float a1 = 0.3f * u[0], a2 = 0.6f * u[0], a3 = 0.9f * u[0], a4 = 0.12f * u[0];
            const float a5 = 1.2f * u[1], a6 = 0.2f * u[1], a7 = 0.3f * u[1], a8 = 1.8f * u[1];
            g[0] = a1; g[1] = a3; g[2] = a5; g[3] = a7;
            g[4] = a2; g[5] = a4; g[6] = a6; g[7] = a8;
            break;
        }
        case OSDE_OU: g[0] = p[2]; break;
        default: g[0] = p[0]; break;
    }
}

static void sde_dims(int sys, int *n, int *m, int *diag) {
    switch (sys) {
        case OSDE_LORENZ: *n = 3; *m = 3; *diag = 1; break;
        case OSDE_NONDIAG: *n = 2; *m = 4; *diag = 0; break;
        default: *n = 1; *m = 1; *diag = 1; break;
    }
}

int deg_oracle_sde_dims(int system, int *n, int *m, int *diag) {
    if (system < 0 || system >= OSDE_COUNT) return -1;
    sde_dims(system, n, m, diag);
    return 0;
}

/* One trajectory. Layout as the kernels: u0[c*N + i], p[j*N + i], ts[row*N + i],
 * us[(row*n + c)*N + i]. method 8 EM, 9 SIEA. wsum (optional) receives sum_j dW_j per noise
 * component in double, wsum[c*N + i]. Returns the per-trajectory return code. */
static int sde_traj(int sys, int method, int64_t N, int64_t i, const float *u0, const float *p, float t0, float tf,
                    float t0_fill, float dt, uint64_t seed, uint64_t id, const float *saveat, int nsaveat, int nsave,
                    int save_everystep, float *ts, float *us, double *wsum, int32_t *naccept) {
    int n, m, diag;
    sde_dims(sys, &n, &m, &diag);
    float u[OS_MAXN], pp[4] = {0, 0, 0, 0};
    for (int c = 0; c < n; ++c) u[c] = u0[(int64_t)c * N + i];
    const int np = (sys == OSDE_GBM) ? 2 : (sys == OSDE_LORENZ || sys == OSDE_OU) ? 3 : (sys == OSDE_BM) ? 1 : 0;
    for (int j = 0; j < np; ++j) pp[j] = p[(int64_t)j * N + i];
    const float sqdt = sqrtf(dt);
    const int use_saveat = saveat != NULL;
    int status = 0;
    const float ratio = fabsf(tf - t0) / fabsf(dt);
    int64_t nsteps = 0;
    if (ratio < 4294967296.0f)
        nsteps = (int64_t)floorf(ratio);
    else
        status = 2;
    int cur = 0;
    if (use_saveat) {
        if (nsaveat > 0 && t0 == saveat[0]) {
            if (nsave > 0) {
                for (int c = 0; c < n; ++c) us[(int64_t)c * N + i] = u[c];
                ts[i] = t0_fill;
            }
            cur = 1;
        }
    } else if (nsave > 0) {
        ts[i] = t0;
        for (int c = 0; c < n; ++c) us[(int64_t)c * N + i] = u[c];
    }
    double wacc[OS_MAXM] = {0};
    float t = t0;
    int64_t row = 1;
    for (int64_t j = 0; j < nsteps; ++j) {
        float uprev[OS_MAXN], w[OS_MAXM];
        memcpy(uprev, u, sizeof(float) * (size_t)n);
        step_normals(seed, id, (uint32_t)j, m, w);
        if (wsum)
            for (int c = 0; c < m; ++c) wacc[c] += (double)(sqdt * w[c]);
        if (method == 8) {
            float fu[OS_MAXN], gu[OS_MAXN * OS_MAXM];
            sde_f(sys, uprev, pp, t, fu);
            sde_g(sys, uprev, pp, t, gu);
            if (diag) {
                for (int c = 0; c < n; ++c) u[c] = (uprev[c] + fu[c] * dt) + (sqdt * gu[c]) * w[c];
            } else {
                for (int r = 0; r < n; ++r) {
                    float acc = (sqdt * gu[r * m]) * w[0];
                    for (int c = 1; c < m; ++c) acc = acc + (sqdt * gu[r * m + c]) * w[c];
                    u[r] = (uprev[r] + fu[r] * dt) + acc;
                }
            }
        } else {
            const float al1 = 0.5f, al2 = 0.5f, ga1 = 0.5f;
            const float la1 = 0.25f, la2 = -0.25f, la3 = 0.25f;
            const float mu1 = 0.25f, mu2 = 0.25f, mu3 = -0.25f;
            const float mu0 = 1.0f, mub0 = 1.0f, la0 = 1.0f, lab0 = 1.0f;
            const float nu1 = 1.0f, nu2 = 0.0f, be2 = 1.0f, be3 = 0.0f, de2 = -1.0f, de3 = 0.0f;
            float k0[OS_MAXN], g0[OS_MAXN], dW[OS_MAXN], W2[OS_MAXN], tmp[OS_MAXN], k1[OS_MAXN], g1[OS_MAXN],
                g2[OS_MAXN];
            sde_f(sys, uprev, pp, t, k0);
            sde_g(sys, uprev, pp, t, g0);
            for (int c = 0; c < n; ++c) {
                dW[c] = sqdt * w[c];
                W2[c] = (dW[c] * dW[c]) / sqdt;
                const float W3 = (nu2 * ((dW[c] * dW[c]) * dW[c])) / dt;
                tmp[c] = ((uprev[c] + (la0 * k0[c]) * dt) + (nu1 * g0[c]) * dW[c]) + g0[c] * W3;
            }
            sde_f(sys, tmp, pp, t + mu0 * dt, k1);
            for (int c = 0; c < n; ++c) tmp[c] = ((uprev[c] + (lab0 * k0[c]) * dt) + (be2 * g0[c]) * sqdt) + (be3 * g0[c]) * W2[c];
            sde_g(sys, tmp, pp, t + mub0 * dt, g1);
            for (int c = 0; c < n; ++c) tmp[c] = ((uprev[c] + (lab0 * k0[c]) * dt) + (de2 * g0[c]) * sqdt) + (de3 * g0[c]) * W2[c];
            sde_g(sys, tmp, pp, t + mub0 * dt, g2);
            for (int c = 0; c < n; ++c) {
                const float ud = uprev[c] + (al1 * k0[c] + al2 * k1[c]) * dt;
                const float A = (ga1 * g0[c]) * dW[c];
                const float B = (((la1 * dW[c]) + la2 * sqdt) + la3 * W2[c]) * g1[c];
                const float C = (((mu1 * dW[c]) + mu2 * sqdt) + mu3 * W2[c]) * g2[c];
                u[c] = ud + ((A + B) + C);
            }
        }
        t = t + dt;
        if (!use_saveat) {
            if (save_everystep) {
                if (row < nsave) {
                    for (int c = 0; c < n; ++c) us[(row * n + c) * N + i] = u[c];
                    ts[row * N + i] = t;
                }
                ++row;
            }
        } else {
            while (cur < nsaveat && saveat[cur] <= t) {
                const float savet = saveat[cur];
                const float th = (savet - (t - dt)) / dt;
                if (cur < nsave) {
                    for (int c = 0; c < n; ++c) us[((int64_t)cur * n + c) * N + i] = uprev[c] + (u[c] - uprev[c]) * th;
                    ts[(int64_t)cur * N + i] = savet;
                }
                ++cur;
            }
        }
    }
    if (use_saveat) {
        for (int64_t r = cur; r < nsave; ++r) ts[r * N + i] = t0_fill;
    } else if (save_everystep) {
        for (int64_t r = row; r < nsave; ++r) ts[r * N + i] = t0_fill;
    } else if (nsave > 1) {
        ts[N + i] = t;
        for (int c = 0; c < n; ++c) us[((int64_t)n + c) * N + i] = u[c];
    }
    if (status == 0) {
        int fin = isfinite(t);
        for (int c = 0; c < n; ++c) fin = fin && isfinite(u[c]);
        if (!fin) status = 3;
    }
    if (wsum)
        for (int c = 0; c < m; ++c) wsum[(int64_t)c * N + i] = wacc[c];
    if (naccept) naccept[i] = (int32_t)(nsteps < 0x7fffffffll ? nsteps : 0x7fffffffll);
    return status;
}

int deg_oracle_sde_solve(int system, int method, int64_t N, const float *u0, const float *p, const float *t0s,
                         const float *tfs, float t0, float tf, float dt, const uint64_t *seeds, uint64_t seed,
                         int64_t traj_base, const float *saveat, int nsaveat, int nsave, int save_everystep, float *ts,
                         float *us, int32_t *retcode, int32_t *naccept, double *wsum, int nthreads) {
    if (system < 0 || system >= OSDE_COUNT || (method != 8 && method != 9) || N < 0) return -1;
    int n, m, diag;
    sde_dims(system, &n, &m, &diag);
    if (method == 9 && !diag) return -2;
    if (!(dt > 0.0f)) return -3;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i) {
        const float a = t0s ? t0s[i] : t0, b = tfs ? tfs[i] : tf;
        const uint64_t sd = seeds ? seeds[i] : seed;
        const int rc = sde_traj(system, method, N, i, u0, p, a, b, t0, dt, sd, (uint64_t)(traj_base + i), saveat,
                                nsaveat, nsave, save_everystep, ts, us, wsum, naccept);
        if (retcode) retcode[i] = rc;
    }
    return 0;
}
