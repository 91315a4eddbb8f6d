/* deg_oracle.c -- CPU restatement of DiffEqGPU's EnsembleGPUKernel ODE path.
 *
 * TEST INFRASTRUCTURE. This library is the checker for the HIP solver: only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg load it. The product
 * never calls it and has no CPU fallback through it.
 *
 * It restates, operation for operation in IEEE Float32 (build with -ffp-contract=off):
 *   fixed-step loop and endpoint rules      src/ensemblegpukernel/kernels.jl:1-72
 *   adaptive loop and endpoint rules        src/ensemblegpukernel/kernels.jl:74-152
 *   Tsit5 fixed / adaptive step             perform_step/gpu_tsit5_perform_step.jl:1-64 / :68-175
 *   Vern7 fixed / adaptive step             perform_step/gpu_vern7_perform_step.jl:1-74 / :78-210
 *   Rosenbrock23 fixed / adaptive step      perform_step/gpu_rosenbrock23_perform_step.jl:1-70 / :74-201
 *   Vern9 fixed / adaptive step             perform_step/gpu_vern9_perform_step.jl:1-124 / :128-319
 *   Rodas4 / Rodas5P fixed / adaptive step  perform_step/gpu_rodas4_perform_step.jl, gpu_rodas5P_perform_step.jl
 *   Kvaerno3 / Kvaerno5 fixed / adaptive    perform_step/gpu_kvaerno3_perform_step.jl, gpu_kvaerno5_perform_step.jl,
 *   (Newton stage solver)                   nlsolve/utils.jl:1-27, nlsolve/type.jl:129-185
 *   PI controller constants                 integrators/integrator_utils.jl:1-11, alg_utils.jl:5-8
 *   saveat / save_everystep writes          integrators/integrator_utils.jl:13-50
 *   dense output                            integrators/nonstiff/interpolants.jl:28-190 (Vern7), :192-370
 *                                           (Vern9), :372-387
 *                                           (Tsit5), integrators/stiff/interpolants.jl (Rosenbrock23)
 *   error norm                              DiffEqBase.ODE_DEFAULT_NORM = sqrt(sum(abs2, x) / n)
 *   tstops                                  perform_step/gpu_tsit5_perform_step.jl:18-27, :158-164
 *   callbacks: handle_callbacks!, find_callback_time, ITP root finder, apply_callback!,
 *   change_t_via_interpolation!, apply_discrete_callback!, terminate!
 *                                           integrators/integrator_utils.jl:52-479
 *   earliest continuous event               DiffEqBase.find_first_continuous_callback (external
 *                                           dependency DiffEqBase 7.2: the earliest event time
 *                                           wins, the first callback on ties)
 * with these stated deviations, shared with the GPU solver:
 *   - a lane that would hit the reference's in-kernel error("dt<dtmin") stops with a
 *     per-trajectory return code instead of aborting the launch;
 *   - a per-trajectory attempt cap (max_attempts) ends a lane that would loop forever;
 *   - an accepted adaptive step that cannot advance t (t + dt rounds to t) and whose next dt
 *     is no larger snaps t to tf when tf is within two ulps of t and otherwise ends the lane
 *     with DtLessThanMin (the reference repeats such a step forever); when the next dt is
 *     larger the lane goes on with t unchanged, as the reference does. A fixed step that
 *     cannot advance t ends the lane with DtLessThanMin;
 *   - Rosenbrock23 forms W = I - gamma J, Rodas4/Rodas5P W = J - I / (dt gamma), Kvaerno3/5 (Newton)
 *     W = -I + gamma dt J, and they solve it in FP64 (LU with partial pivoting) instead of Float32
 *     Cramer's rule (linalg/linsolve.jl:31-57);
 *   - optional compensated (Kahan) accumulation of t (flag DEG_O_KAHAN).
 * Dense-output polynomials use fused multiply-add Horner steps (the reference's
 * @muladd/@evalpoly), vector combinations are left folds without fusion.
 *
 * An FP64 Vern9 integrator (tableaus/verner_tableaus.jl:1198-1332, same PI controller
 * in double) provides the high-accuracy reference solutions.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "tableaus.h"
#include "systems.h"

#define OMAXN 24
#define OMAXS 17

enum { OM_TSIT5 = 0, OM_VERN7 = 1, OM_RB23 = 2, OM_VERN9 = 3, OM_RODAS4 = 4, OM_RODAS5P = 5, OM_KVAERNO3 = 6,
       OM_KVAERNO5 = 7 };
/* DEG_O_AUGMENTED: a dual-number system solved as a plain system of its n (1 + P) entries (the full
 * augmented Jacobian and W, the plain norm); a pin of the dual W factor, not a reference behaviour.
 * DEG_O_PIVLU: W by LU with partial pivoting also for n <= 3; a pin of the Cramer solve, likewise */
enum { DEG_O_KAHAN = 1, DEG_O_DETPOW = 2, DEG_O_FDJAC = 4, DEG_O_AUGMENTED = 8, DEG_O_PIVLU = 16 };

/* autodiff = false: J by forward differences, column i = (f(u + dx e_i) - f(u)) / dx with the
 * perturbed entry rounded to Float32 (finite_diff_jac, alg_utils.jl:17-27), and the time gradient
 * (f(u, p, t + dx) - f(u, p, t)) / dx (nlsolve/type.jl:150-154); dx = sqrt(eps(Float32)) = 2^-11.5
 * rounded to Float32. */
static float fd_step(void) {
    const uint32_t b = 0x39B504F3u;
    float x;
    memcpy(&x, &b, sizeof x);
    return x;
}

/* Over a dual-number state (S->dualp > 0) J has the columns of the dualnv values only,
 * J[r * dualnv + c] = d f_r / d value_c for all n rows: rows 0..dualnv-1 are the Jacobian's values and
 * row dualnv + r P + j its partial j (the forward difference of the dual right-hand side is taken in
 * dual arithmetic, finite_diff_jac on a dual u: the step dx perturbs the value alone). */
static int jac_cols(const osys *S) { return S->dualp > 0 ? S->dualnv : S->n; }

static void sys_jac(const osys *S, const float *u, const float *p, float t, float *J) {
    if (!S->fd) {
        S->jac(u, p, t, J);
        return;
    }
    const int n = S->n, m = jac_cols(S);
    const float dx = fd_step();
    float f0[OMAXN], f1[OMAXN], ux[OMAXN];
    S->f(u, p, t, f0);
    for (int i = 0; i < m; ++i) {
        for (int r = 0; r < n; ++r) ux[r] = u[r];
        ux[i] = u[i] + dx;
        S->f(ux, p, t, f1);
        for (int r = 0; r < n; ++r) J[r * m + i] = (f1[r] - f0[r]) / dx;
    }
}

static void sys_tgrad(const osys *S, const float *u, const float *p, float t, float *dT) {
    if (!S->fd) {
        S->tgrad(u, p, t, dT);
        return;
    }
    const int n = S->n;
    const float dx = fd_step();
    float f0[OMAXN], f1[OMAXN];
    S->f(u, p, t, f0);
    S->f(u, p, t + dx, f1);
    for (int r = 0; r < n; ++r) dT[r] = (f1[r] - f0[r]) / dx;
}
/* per-trajectory return codes (same values as the C ABI of the HIP library) */
enum { ORET_SUCCESS = 0, ORET_DTMIN = 1, ORET_MAXITERS = 2, ORET_NONFINITE = 3, ORET_TERMINATED = 4,
       ORET_FAILURE = 5 };

/* ------------------------------------------------------------------------------------ */
/* Float32 tableau in the shape every explicit method here uses                         */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int s;
    int fsal;                 /* Tsit5: the last stage is evaluated at the new solution */
    float c[OMAXS + 1];
    float a[OMAXS + 1][OMAXS + 1];
    unsigned char am[OMAXS + 1][OMAXS + 1];
    float b[OMAXS + 1];
    unsigned char bm[OMAXS + 1];
    float bt[OMAXS + 1];
    unsigned char btm[OMAXS + 1];
} erk32;

static void erk32_load(erk32 *T, int s, int fsal, const double *c, const double *a, const unsigned char *am,
                       const double *b, const unsigned char *bm, const double *bt, const unsigned char *btm) {
    memset(T, 0, sizeof(*T));
    T->s = s;
    T->fsal = fsal;
    for (int i = 0; i <= s; ++i) {
        T->c[i] = (float)c[i];
        T->b[i] = (float)b[i];
        T->bm[i] = bm[i];
        T->bt[i] = (float)bt[i];
        T->btm[i] = btm[i];
        for (int j = 0; j <= s; ++j) {
            T->a[i][j] = (float)a[i * (s + 1) + j];
            T->am[i][j] = am[i * (s + 1) + j];
        }
    }
}

static erk32 TAB_T5, TAB_V7, TAB_V9;
static void init_v7x(void);
static void init_v9x(void);
static float T5R[8][5];
static int tabs_ready = 0;

static void init_tabs(void) {
    if (tabs_ready) return;
    erk32_load(&TAB_T5, OT5_S, 1, OT5_C, &OT5_A[0][0], &OT5_AM[0][0], OT5_B, OT5_BM, OT5_BT, OT5_BTM);
    erk32_load(&TAB_V7, OV7_S, 0, OV7_C, &OV7_A[0][0], &OV7_AM[0][0], OV7_B, OV7_BM, OV7_BT, OV7_BTM);
    erk32_load(&TAB_V9, OV9_S, 0, OV9_C, &OV9_A[0][0], &OV9_AM[0][0], OV9_B, OV9_BM, OV9_BT, OV9_BTM);
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 5; ++k) T5R[i][k] = (float)OT5_R[i][k];
    init_v7x();
    init_v9x();
    tabs_ready = 1;
}

/* ------------------------------------------------------------------------------------ */
/* Julia Float32 semantics helpers                                                       */
/* ------------------------------------------------------------------------------------ */
/* Julia's min/max propagate NaN (Base.min for IEEE floats). */
static inline float jl_min(float a, float b) {
    if (a != a || b != b) return a + b;
    return (b < a) ? b : a;
}
static inline float jl_max(float a, float b) {
    if (a != a || b != b) return a + b;
    return (b > a) ? b : a;
}

/* Deterministic power x^y for x >= 0 built only from IEEE +,-,*,/ and exponent bit
 * manipulation: log2 by the atanh series on a mantissa in [sqrt(1/2), sqrt(2)], exp2 by a
 * degree-7 Taylor polynomial on [-1/2, 1/2] ln 2. Relative error about 1e-6, which is
 * irrelevant for a step-size controller. The HIP solver carries a bit-identical copy so
 * that adaptive runs can be compared bit for bit (flag DEG_O_DETPOW); the default path
 * uses the platform powf. */
static inline float det_exp2i(int k) {
    uint32_t b = (uint32_t)(k + 127) << 23;
    float r;
    memcpy(&r, &b, 4);
    return r;
}
static float det_pow(float x, float y) {
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return 0.0f;
    if (x == INFINITY) return INFINITY;
    uint32_t b;
    memcpy(&b, &x, 4);
    int e = (int)((b >> 23) & 0xffu);
    if (e == 0) {
        x = x * 16777216.0f;
        memcpy(&b, &x, 4);
        e = (int)((b >> 23) & 0xffu) - 24;
    }
    e -= 127;
    uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &mb, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float lnm = (2.0f * s) * (1.0f + s2 * (0.333333343f + s2 * (0.2f + s2 * (0.142857149f + s2 * 0.111111112f))));
    float l2 = (float)e + lnm * 1.44269502f;
    float z = y * l2;
    if (z > 128.0f) return INFINITY;
    if (z < -149.0f) return 0.0f;
    float n = rintf(z);
    float g = (z - n) * 0.693147182f;
    float r = 1.0f + g * (1.0f + g * (0.5f + g * (0.166666672f + g * (0.0416666679f + g * (0.00833333377f + g * (0.00138888892f + g * 0.000198412701f))))));
    int ni = (int)n;
    int h1 = ni / 2;
    int h2 = ni - h1;
    return (r * det_exp2i(h1)) * det_exp2i(h2);
}

static inline float ctl_pow(float x, float y, int detpow) { return detpow ? det_pow(x, y) : powf(x, y); }

/* ODE_DEFAULT_NORM for an SVector: sqrt(((x1^2 + x2^2) + ...) / n) */
static inline float rms_norm(const float *x, int n) {
    float acc = x[0] * x[0];
    for (int c = 1; c < n; ++c) acc = acc + x[c] * x[c];
    return sqrtf(acc / (float)n);
}

/* ------------------------------------------------------------------------------------ */
/* Explicit RK stage combinations in the reference's association order                  */
/* ------------------------------------------------------------------------------------ */
/* stage i input: uprev + dt * (sum_j a_ij k_j), except stage 2: uprev + (dt*a21) * k1 */
static void stage_input(const erk32 *T, int i, int n, const float *uprev, float dt,
                        float k[][OMAXN], float *out) {
    if (i == 2) {
        float a = dt * T->a[2][1];
        for (int c = 0; c < n; ++c) out[c] = uprev[c] + a * k[1][c];
        return;
    }
    for (int c = 0; c < n; ++c) {
        float acc = 0.0f;
        int first = 1;
        for (int j = 1; j < i; ++j) {
            if (!T->am[i][j]) continue;
            float term = T->a[i][j] * k[j][c];
            acc = first ? term : acc + term;
            first = 0;
        }
        out[c] = uprev[c] + dt * acc;
    }
}

static void weighted_sum(const float *w, const unsigned char *wm, int s, int n, float dt,
                         float k[][OMAXN], float *out) {
    for (int c = 0; c < n; ++c) {
        float acc = 0.0f;
        int first = 1;
        for (int j = 1; j <= s; ++j) {
            if (!wm[j]) continue;
            float term = w[j] * k[j][c];
            acc = first ? term : acc + term;
            first = 0;
        }
        out[c] = dt * acc;
    }
}

/* One explicit RK attempt from (uprev, t) with step dt. k[1] must hold f(uprev, t).
 * Fills k[2..s], unew. */
static void erk_attempt(const erk32 *T, const osys *S, const float *p, const float *uprev, float t,
                        float dt, float k[][OMAXN], float *unew) {
    int n = S->n;
    float tmp[OMAXN];
    int s = T->s;
    for (int i = 2; i <= s; ++i) {
        stage_input(T, i, n, uprev, dt, k, tmp);
        if (T->fsal && i == s) {
            for (int c = 0; c < n; ++c) unew[c] = tmp[c];
        }
        S->f(tmp, p, t + T->c[i] * dt, k[i]);
    }
    if (!T->fsal) {
        float inc[OMAXN];
        weighted_sum(T->b, T->bm, s, n, dt, k, inc);
        for (int c = 0; c < n; ++c) unew[c] = uprev[c] + inc[c];
    }
}

/* The same error norm over a dual-number state (S->dualnv values with S->dualp partials each,
 * component i's partials at dualnv + i dualp + j): the scaling e ./ (abstol .+ max.(|uprev|,|u|) *
 * reltol) in dual arithmetic with the derivative rules of DiffRules (a ForwardDiff 1.3.3 dependency,
 * not vendored in the reference): abs negates the partials where signbit(value) (so -0.0 negates too),
 * max(x, y) takes x's partials only when x > y and y's otherwise (ties and NaN take y = |u|); e / s
 * has partials e' (1 / s) + s' (-(e / s^2)) (ForwardDiff's _div_partials); then the root mean square
 * over all values and partials, summed as value^2 + (partial_1^2 + ... + partial_P^2) per component. */
static float dual_eest(int nv, int P, int n, const float *e, const float *uprev, const float *u, float abstol,
                       float reltol) {
    float acc = 0.0f;
    for (int i = 0; i < nv; ++i) {
        const float av = uprev[i], bv = u[i];
        const float aa = fabsf(av), ba = fabsf(bv);
        const int nan = (aa != aa) || (ba != ba);
        const int pick_u = !(aa > ba);
        const float mv = nan ? aa + ba : (pick_u ? ba : aa);
        const int neg = pick_u ? (signbit(bv) != 0) : (signbit(av) != 0);
        const float sv = abstol + mv * reltol;
        const float ev = e[i];
        const float qv = ev / sv;
        const float inv = 1.0f / sv;
        const float c = -(ev / (sv * sv));
        float ps = 0.0f;
        for (int j = 0; j < P; ++j) {
            const int q = nv + i * P + j;
            const float md = pick_u ? u[q] : uprev[q];
            const float sd = (neg ? -md : md) * reltol;
            const float qd = e[q] * inv + sd * c;
            ps = (j == 0) ? qd * qd : ps + qd * qd;
        }
        const float term = qv * qv + ps;
        acc = (i == 0) ? term : acc + term;
    }
    return sqrtf(acc / (float)n);
}

/* EEst = norm(e ./ (abstol .+ max.(|uprev|,|u|) * reltol)), the dual norm over a dual state (e is scaled
 * in place on the plain path) */
static float scaled_eest(const osys *S, float *e, const float *uprev, const float *u, float abstol, float reltol) {
    const int n = S->n;
    if (S->dualp > 0) return dual_eest(S->dualnv, S->dualp, n, e, uprev, u, abstol, reltol);
    for (int c = 0; c < n; ++c) {
        float sc = abstol + jl_max(fabsf(uprev[c]), fabsf(u[c])) * reltol;
        e[c] = e[c] / sc;
    }
    return rms_norm(e, n);
}

/* scaled error  EEst = norm( dt*(sum btilde_j k_j) ./ (abstol .+ max.(|uprev|,|u|) * reltol) ) */
static float erk_eest(const erk32 *T, int n, float dt, float k[][OMAXN], const float *uprev,
                      const float *u, float abstol, float reltol, int dualnv, int dualp) {
    float e[OMAXN];
    weighted_sum(T->bt, T->btm, T->s, n, dt, k, e);
    if (dualp > 0) return dual_eest(dualnv, dualp, n, e, uprev, u, abstol, reltol);
    for (int c = 0; c < n; ++c) {
        float sc = abstol + jl_max(fabsf(uprev[c]), fabsf(u[c])) * reltol;
        e[c] = e[c] / sc;
    }
    return rms_norm(e, n);
}

/* Tsit5 dense output: y0 + dt*(b1(th) k1 + ... + b7(th) k7) */
static void tsit5_interp(float th, float dt, int n, const float *y0, float k[][OMAXN], float *out) {
    float bth[8];
    bth[1] = th * fmaf(th, fmaf(th, fmaf(th, T5R[1][4], T5R[1][3]), T5R[1][2]), T5R[1][1]);
    for (int i = 2; i <= 7; ++i) bth[i] = th * (th * fmaf(th, fmaf(th, T5R[i][4], T5R[i][3]), T5R[i][2]));
    for (int c = 0; c < n; ++c) {
        float acc = bth[1] * k[1][c];
        for (int i = 2; i <= 7; ++i) acc = acc + bth[i] * k[i][c];
        out[c] = y0[c] + dt * acc;
    }
}

/* Rosenbrock23 dense output: y0 + dt*(c1 k1 + c2 k2), c1 = th(1-th)/(1-2d), c2 = th(th-2d)/(1-2d) */
static void rb23_interp(float th, float dt, int n, float d, const float *y0, const float *k1,
                        const float *k2, float *out) {
    float den = fmaf(-2.0f, d, 1.0f);
    float c1 = (th * (1.0f - th)) / den;
    float c2 = (th * fmaf(-2.0f, d, th)) / den;
    for (int c = 0; c < n; ++c) out[c] = y0[c] + dt * (c1 * k1[c] + c2 * k2[c]);
}

/* Vern7 dense output (interpolants.jl:28-190): six extra stages k11..k16 from uprev, then
 * y0 + th*step_sum + th*(th-1)*correction with q_i(th) from the synthetic division of the
 * reference's r coefficients. Scalar muladds (stage times, polynomial Horner steps) are fused;
 * vector combinations are left folds without fusion. */
static float V7X_A[17][17], V7X_C[17], V7R[17][8];
static unsigned char V7X_AM[17][17];
static const int V7_QSTAGES[13] = {1, 4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15, 16};

static void init_v7x(void) {
    for (int i = 0; i < 17; ++i) {
        V7X_C[i] = (float)OV7X_C[i];
        for (int j = 0; j < 17; ++j) {
            V7X_A[i][j] = (float)OV7X_A[i][j];
            V7X_AM[i][j] = OV7X_AM[i][j];
        }
        for (int k = 0; k < 8; ++k) V7R[i][k] = (float)OV7R[i][k];
    }
}

static float vern7_q(int i, float th) {
    const float *r = V7R[i];
    float s5 = r[7];
    float s4 = r[6] + s5;
    float s3 = r[5] + s4;
    float s2 = r[4] + s3;
    float s1 = r[3] + s2;
    float c0 = r[2] + s1;
    return fmaf(th, fmaf(th, fmaf(th, fmaf(th, fmaf(th, s5, s4), s3), s2), s1), c0);
}

static void vern7_interp(const osys *S, const float *p, float th, float dt, float tprev, int n, const float *y0,
                         float k[][OMAXN], float *out) {
    float kx[17][OMAXN];
    for (int s = 1; s <= 10; ++s)
        for (int c = 0; c < n; ++c) kx[s][c] = k[s][c];
    float tmp[OMAXN];
    for (int i = 11; i <= 16; ++i) {
        for (int c = 0; c < n; ++c) {
            float acc = 0.0f;
            int first = 1;
            for (int j = 1; j < i; ++j) {
                if (!V7X_AM[i][j]) continue;
                float term = V7X_A[i][j] * kx[j][c];
                acc = first ? term : acc + term;
                first = 0;
            }
            tmp[c] = y0[c] + dt * acc;
        }
        S->f(tmp, p, fmaf(V7X_C[i], dt, tprev), kx[i]);
    }
    float q[17];
    for (int m = 0; m < 13; ++m) q[V7_QSTAGES[m]] = vern7_q(V7_QSTAGES[m], th);
    const erk32 *T = &TAB_V7;
    for (int c = 0; c < n; ++c) {
        float ss = T->b[1] * kx[1][c];
        const int bi[6] = {4, 5, 6, 7, 8, 9};
        for (int m = 0; m < 6; ++m) ss = ss + T->b[bi[m]] * kx[bi[m]][c];
        float step_sum = dt * ss;
        float cs = kx[1][c] * q[1];
        for (int m = 1; m < 13; ++m) cs = cs + kx[V7_QSTAGES[m]][c] * q[V7_QSTAGES[m]];
        float corr = dt * cs;
        out[c] = (y0[c] + th * step_sum) + (th * (th - 1.0f)) * corr;
    }
}

/* Vern9 dense output (interpolants.jl:192-370): extra stages 17..26 from uprev and the step's
 * stages 1, 8..15, then y0 + dt * (sum of b_i(th) k_i) with b_1 = th*H(r_11..r_19) and
 * b_i = th*(th*H(r_i2..r_i9)), H a fused Horner evaluation (@evalpoly). Stage times are fused
 * muladds; vector sums are unfused left folds. */
static float V9X_A[27][27], V9X_C[27], V9R[27][10];
static unsigned char V9X_AM[27][27];
static const int V9_BSTAGES[19] = {1, 8, 9, 10, 11, 12, 13, 14, 15, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26};

static void init_v9x(void) {
    for (int i = 0; i < 27; ++i) {
        V9X_C[i] = (float)OV9X_C[i];
        for (int j = 0; j < 27; ++j) {
            V9X_A[i][j] = (float)OV9X_A[i][j];
            V9X_AM[i][j] = OV9X_AM[i][j];
        }
        for (int k = 0; k < 10; ++k) V9R[i][k] = (float)OV9R[i][k];
    }
}

static float vern9_b(int i, float th) {
    const float *r = V9R[i];
    float h = fmaf(th, r[9], r[8]);
    for (int k = 7; k >= 2; --k) h = fmaf(th, h, r[k]);
    if (i == 1) return th * fmaf(th, h, r[1]);
    return th * (th * h);
}

static void vern9_interp(const osys *S, const float *p, float th, float dt, float tprev, int n, const float *y0,
                         float k[][OMAXN], float *out) {
    float kx[27][OMAXN];
    for (int s = 1; s <= 16; ++s)
        for (int c = 0; c < n; ++c) kx[s][c] = k[s][c];
    float tmp[OMAXN];
    for (int i = 17; i <= 26; ++i) {
        for (int c = 0; c < n; ++c) {
            float acc = 0.0f;
            int first = 1;
            for (int j = 1; j < i; ++j) {
                if (!V9X_AM[i][j]) continue;
                float term = V9X_A[i][j] * kx[j][c];
                acc = first ? term : acc + term;
                first = 0;
            }
            tmp[c] = y0[c] + dt * acc;
        }
        S->f(tmp, p, fmaf(V9X_C[i], dt, tprev), kx[i]);
    }
    float b[27];
    for (int m = 0; m < 19; ++m) b[V9_BSTAGES[m]] = vern9_b(V9_BSTAGES[m], th);
    for (int c = 0; c < n; ++c) {
        float acc = kx[1][c] * b[1];
        for (int m = 1; m < 19; ++m) acc = acc + kx[V9_BSTAGES[m]][c] * b[V9_BSTAGES[m]];
        out[c] = y0[c] + dt * acc;
    }
}

/* ------------------------------------------------------------------------------------ */
/* FP64 LU with partial pivoting for W = I - gamma*J (Rosenbrock23)                     */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    double lu[OMAXN][OMAXN];  /* LU factors, or for n <= 3 the cofactor rows of Cramer's rule */
    int perm[OMAXN];
    int cramer;               /* n <= 3: x_i = (sum_j lu[i][j] b_j) / det */
    double det;
    /* dual-number state (P > 0): lu factors the nv x nv values of W = alpha M + beta J, and the
     * partials of W are beta times rows nv.. of the dual Jacobian, kept in jp */
    int nv, P;
    double beta;
    float jp[OMAXN][OMAXN];
} olu;

/* Record the dual part of a factor built from J (the n x nv dual Jacobian of sys_jac) */
static void w_dual(olu *F, const osys *S, double beta, const float *J) {
    F->nv = S->dualp > 0 ? S->dualnv : 0;
    F->P = S->dualp;
    F->beta = beta;
    if (F->P == 0) return;
    for (int r = F->nv; r < S->n; ++r)
        for (int c = 0; c < F->nv; ++c) F->jp[r][c] = J[r * F->nv + c];
}

/* Cramer's rule for n <= 3, the reference's own solve for these sizes (linalg/linsolve.jl:8-57), in FP64:
 * the cofactor rows of its formulas and StaticArrays' det (3 x 3 expanded along the first row) */
static void w_cramer(olu *F, int n) {
    double (*W)[OMAXN] = F->lu;
    F->cramer = 1;
    if (n == 1) {
        F->det = W[0][0];
        W[0][0] = 1.0;
    } else if (n == 2) {
        const double a00 = W[0][0], a01 = W[0][1], a10 = W[1][0], a11 = W[1][1];
        F->det = a00 * a11 - a01 * a10;
        W[0][0] = a11;
        W[0][1] = -a01;
        W[1][0] = -a10;
        W[1][1] = a00;
    } else {
        const double a00 = W[0][0], a01 = W[0][1], a02 = W[0][2];
        const double a10 = W[1][0], a11 = W[1][1], a12 = W[1][2];
        const double a20 = W[2][0], a21 = W[2][1], a22 = W[2][2];
        F->det = (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20)) + a02 * (a10 * a21 - a11 * a20);
        W[0][0] = a11 * a22 - a12 * a21;
        W[0][1] = a02 * a21 - a01 * a22;
        W[0][2] = a01 * a12 - a02 * a11;
        W[1][0] = a12 * a20 - a10 * a22;
        W[1][1] = a00 * a22 - a02 * a20;
        W[1][2] = a02 * a10 - a00 * a12;
        W[2][0] = a10 * a21 - a11 * a20;
        W[2][1] = a01 * a20 - a00 * a21;
        W[2][2] = a00 * a11 - a01 * a10;
    }
}

static void w_lu(olu *F, int n, int pivlu) {
    if (n <= 3 && !pivlu) {
        w_cramer(F, n);
        return;
    }
    F->cramer = 0;
    for (int r = 0; r < n; ++r) F->perm[r] = r;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double mx = fabs(F->lu[k][k]);
        for (int i = k + 1; i < n; ++i) {
            double v = fabs(F->lu[i][k]);
            if (v > mx) {
                mx = v;
                piv = i;
            }
        }
        if (piv != k) {
            for (int c = 0; c < n; ++c) {
                double tv = F->lu[k][c];
                F->lu[k][c] = F->lu[piv][c];
                F->lu[piv][c] = tv;
            }
            int tp = F->perm[k];
            F->perm[k] = F->perm[piv];
            F->perm[piv] = tp;
        }
        for (int i = k + 1; i < n; ++i) {
            double l = F->lu[i][k] / F->lu[k][k];
            F->lu[i][k] = l;
            for (int c = k + 1; c < n; ++c) F->lu[i][c] = F->lu[i][c] - l * F->lu[k][c];
        }
    }
}

/* M v for a mass matrix (row sums left to right, unfused), the identity when M is NULL */
static void mass_mul(const float *M, int n, const float *v, float *out) {
    for (int r = 0; r < n; ++r) {
        if (!M) {
            out[r] = v[r];
            continue;
        }
        float acc = M[r * n] * v[0];
        for (int c = 1; c < n; ++c) acc = acc + M[r * n + c] * v[c];
        out[r] = acc;
    }
}

/* Rosenbrock23: W = M - gamma J (gpu_rosenbrock23_perform_step.jl:44-45), M = I without a mass matrix */
static void w_factor(olu *F, const osys *S, float gam, const float *J, const float *M) {
    const int m = jac_cols(S);
    double g = (double)gam;
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c)
            F->lu[r][c] = (M ? (double)M[r * m + c] : (r == c ? 1.0 : 0.0)) - g * (double)J[r * m + c];
    w_lu(F, m, S->pivlu);
    w_dual(F, S, -g, J);
}

/* Rodas: W = J - M inv(dt gamma) (gpu_rodas4_perform_step.jl:66-67), the reciprocal and its product
 * with M taken in Float32 */
static void w_factor_rodas(olu *F, const osys *S, float invg, const float *J, const float *M) {
    const int m = jac_cols(S);
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c)
            F->lu[r][c] = (double)J[r * m + c] - (M ? (double)(M[r * m + c] * invg) : (r == c ? (double)invg : 0.0));
    w_lu(F, m, S->pivlu);
    w_dual(F, S, 1.0, J);
}

/* y = LU \ b in FP64 (b in FP64, overwritten) */
static void lu_solve_d(const olu *F, int n, double *b) {
    double y[OMAXN];
    if (F->cramer) {
        for (int i = 0; i < n; ++i) {
            double v = F->lu[i][0] * b[0];
            for (int j = 1; j < n; ++j) v = v + F->lu[i][j] * b[j];
            y[i] = v / F->det;
        }
        for (int i = 0; i < n; ++i) b[i] = y[i];
        return;
    }
    for (int i = 0; i < n; ++i) {
        double v = b[F->perm[i]];
        for (int j = 0; j < i; ++j) v = v - F->lu[i][j] * y[j];
        y[i] = v;
    }
    for (int i = n - 1; i >= 0; --i) {
        double v = y[i];
        for (int j = i + 1; j < n; ++j) v = v - F->lu[i][j] * y[j];
        y[i] = v / F->lu[i][i];
    }
    for (int i = 0; i < n; ++i) b[i] = y[i];
}

/* x = W \ b. Over a dual-number state the solve is the dual one, value LU plus one solve per
 * partial: x_v = W_v \ b_v and x_j = W_v \ (b_j - W_j x_v), with W_j = beta J_j and x_v kept in FP64
 * for the product (the derivative of W^-1 b: W^-1 (b' - W' W^-1 b)). */
static void w_solve(const olu *F, int n, const float *b, float *x) {
    const int m = F->P > 0 ? F->nv : n;
    double y[OMAXN];
    for (int i = 0; i < m; ++i) y[i] = (double)b[i];
    lu_solve_d(F, m, y);
    for (int i = 0; i < m; ++i) x[i] = (float)y[i];
    for (int j = 0; j < F->P; ++j) {
        double z[OMAXN];
        for (int r = 0; r < m; ++r) {
            const int q = m + r * F->P + j;
            double v = (double)b[q];
            for (int c = 0; c < m; ++c) v = v - (F->beta * (double)F->jp[q][c]) * y[c];
            z[r] = v;
        }
        lu_solve_d(F, m, z);
        for (int r = 0; r < m; ++r) x[m + r * F->P + j] = (float)z[r];
    }
}

/* ------------------------------------------------------------------------------------ */
/* Rodas4 / Rodas5P (gpu_rodas4_perform_step.jl:1-277, gpu_rodas5P_perform_step.jl:1-351)    */
/* ------------------------------------------------------------------------------------ */
/* One step from (uprev, t) with step dt: unew, the dense-output combinations k[1..2] (Rodas4)
 * or k[1..3] (Rodas5P), and the last stage (the error estimate) in kerr. Vector sums keep the
 * reference's association, including its term order in stages 5 and 6. */
#define R4(x) ((float)ORO4_##x)
#define R5(x) ((float)ORO5_##x)
static void rodas4_step(const osys *S, const float *p, const float *uprev, float t, float dt, float k[][OMAXN],
                        float *unew, float *kerr) {
    const int n = S->n;
    float Jm[OMAXN * OMAXN], dT[OMAXN], du[OMAXN], lt[OMAXN] = {0}, u[OMAXN], cm[OMAXN] = {0}, mc[OMAXN] = {0};
    float K1[OMAXN], K2[OMAXN], K3[OMAXN], K4[OMAXN], K5[OMAXN], K6[OMAXN];
    sys_jac(S, uprev, p, t, Jm);
    sys_tgrad(S, uprev, p, t, dT);
    const float dtC21 = R4(C21) / dt, dtC31 = R4(C31) / dt, dtC32 = R4(C32) / dt;
    const float dtC41 = R4(C41) / dt, dtC42 = R4(C42) / dt, dtC43 = R4(C43) / dt;
    const float dtC51 = R4(C51) / dt, dtC52 = R4(C52) / dt, dtC53 = R4(C53) / dt, dtC54 = R4(C54) / dt;
    const float dtC61 = R4(C61) / dt, dtC62 = R4(C62) / dt, dtC63 = R4(C63) / dt, dtC64 = R4(C64) / dt;
    const float dtC65 = R4(C65) / dt;
    const float dtd1 = dt * R4(d1), dtd2 = dt * R4(d2), dtd3 = dt * R4(d3), dtd4 = dt * R4(d4);
    const float dtgamma = dt * R4(gamma);
    olu F;
    w_factor_rodas(&F, S, 1.0f / dtgamma, Jm, S->M);
    S->f(uprev, p, t, du);
    for (int c = 0; c < n; ++c) lt[c] = -(du[c] + dtd1 * dT[c]);
    w_solve(&F, n, lt, K1);
    for (int c = 0; c < n; ++c) u[c] = uprev[c] + R4(a21) * K1[c];
    S->f(u, p, t + R4(c2) * dt, du);
    for (int c = 0; c < n; ++c) cm[c] = dtC21 * K1[c];
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -((du[c] + dtd2 * dT[c]) + mc[c]);
    w_solve(&F, n, lt, K2);
    for (int c = 0; c < n; ++c) u[c] = (uprev[c] + R4(a31) * K1[c]) + R4(a32) * K2[c];
    S->f(u, p, t + R4(c3) * dt, du);
    for (int c = 0; c < n; ++c) cm[c] = (dtC31 * K1[c] + dtC32 * K2[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -((du[c] + dtd3 * dT[c]) + mc[c]);
    w_solve(&F, n, lt, K3);
    for (int c = 0; c < n; ++c) u[c] = ((uprev[c] + R4(a41) * K1[c]) + R4(a42) * K2[c]) + R4(a43) * K3[c];
    S->f(u, p, t + R4(c4) * dt, du);
    for (int c = 0; c < n; ++c) cm[c] = ((dtC41 * K1[c] + dtC42 * K2[c]) + dtC43 * K3[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -((du[c] + dtd4 * dT[c]) + mc[c]);
    w_solve(&F, n, lt, K4);
    for (int c = 0; c < n; ++c)
        u[c] = (((uprev[c] + R4(a51) * K1[c]) + R4(a52) * K2[c]) + R4(a53) * K3[c]) + R4(a54) * K4[c];
    S->f(u, p, t + dt, du);
    for (int c = 0; c < n; ++c) cm[c] = (((dtC52 * K2[c] + dtC54 * K4[c]) + dtC51 * K1[c]) + dtC53 * K3[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -(du[c] + mc[c]);
    w_solve(&F, n, lt, K5);
    for (int c = 0; c < n; ++c) u[c] = u[c] + K5[c];
    S->f(u, p, t + dt, du);
    for (int c = 0; c < n; ++c) cm[c] = ((((dtC61 * K1[c] + dtC62 * K2[c]) + dtC65 * K5[c]) + dtC64 * K4[c]) + dtC63 * K3[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -(du[c] + mc[c]);
    w_solve(&F, n, lt, K6);
    for (int c = 0; c < n; ++c) {
        unew[c] = u[c] + K6[c];
        kerr[c] = K6[c];
        k[1][c] = (((R4(h21) * K1[c] + R4(h22) * K2[c]) + R4(h23) * K3[c]) + R4(h24) * K4[c]) + R4(h25) * K5[c];
        k[2][c] = (((R4(h31) * K1[c] + R4(h32) * K2[c]) + R4(h33) * K3[c]) + R4(h34) * K4[c]) + R4(h35) * K5[c];
    }
}

static void rodas5p_step(const osys *S, const float *p, const float *uprev, float t, float dt, float k[][OMAXN],
                         float *unew, float *kerr) {
    const int n = S->n;
    float Jm[OMAXN * OMAXN], dT[OMAXN], du[OMAXN], lt[OMAXN] = {0}, u[OMAXN], cm[OMAXN] = {0}, mc[OMAXN] = {0};
    float K1[OMAXN], K2[OMAXN], K3[OMAXN], K4[OMAXN], K5[OMAXN], K6[OMAXN], K7[OMAXN], K8[OMAXN];
    sys_jac(S, uprev, p, t, Jm);
    sys_tgrad(S, uprev, p, t, dT);
    const float dtC21 = R5(C21) / dt, dtC31 = R5(C31) / dt, dtC32 = R5(C32) / dt;
    const float dtC41 = R5(C41) / dt, dtC42 = R5(C42) / dt, dtC43 = R5(C43) / dt;
    const float dtC51 = R5(C51) / dt, dtC52 = R5(C52) / dt, dtC53 = R5(C53) / dt, dtC54 = R5(C54) / dt;
    const float dtC61 = R5(C61) / dt, dtC62 = R5(C62) / dt, dtC63 = R5(C63) / dt, dtC64 = R5(C64) / dt;
    const float dtC65 = R5(C65) / dt;
    const float dtC71 = R5(C71) / dt, dtC72 = R5(C72) / dt, dtC73 = R5(C73) / dt, dtC74 = R5(C74) / dt;
    const float dtC75 = R5(C75) / dt, dtC76 = R5(C76) / dt;
    const float dtC81 = R5(C81) / dt, dtC82 = R5(C82) / dt, dtC83 = R5(C83) / dt, dtC84 = R5(C84) / dt;
    const float dtC85 = R5(C85) / dt, dtC86 = R5(C86) / dt, dtC87 = R5(C87) / dt;
    const float dtd1 = dt * R5(d1), dtd2 = dt * R5(d2), dtd3 = dt * R5(d3), dtd4 = dt * R5(d4), dtd5 = dt * R5(d5);
    const float dtgamma = dt * R5(gamma);
    olu F;
    w_factor_rodas(&F, S, 1.0f / dtgamma, Jm, S->M);
    S->f(uprev, p, t, du);
    for (int c = 0; c < n; ++c) lt[c] = -(du[c] + dtd1 * dT[c]);
    w_solve(&F, n, lt, K1);
    for (int c = 0; c < n; ++c) u[c] = uprev[c] + R5(a21) * K1[c];
    S->f(u, p, t + R5(c2) * dt, du);
    for (int c = 0; c < n; ++c) cm[c] = dtC21 * K1[c];
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -((du[c] + dtd2 * dT[c]) + mc[c]);
    w_solve(&F, n, lt, K2);
    for (int c = 0; c < n; ++c) u[c] = (uprev[c] + R5(a31) * K1[c]) + R5(a32) * K2[c];
    S->f(u, p, t + R5(c3) * dt, du);
    for (int c = 0; c < n; ++c) cm[c] = (dtC31 * K1[c] + dtC32 * K2[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -((du[c] + dtd3 * dT[c]) + mc[c]);
    w_solve(&F, n, lt, K3);
    for (int c = 0; c < n; ++c) u[c] = ((uprev[c] + R5(a41) * K1[c]) + R5(a42) * K2[c]) + R5(a43) * K3[c];
    S->f(u, p, t + R5(c4) * dt, du);
    for (int c = 0; c < n; ++c) cm[c] = ((dtC41 * K1[c] + dtC42 * K2[c]) + dtC43 * K3[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -((du[c] + dtd4 * dT[c]) + mc[c]);
    w_solve(&F, n, lt, K4);
    for (int c = 0; c < n; ++c)
        u[c] = (((uprev[c] + R5(a51) * K1[c]) + R5(a52) * K2[c]) + R5(a53) * K3[c]) + R5(a54) * K4[c];
    S->f(u, p, t + R5(c5) * dt, du);
    for (int c = 0; c < n; ++c) cm[c] = (((dtC52 * K2[c] + dtC54 * K4[c]) + dtC51 * K1[c]) + dtC53 * K3[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -((du[c] + dtd5 * dT[c]) + mc[c]);
    w_solve(&F, n, lt, K5);
    for (int c = 0; c < n; ++c)
        u[c] = ((((uprev[c] + R5(a61) * K1[c]) + R5(a62) * K2[c]) + R5(a63) * K3[c]) + R5(a64) * K4[c]) + R5(a65) * K5[c];
    S->f(u, p, t + dt, du);
    for (int c = 0; c < n; ++c) cm[c] = ((((dtC61 * K1[c] + dtC62 * K2[c]) + dtC63 * K3[c]) + dtC64 * K4[c]) + dtC65 * K5[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -(du[c] + mc[c]);
    w_solve(&F, n, lt, K6);
    for (int c = 0; c < n; ++c) u[c] = u[c] + K6[c];
    S->f(u, p, t + dt, du);
    for (int c = 0; c < n; ++c) cm[c] = (((((dtC71 * K1[c] + dtC72 * K2[c]) + dtC73 * K3[c]) + dtC74 * K4[c]) + dtC75 * K5[c]) + dtC76 * K6[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -(du[c] + mc[c]);
    w_solve(&F, n, lt, K7);
    for (int c = 0; c < n; ++c) u[c] = u[c] + K7[c];
    S->f(u, p, t + dt, du);
    for (int c = 0; c < n; ++c) cm[c] = ((((((dtC81 * K1[c] + dtC82 * K2[c]) + dtC83 * K3[c]) + dtC84 * K4[c]) + dtC85 * K5[c]) + dtC86 * K6[c]) + dtC87 * K7[c]);
    mass_mul(S->M, n, cm, mc);
    for (int c = 0; c < n; ++c) lt[c] = -(du[c] + mc[c]);
    w_solve(&F, n, lt, K8);
    for (int c = 0; c < n; ++c) {
        unew[c] = u[c] + K8[c];
        kerr[c] = K8[c];
        k[1][c] = ((((((R5(h21) * K1[c] + R5(h22) * K2[c]) + R5(h23) * K3[c]) + R5(h24) * K4[c]) + R5(h25) * K5[c]) +
                    R5(h26) * K6[c]) + R5(h27) * K7[c]) + R5(h28) * K8[c];
        k[2][c] = ((((((R5(h31) * K1[c] + R5(h32) * K2[c]) + R5(h33) * K3[c]) + R5(h34) * K4[c]) + R5(h35) * K5[c]) +
                    R5(h36) * K6[c]) + R5(h37) * K7[c]) + R5(h38) * K8[c];
        k[3][c] = ((((((R5(h41) * K1[c] + R5(h42) * K2[c]) + R5(h43) * K3[c]) + R5(h44) * K4[c]) + R5(h45) * K5[c]) +
                    R5(h46) * K6[c]) + R5(h47) * K7[c]) + R5(h48) * K8[c];
    }
}

/* Rodas dense output (stiff/interpolants.jl:1-23): (1-th) y0 + th (y1 + (1-th)(k1 + th k2 [+ th k3]))
 * with y1 the integrator's current u */
static void rodas_interp(int method, float th, int n, const float *y0, const float *y1, float k[][OMAXN], float *out) {
    const float th1 = 1.0f - th;
    for (int c = 0; c < n; ++c) {
        if (method == OM_RODAS4) out[c] = th1 * y0[c] + th * (y1[c] + th1 * (k[1][c] + th * k[2][c]));
        else out[c] = th1 * y0[c] + th * (y1[c] + th1 * (k[1][c] + th * (k[2][c] + th * k[3][c])));
    }
}

/* ------------------------------------------------------------------------------------ */
/* Kvaerno3 / Kvaerno5 (gpu_kvaerno3_perform_step.jl, gpu_kvaerno5_perform_step.jl)             */
/* ------------------------------------------------------------------------------------ */
/* W = -M + (gamma dt) J(x, tc) (nlsolve/type.jl:129-140), M = I without a mass matrix, factored in FP64 */
static void kv_w(olu *F, const osys *S, const float *p, float gdt, const float *x, float tc) {
    const int m = jac_cols(S);
    float Jm[OMAXN * OMAXN];
    sys_jac(S, x, p, tc, Jm);
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c)
            F->lu[r][c] = (S->M ? -(double)S->M[r * m + c] : (r == c ? -1.0 : 0.0)) + (double)gdt * (double)Jm[r * m + c];
    w_lu(F, m, S->pivlu);
    w_dual(F, S, (double)gdt, Jm);
}

/* nlsolve (nlsolve/utils.jl:1-27): Newton on z for x = tmp + gamma z at tc = tnl + c dtnl */
static void kv_newton(const osys *S, const float *p, float gamma, float *z, const float *tmp, float tnl, float dtnl,
                      float c) {
    const int n = S->n;
    const float tc = tnl + c * dtnl;
    const float gdt = gamma * dtnl;
    for (int it = 0; it < 30; ++it) {
        float x[OMAXN], fe[OMAXN], r[OMAXN] = {0}, dz[OMAXN], mz[OMAXN];
        for (int q = 0; q < n; ++q) x[q] = tmp[q] + gamma * z[q];
        olu F;
        kv_w(&F, S, p, gdt, x, tc);
        S->f(x, p, tc, fe);
        mass_mul(S->M, n, z, mz); /* dt f - M z (nlsolve/utils.jl:17, :21) */
        for (int q = 0; q < n; ++q) r[q] = dtnl * fe[q] - mz[q];
        w_solve(&F, n, r, dz);
        for (int q = 0; q < n; ++q) z[q] = z[q] - dz[q];
        for (int q = 0; q < n; ++q) x[q] = tmp[q] + gamma * z[q];
        S->f(x, p, tc, fe);
        mass_mul(S->M, n, z, mz);
        for (int q = 0; q < n; ++q) r[q] = dtnl * fe[q] - mz[q];
        /* diffeqgpunorm (utils.jl:1-4): over a dual array, the values alone */
        if (rms_norm(r, S->dualp > 0 ? S->dualnv : n) < 100.0f * 1.1920928955078125e-7f) break;
    }
}

/* One Kvaerno step: z1 = dt k1, implicit stages at the solver's (tnl, dtnl); zl the last stage;
 * when err is non-NULL it receives W(u, t + dt) \ (btilde . z). */
#define K3(x) OKV3_##x
#define K5(x) OKV5_##x
static void kvaerno_step(int method, const osys *S, const float *p, const float *uprev, float t, float dt, float tnl,
                         float dtnl, const float *k1, float *zl, float *unew, float *err) {
    const int n = S->n;
    float z1[OMAXN] = {0}, z2[OMAXN] = {0}, z3[OMAXN] = {0}, tmp[OMAXN] = {0}, ev[OMAXN] = {0};
    const float g = (method == OM_KVAERNO3) ? K3(gamma) : K5(gamma);
    for (int q = 0; q < n; ++q) z1[q] = dt * k1[q];
    for (int q = 0; q < n; ++q) {
        z2[q] = z1[q];
        tmp[q] = uprev[q] + g * z1[q];
    }
    kv_newton(S, p, g, z2, tmp, tnl, dtnl, g);
    if (method == OM_KVAERNO3) {
        for (int q = 0; q < n; ++q) {
            z3[q] = K3(alpha31) * z1[q] + K3(alpha32) * z2[q];
            tmp[q] = (uprev[q] + K3(a31) * z1[q]) + K3(a32) * z2[q];
        }
        kv_newton(S, p, g, z3, tmp, tnl, dtnl, K3(c3));
        for (int q = 0; q < n; ++q) {
            zl[q] = (K3(a31) * z1[q] + K3(a32) * z2[q]) + g * z3[q];
            tmp[q] = ((uprev[q] + K3(a41) * z1[q]) + K3(a42) * z2[q]) + K3(a43) * z3[q];
        }
        kv_newton(S, p, g, zl, tmp, tnl, dtnl, 1.0f);
        for (int q = 0; q < n; ++q) unew[q] = tmp[q] + g * zl[q];
        for (int q = 0; q < n; ++q)
            ev[q] = ((K3(btilde1) * z1[q] + K3(btilde2) * z2[q]) + K3(btilde3) * z3[q]) + K3(btilde4) * zl[q];
    } else {
        float z4[OMAXN], z5[OMAXN], z6[OMAXN];
        for (int q = 0; q < n; ++q) {
            z3[q] = K5(alpha31) * z1[q] + K5(alpha32) * z2[q];
            tmp[q] = (uprev[q] + K5(a31) * z1[q]) + K5(a32) * z2[q];
        }
        kv_newton(S, p, g, z3, tmp, tnl, dtnl, K5(c3));
        for (int q = 0; q < n; ++q) {
            z4[q] = (K5(alpha41) * z1[q] + K5(alpha42) * z2[q]) + K5(alpha43) * z3[q];
            tmp[q] = ((uprev[q] + K5(a41) * z1[q]) + K5(a42) * z2[q]) + K5(a43) * z3[q];
        }
        kv_newton(S, p, g, z4, tmp, tnl, dtnl, K5(c4));
        for (int q = 0; q < n; ++q) {
            z5[q] = (K5(alpha51) * z1[q] + K5(alpha52) * z2[q]) + K5(alpha53) * z3[q];
            tmp[q] = (((uprev[q] + K5(a51) * z1[q]) + K5(a52) * z2[q]) + K5(a53) * z3[q]) + K5(a54) * z4[q];
        }
        kv_newton(S, p, g, z5, tmp, tnl, dtnl, K5(c5));
        for (int q = 0; q < n; ++q) {
            z6[q] = (K5(alpha61) * z1[q] + K5(alpha62) * z2[q]) + K5(alpha63) * z3[q];
            tmp[q] = (((uprev[q] + K5(a61) * z1[q]) + K5(a63) * z3[q]) + K5(a64) * z4[q]) + K5(a65) * z5[q];
        }
        kv_newton(S, p, g, z6, tmp, tnl, dtnl, K5(c6));
        for (int q = 0; q < n; ++q) {
            zl[q] = (((K5(a61) * z1[q] + K5(a63) * z3[q]) + K5(a64) * z4[q]) + K5(a65) * z5[q]) + g * z6[q];
            tmp[q] = ((((uprev[q] + K5(a71) * z1[q]) + K5(a73) * z3[q]) + K5(a74) * z4[q]) + K5(a75) * z5[q]) +
                     K5(a76) * z6[q];
        }
        kv_newton(S, p, g, zl, tmp, tnl, dtnl, 1.0f);
        for (int q = 0; q < n; ++q) unew[q] = tmp[q] + g * zl[q];
        for (int q = 0; q < n; ++q)
            ev[q] = ((((K5(btilde1) * z1[q] + K5(btilde3) * z3[q]) + K5(btilde4) * z4[q]) + K5(btilde5) * z5[q]) +
                     K5(btilde6) * z6[q]) + K5(btilde7) * zl[q];
    }
    if (err) {
        olu F;
        kv_w(&F, S, p, g * dt, unew, t + 1.0f * dt);
        w_solve(&F, n, ev, err);
    }
}

/* Hermite dense output (nonstiff/interpolants.jl:1-20) with y1 the integrator's current u */
static void hermite_interp(float th, float dt, int n, const float *y0, const float *y1, float k[][OMAXN], float *out) {
    const float a0 = 1.0f - th;
    const float s = th * (th - 1.0f);
    const float w = 1.0f - 2.0f * th;
    const float d1 = (th - 1.0f) * dt;
    const float d2 = th * dt;
    for (int c = 0; c < n; ++c) out[c] = (a0 * y0[c] + th * y1[c]) + s * ((w * (y1[c] - y0[c]) + d1 * k[1][c]) + d2 * k[2][c]);
}

/* ------------------------------------------------------------------------------------ */
/* Per-trajectory solver                                                                 */
/* ------------------------------------------------------------------------------------ */
/* Callback descriptors: the conditions and affects the tests use, restated in C. They play the
 * role of the user functions of a GPUDiscreteCallback / GPUContinuousCallback
 * (src/ensemblegpukernel/callbacks.jl:1-124). */
enum { OCB_DISCRETE = 0, OCB_CONTINUOUS = 1 };
enum { OCOND_T_EQ = 0,      /* t == cval                 (discrete) */
       OCOND_U0 = 1,        /* u[0]                      (continuous) */
       OCOND_U0_MINUS = 2,  /* u[0] - cval               (continuous) */
       OCOND_U0_LT = 3 };   /* u[0] < cval               (discrete) */
enum { OAFF_ADD = 1,        /* u[c] = u[c] + aval[c]            (integrator.u += SVector(aval)) */
       OAFF_SCALE = 2,      /* u[c] = u[c] + aval[c] * u[c]     (integrator.u += aval .* integrator.u) */
       OAFF_TERMINATE = 4,  /* terminate!(integrator) */
       OAFF_KEEP_FSAL = 8 };/* u_modified!(integrator, false) */
typedef struct {
    int32_t kind, cond;
    int32_t affect, affect_neg; /* OAFF_* bits, -1: nothing (continuous affect_neg! = nothing) */
    int32_t rootfind;           /* 0 NoRootFind, 1 LeftRootFind, 2 RightRootFind */
    float cval;
    float aval[4];
    float abstol, nudge, dtrelax;
} ocb;

typedef struct {
    const osys *S;
    int method;
    int adaptive;
    int64_t N;
    float dt, abstol, reltol;
    const float *saveat;
    int nsaveat, nsave, save_everystep;
    int kahan, detpow;
    int64_t max_attempts;
    float *ts, *us;
    const float *tstops;
    int ntstops;
    const ocb *cbs;
    int ncb;
} ojob;

/* One trajectory's integrator state (the fields of GPUTsit5Integrator & co. that the step,
 * the saves and the callbacks use; integrators/nonstiff/types.jl:3-38). */
typedef struct {
    int n;
    float p[OMAXN];
    float u[OMAXN], uprev[OMAXN];
    float k[OMAXS + 1][OMAXN];
    float rk1[OMAXN], rk2[OMAXN]; /* Rosenbrock23 interpolation slopes */
    float t, tprev, tf, tcomp;
    float dtstep;                 /* integrator.dt: fixed dt, or the last accepted step */
    float dtnew;                  /* adaptive: the next step's dt */
    int cur;                      /* 0-based saveat cursor */
    int step_idx;                 /* 0-based row of the next save_everystep write */
    int tsi;                      /* 0-based tstops_idx */
    int evlast;                   /* event_last_time (1-based callback index, 0 none) */
    int umod;                     /* u_modified */
    int retcode;
} ost;

static inline void put_u(const ojob *J, int row, int64_t i, const float *u) {
    int n = J->S->n;
    for (int c = 0; c < n; ++c) J->us[((int64_t)row * n + c) * J->N + i] = u[c];
}

/* t <- t + dt, optionally compensated */
static inline float advance_t(float t, float dt, float *comp, int kahan) {
    if (!kahan) return t + dt;
    float y = dt - *comp;
    float tn = t + y;
    *comp = (tn - t) - y;
    return tn;
}

/* the next float above x (x finite) */
static inline float next_up(float x) {
    if (x == 0.0f) return 1.4e-45f;
    uint32_t b;
    memcpy(&b, &x, 4);
    b = (x > 0.0f) ? b + 1u : b - 1u;
    float r;
    memcpy(&r, &b, 4);
    return r;
}

/* An accepted step that leaves t unchanged (t + dt rounds to t) and proposes no larger dt
 * would repeat forever in the reference's loop. Here the lane snaps to tf when tf is within
 * two units of t's resolution, and otherwise stops with DtLessThanMin. Returns 1 when the lane
 * must stop. */
static inline int no_progress(float *t, float tf) {
    float ulp = next_up(*t) - *t;
    if (tf - *t <= 2.0f * ulp) {
        *t = tf;
        return 0;
    }
    return 1;
}

/* integrator(x): dense output of the last step at time x, Θ = (x - tprev) / dt
 * (nonstiff/types.jl:41-44, stiff/types.jl:921-939) */
static void interp_at(const ojob *J, const ost *s, float x, float *out) {
    const float th = (x - s->tprev) / s->dtstep;
    const float d = 1.0f / (2.0f + sqrtf(2.0f));
    if (J->method == OM_RODAS4 || J->method == OM_RODAS5P) rodas_interp(J->method, th, s->n, s->uprev, s->u, (float(*)[OMAXN])s->k, out);
    else if (J->method == OM_KVAERNO3 || J->method == OM_KVAERNO5) hermite_interp(th, s->dtstep, s->n, s->uprev, s->u, (float(*)[OMAXN])s->k, out);
    else if (J->method == OM_TSIT5) tsit5_interp(th, s->dtstep, s->n, s->uprev, (float(*)[OMAXN])s->k, out);
    else if (J->method == OM_RB23) rb23_interp(th, s->dtstep, s->n, d, s->uprev, s->rk1, s->rk2, out);
    else if (J->method == OM_VERN9) vern9_interp(J->S, s->p, th, s->dtstep, s->tprev, s->n, s->uprev, (float(*)[OMAXN])s->k, out);
    else vern7_interp(J->S, s->p, th, s->dtstep, s->tprev, s->n, s->uprev, (float(*)[OMAXN])s->k, out);
}

/* savevalues! (integrator_utils.jl:13-50) */
static void savevalues(const ojob *J, int64_t i, ost *s) {
    if (J->saveat) {
        while (s->cur < J->nsaveat && J->saveat[s->cur] <= s->t) {
            const float st = J->saveat[s->cur];
            float out[OMAXN];
            interp_at(J, s, st, out);
            if (s->cur < J->nsave) {
                put_u(J, s->cur, i, out);
                J->ts[(int64_t)s->cur * J->N + i] = st;
            }
            ++s->cur;
        }
    } else if (J->save_everystep && !J->adaptive) {
        if (s->step_idx >= 0 && s->step_idx < J->nsave) {
            put_u(J, s->step_idx, i, s->u);
            J->ts[(int64_t)s->step_idx * J->N + i] = s->t;
        }
        ++s->step_idx;
    }
}

/* ---- callbacks (integrator_utils.jl:52-479) ---- */
static inline float jl_sign(float x) { return (x < 0.0f) ? -1.0f : ((x > 0.0f) ? 1.0f : x); }

/* Julia eps(x::Float32) */
static float jl_eps(float x) {
    float ax = fabsf(x);
    if (!isfinite(ax)) return NAN;
    if (ax < FLT_MIN) return 1.4e-45f;
    int e;
    frexpf(ax, &e); /* ax = m 2^e, m in [0.5, 1): exponent(x) = e - 1 */
    return ldexpf(1.0f, (e - 1) - 23);
}

/* Julia nextfloat(x::Float32) */
static float jl_nextfloat(float x) {
    if (isnan(x) || x == INFINITY) return x;
    return next_up(x);
}

static float ocb_ccond(const ocb *c, const float *u, float t) {
    (void)t;
    switch (c->cond) {
        case OCOND_U0: return u[0];
        case OCOND_U0_MINUS: return u[0] - c->cval;
        default: return NAN;
    }
}

static int ocb_dcond(const ocb *c, const float *u, float t) {
    switch (c->cond) {
        case OCOND_T_EQ: return t == c->cval;
        case OCOND_U0_LT: return u[0] < c->cval;
        default: return 0;
    }
}

/* applies affect bits to u; returns the bits */
static int ocb_apply(const ocb *c, int bits, float *u, int n) {
    if (bits & OAFF_ADD)
        for (int q = 0; q < n && q < 4; ++q) u[q] = u[q] + c->aval[q];
    if (bits & OAFF_SCALE)
        for (int q = 0; q < n && q < 4; ++q) u[q] = u[q] + c->aval[q] * u[q];
    return bits;
}

static void apply_affect_flags(ost *s, int bits) {
    if (bits & OAFF_TERMINATE) s->retcode = ORET_TERMINATED; /* terminate! (integrator_utils.jl:52-67) */
    if (bits & OAFF_KEEP_FSAL) s->umod = 0;                  /* u_modified!(integrator, false) */
}

/* DiffEqBase.get_condition (integrator_utils.jl:460-478) */
static float cond_at(const ojob *J, const ost *s, const ocb *c, float x) {
    if (x == s->t) return ocb_ccond(c, s->u, x);
    if (x == s->tprev) return ocb_ccond(c, s->uprev, x);
    float tmp[OMAXN];
    interp_at(J, s, x, tmp);
    return ocb_ccond(c, tmp, x);
}

/* gpu_find_root (integrator_utils.jl:327-384): ITP with scaled_k1 = 0.2, k2 = 2, n0 = 10 */
static float find_root(const ojob *J, const ost *s, const ocb *c, float left, float right, int rootfind) {
    float fl = cond_at(J, s, c, left);
    float fr = cond_at(J, s, c, right);
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
        const float xt = (delta <= fabsf(diff)) ? xf + copysignf(delta, diff) : mid;
        const float xp = (fabsf(xt - mid) <= r) ? xt : mid - copysignf(r, diff);
        const float yp = cond_at(J, s, c, xp);
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
    return (rootfind == 1) ? left : right;
}

/* handle_callbacks! (integrator_utils.jl:274-325). Returns saved_in_cb. */
static int handle_callbacks(const ojob *J, int64_t i, ost *s) {
    int csaved = 0, ndisc = 0, ncont = 0;
    for (int q = 0; q < J->ncb; ++q) {
        if (J->cbs[q].kind == OCB_DISCRETE) ++ndisc;
        else ++ncont;
    }
    if (ncont > 0) {
        /* DiffEqBase.find_first_continuous_callback over find_callback_time (:386-446) */
        float tmin = s->t, upc = 0.0f;
        int occ = 0, idx = 1, j = 0;
        const ocb *first = NULL;
        for (int q = 0; q < J->ncb; ++q) {
            const ocb *c = &J->cbs[q];
            if (c->kind != OCB_CONTINUOUS) continue;
            if (!first) first = c;
            float bt = s->tprev;
            float bc = ocb_ccond(c, s->uprev, s->tprev);
            if (s->evlast == j + 1 && fabsf(bc - 0.0f) <= c->abstol) {
                bt = s->tprev + s->dtstep * c->nudge;
                bc = cond_at(J, s, c, bt);
            }
            const float bs = jl_sign(bc);
            const float tsg = jl_sign(cond_at(J, s, c, s->t));
            const int oc = ((bs < 0.0f && c->affect >= 0) || (bs > 0.0f && c->affect_neg >= 0)) && bs * tsg <= 0.0f;
            float tj = s->t;
            if (oc && c->rootfind != 0 && tsg != 0.0f) tj = find_root(J, s, c, bt, s->t, c->rootfind);
            if (j == 0) {
                tmin = tj;
                upc = bs;
                occ = oc;
                idx = 1;
            } else if (oc && (tj < tmin || !occ)) {
                tmin = tj;
                upc = bs;
                occ = 1;
                idx = j + 1;
            }
            ++j;
        }
        if (occ) {
            s->evlast = idx;
            /* apply_callback! (:232-272) with the first continuous callback's affects */
            if (tmin < s->tprev) {
                s->retcode = ORET_FAILURE; /* error("Current interpolant only works between tprev and t") */
            } else {
                if (tmin != s->t) {
                    /* _change_t_via_interpolation! (:174-201) */
                    float tmp[OMAXN];
                    interp_at(J, s, tmin, tmp);
                    memcpy(s->u, tmp, sizeof(float) * s->n);
                    s->step_idx -= (int)rintf((s->t - tmin) / s->dtstep);
                    s->t = tmin;
                    s->tcomp = 0.0f;
                }
                if (J->adaptive && s->dtnew < 1.0e-12f) s->dtnew = jl_min(first->dtrelax * s->dtstep, fabsf(s->tf - s->t));
                savevalues(J, i, s);
                csaved = 1;
                s->umod = 1;
                if (upc < 0.0f) {
                    if (first->affect >= 0) apply_affect_flags(s, ocb_apply(first, first->affect, s->u, s->n));
                    else s->umod = 0;
                } else if (upc > 0.0f) {
                    if (first->affect_neg >= 0) apply_affect_flags(s, ocb_apply(first, first->affect_neg, s->u, s->n));
                    else s->umod = 0;
                }
            }
        } else {
            s->evlast = 0;
        }
    }
    if (ndisc > 0) {
        /* apply_discrete_callback! (:70-149), in order; its result replaces the continuous one */
        int dsaved = 0;
        for (int q = 0; q < J->ncb; ++q) {
            const ocb *c = &J->cbs[q];
            if (c->kind != OCB_DISCRETE) continue;
            if (ocb_dcond(c, s->u, s->t)) {
                savevalues(J, i, s);
                dsaved = 1;
                s->umod = 1;
                apply_affect_flags(s, ocb_apply(c, c->affect, s->u, s->n));
            }
        }
        return dsaved;
    }
    return csaved;
}

#define O_TSTOP_SLACK (100.0f * 1.1920928955078125e-7f) /* 100 eps(Float32) */

static int solve_traj(const ojob *J, int64_t i, const float *u0, const float *p, float t0, float tf,
                      int32_t *nacc_out, int32_t *nrej_out) {
    const osys *S = J->S;
    const int n = S->n;
    const erk32 *T = (J->method == OM_TSIT5) ? &TAB_T5 : (J->method == OM_VERN9 ? &TAB_V9 : &TAB_V7);
    ost st;
    ost *s = &st;
    memset(s, 0, sizeof(st));
    s->n = n;
    memcpy(s->p, p, sizeof(float) * S->np);
    float unew[OMAXN];
    float (*k)[OMAXN] = s->k;
    memcpy(s->u, u0, sizeof(float) * n);
    memcpy(s->uprev, u0, sizeof(float) * n);
    s->retcode = ORET_SUCCESS;
    s->tf = tf;
    s->evlast = 1; /* nonstiff/types.jl:318 */
    s->umod = 1;   /* u_modified = true at init: the first step evaluates k1 */
    int32_t nacc = 0, nrej = 0;
    int64_t attempts = 0;
    const int use_saveat = (J->saveat != NULL);
    s->step_idx = 1;

    /* kernels.jl:35-45 / :116-126: initial save */
    if (use_saveat) {
        if (J->nsaveat > 0 && t0 == J->saveat[0]) {
            put_u(J, 0, i, u0);
            s->cur = 1;
        }
    } else {
        J->ts[i] = t0;
        put_u(J, 0, i, u0);
    }

    s->t = t0;
    s->tprev = t0;
    s->tcomp = 0.0f;
    const float d = 1.0f / (2.0f + sqrtf(2.0f));
    s->dtstep = J->dt; /* fixed dt, or the accepted dt of the last adaptive step */

    if (!J->adaptive) {
        /* ---------------- fixed step: kernels.jl:49-52 -------------------------------- */
        const float dt = J->dt;
        while (s->t < tf && s->retcode != ORET_TERMINATED) {
            if (attempts >= J->max_attempts) { s->retcode = ORET_MAXITERS; break; }
            ++attempts;
            memcpy(s->uprev, s->u, sizeof(float) * n);
            s->tprev = s->t;
            const float told = s->t;
            float dts = dt;
            /* gpu_tsit5_perform_step.jl:18-27: a step reaching the next stopping time ends there;
             * integrator.dt (the dense output's) stays dt */
            if (s->tsi < J->ntstops && ((J->tstops[s->tsi] - s->t) - dt) - O_TSTOP_SLACK < 0.0f) {
                s->t = J->tstops[s->tsi];
                dts = s->t - told;
                ++s->tsi;
                s->tcomp = 0.0f;
            } else {
                s->t = advance_t(s->t, dt, &s->tcomp, J->kahan);
                if (s->t == told && !J->kahan) {
                    s->retcode = ORET_DTMIN; /* the fixed step is below the resolution of t */
                    --attempts;
                    break;
                }
            }
            if (J->method == OM_KVAERNO3 || J->method == OM_KVAERNO5) {
                /* gpu_kvaerno*_perform_step.jl fixed: the nonlinear solver gets the advanced t and the
                 * integrator's dt; k1 is the last step's f(u, t) unless u was modified */
                float zl[OMAXN];
                if (s->umod) S->f(s->uprev, s->p, told, k[1]);
                kvaerno_step(J->method, S, s->p, s->uprev, told, dts, s->t, dt, k[1], zl, s->u, NULL);
                for (int c = 0; c < n; ++c) k[2][c] = zl[c] / dts;
                S->f(s->u, s->p, told, k[1]);
            } else if (J->method == OM_RODAS4 || J->method == OM_RODAS5P) {
                float kerr[OMAXN];
                if (J->method == OM_RODAS4) rodas4_step(S, s->p, s->uprev, told, dts, k, s->u, kerr);
                else rodas5p_step(S, s->p, s->uprev, told, dts, k, s->u, kerr);
            } else if (J->method == OM_RB23) {
                /* gpu_rosenbrock23_perform_step.jl:1-70 */
                float Jm[OMAXN * OMAXN], dT[OMAXN], F0[OMAXN], rhs[OMAXN], tmp[OMAXN], F1[OMAXN], x[OMAXN];
                float gam = dts * d;
                float dto2 = dts / 2.0f;
                sys_jac(S, s->uprev, s->p, told, Jm);
                sys_tgrad(S, s->uprev, s->p, told, dT);
                olu F;
                w_factor(&F, S, gam, Jm, S->M);
                S->f(s->uprev, s->p, told, F0);
                for (int c = 0; c < n; ++c) rhs[c] = F0[c] + gam * dT[c];
                w_solve(&F, n, rhs, s->rk1);
                for (int c = 0; c < n; ++c) tmp[c] = s->uprev[c] + dto2 * s->rk1[c];
                S->f(tmp, s->p, told + dto2, F1);
                /* F1 - k1, or F1 - M k1 with a mass matrix (:54-58) */
                mass_mul(S->M, n, s->rk1, x);
                for (int c = 0; c < n; ++c) rhs[c] = F1[c] - x[c];
                w_solve(&F, n, rhs, x);
                for (int c = 0; c < n; ++c) s->rk2[c] = x[c] + s->rk1[c];
                for (int c = 0; c < n; ++c) s->u[c] = s->uprev[c] + dts * s->rk2[c];
            } else {
                /* gpu_tsit5_perform_step.jl:1-64, gpu_vern7_perform_step.jl:1-74 */
                if (T->fsal && !s->umod) {
                    memcpy(k[1], k[T->s], sizeof(float) * n);
                } else {
                    S->f(s->uprev, s->p, told, k[1]);
                }
                erk_attempt(T, S, s->p, s->uprev, told, dts, k, unew);
                memcpy(s->u, unew, sizeof(float) * n);
            }
            s->umod = 0;
            ++nacc;
            int saved_in_cb = (J->ncb > 0) ? handle_callbacks(J, i, s) : 0;
            if (s->retcode == ORET_FAILURE) break;
            if (!saved_in_cb) savevalues(J, i, s);
        }
        /* kernels.jl:53-62 */
        if (s->t > tf && !use_saveat && J->save_everystep &&
            (s->retcode == ORET_SUCCESS || s->retcode == ORET_TERMINATED)) {
            float out[OMAXN];
            int row = J->nsave - 1;
            interp_at(J, s, tf, out);
            put_u(J, row, i, out);
            J->ts[(int64_t)row * J->N + i] = tf;
        }
    } else {
        /* ---------------- adaptive: kernels.jl:128-131 -------------------------------- */
        static const int orders[8] = {5, 7, 2, 9, 4, 5, 3, 5};
        const int order = orders[J->method];
        /* integrator_utils.jl:1-11 */
        const float beta1 = (float)(7.0 / (10.0 * order));
        const float beta2 = (float)(2.0 / (5.0 * order));
        const float qmax = 10.0f, qmin = (float)(1.0 / 5.0), gamma = (float)(9.0 / 10.0);
        const float qoldinit = (float)1.0e-4;
        float qold = qoldinit;
        float dt = J->dt; /* integ.dtnew */
        const float e32 = 6.0f + sqrtf(2.0f);
        while (s->t < tf && s->retcode != ORET_TERMINATED) {
            memcpy(s->uprev, s->u, sizeof(float) * n);
            const float t = s->t;
            float EEst = INFINITY;
            /* FSAL: gpu_tsit5_perform_step.jl:92-97 */
            if (J->method == OM_TSIT5) {
                if (!s->umod) memcpy(k[1], k[7], sizeof(float) * n);
                else S->f(s->uprev, s->p, t, k[1]);
            }
            s->umod = 0;
            int stop = 0;
            while (EEst > 1.0f) {
                if (dt < 1.0e-14f) { s->retcode = ORET_DTMIN; stop = 1; break; }
                if (attempts >= J->max_attempts) { s->retcode = ORET_MAXITERS; stop = 1; break; }
                ++attempts;
                float q11 = 0.0f, q;
                if (J->method == OM_KVAERNO3 || J->method == OM_KVAERNO5) {
                    /* gpu_kvaerno*_perform_step.jl adaptive: k1 = f(uprev, t) in every attempt */
                    float zl[OMAXN], e[OMAXN];
                    S->f(s->uprev, s->p, t, k[1]);
                    kvaerno_step(J->method, S, s->p, s->uprev, t, dt, t, dt, k[1], zl, unew, e);
                    for (int c = 0; c < n; ++c) k[2][c] = zl[c] / dt;
                    EEst = scaled_eest(S, e, s->uprev, unew, J->abstol, J->reltol);
                } else if (J->method == OM_RODAS4 || J->method == OM_RODAS5P) {
                    /* gpu_rodas4_perform_step.jl:158-238, gpu_rodas5P_perform_step.jl:186-296 */
                    float e[OMAXN];
                    if (J->method == OM_RODAS4) rodas4_step(S, s->p, s->uprev, t, dt, k, unew, e);
                    else rodas5p_step(S, s->p, s->uprev, t, dt, k, unew, e);
                    EEst = scaled_eest(S, e, s->uprev, unew, J->abstol, J->reltol);
                } else if (J->method == OM_RB23) {
                    /* gpu_rosenbrock23_perform_step.jl:106-153 */
                    float Jm[OMAXN * OMAXN], dT[OMAXN], F0[OMAXN], F1[OMAXN], F2[OMAXN], rhs[OMAXN], tmp[OMAXN], x[OMAXN];
                    float gam = dt * d;
                    float dto2 = dt / 2.0f;
                    float dto6 = dt / 6.0f;
                    sys_jac(S, s->uprev, s->p, t, Jm);
                    sys_tgrad(S, s->uprev, s->p, t, dT);
                    olu F;
                    w_factor(&F, S, gam, Jm, S->M);
                    S->f(s->uprev, s->p, t, F0);
                    for (int c = 0; c < n; ++c) rhs[c] = F0[c] + gam * dT[c];
                    w_solve(&F, n, rhs, s->rk1);
                    for (int c = 0; c < n; ++c) tmp[c] = s->uprev[c] + dto2 * s->rk1[c];
                    S->f(tmp, s->p, t + dto2, F1);
                    /* F1 - k1, or F1 - M k1 with a mass matrix (:129-133) */
                    mass_mul(S->M, n, s->rk1, x);
                    for (int c = 0; c < n; ++c) rhs[c] = F1[c] - x[c];
                    w_solve(&F, n, rhs, x);
                    for (int c = 0; c < n; ++c) s->rk2[c] = x[c] + s->rk1[c];
                    for (int c = 0; c < n; ++c) unew[c] = s->uprev[c] + dt * s->rk2[c];
                    S->f(unew, s->p, t + dt, F2);
                    if (!S->M) {
                        for (int c = 0; c < n; ++c)
                            rhs[c] = ((F2[c] - e32 * (s->rk2[c] - F1[c])) - 2.0f * (s->rk1[c] - F0[c])) + dt * dT[c];
                    } else {
                        /* F2 - M (e32 k2 + 2 k1) + e32 F1 + 2 F0 + dt dT (:144-148) */
                        float v[OMAXN], mv[OMAXN];
                        for (int c = 0; c < n; ++c) v[c] = e32 * s->rk2[c] + 2.0f * s->rk1[c];
                        mass_mul(S->M, n, v, mv);
                        for (int c = 0; c < n; ++c) rhs[c] = (((F2[c] - mv[c]) + e32 * F1[c]) + 2.0f * F0[c]) + dt * dT[c];
                    }
                    float k3[OMAXN];
                    w_solve(&F, n, rhs, k3);
                    float e[OMAXN] = {0};
                    for (int c = 0; c < n; ++c) e[c] = dto6 * ((s->rk1[c] - 2.0f * s->rk2[c]) + k3[c]);
                    EEst = scaled_eest(S, e, s->uprev, unew, J->abstol, J->reltol);
                } else {
                    /* gpu_vern7_perform_step.jl:117, gpu_vern9_perform_step.jl:170: k1 in every attempt */
                    if (J->method == OM_VERN7 || J->method == OM_VERN9) S->f(s->uprev, s->p, t, k[1]);
                    erk_attempt(T, S, s->p, s->uprev, t, dt, k, unew);
                    EEst = erk_eest(T, n, dt, k, s->uprev, unew, J->abstol, J->reltol, S->dualnv, S->dualp);
                }
                /* PI controller: gpu_tsit5_perform_step.jl:124-137 */
                if (EEst == 0.0f) {
                    q = 1.0f / qmax;
                } else {
                    q11 = ctl_pow(EEst, beta1, J->detpow);
                    q = q11 / ctl_pow(qold, beta2, J->detpow);
                }
                if (EEst > 1.0f) {
                    dt = dt / jl_min(1.0f / qmin, q11 / gamma);
                    ++nrej;
                } else {
                    q = jl_max(1.0f / qmax, jl_min(1.0f / qmin, q / gamma));
                    qold = jl_max(EEst, qoldinit);
                    float dtnew = dt / q;
                    dtnew = jl_min(fabsf(dtnew), fabsf((tf - t) - dt));
                    s->dtstep = dt;
                    s->tprev = t;
                    memcpy(s->u, unew, sizeof(float) * n);
                    if ((tf - t) - dt < 1.0e-14f) {
                        s->t = tf;
                        s->tcomp = 0.0f;
                    } else if (s->tsi < J->ntstops && ((J->tstops[s->tsi] - t) - dt) - O_TSTOP_SLACK < 0.0f) {
                        /* gpu_tsit5_perform_step.jl:158-164: end at the stopping time, state from
                         * this step's dense output */
                        s->t = J->tstops[s->tsi];
                        s->tcomp = 0.0f;
                        ++s->tsi;
                        float tmp[OMAXN];
                        interp_at(J, s, s->t, tmp);
                        memcpy(s->u, tmp, sizeof(float) * n);
                    } else {
                        float tn = advance_t(t, dt, &s->tcomp, J->kahan);
                        if (tn == t && !J->kahan && !(dtnew > dt) && no_progress(&tn, tf)) {
                            s->retcode = ORET_DTMIN;
                            stop = 1;
                        }
                        s->t = tn;
                    }
                    dt = dtnew;
                    ++nacc;
                    if (stop) break;
                }
            }
            if (stop) break;
            s->dtnew = dt;
            int saved_in_cb = (J->ncb > 0) ? handle_callbacks(J, i, s) : 0;
            dt = s->dtnew;
            if (s->retcode == ORET_FAILURE) break;
            /* savevalues! after the accepted step */
            if (!saved_in_cb) savevalues(J, i, s);
        }
    }
    /* kernels.jl:59-62: endpoint-only output keeps the last state */
    if (!use_saveat && !J->save_everystep) {
        put_u(J, 1, i, s->u);
        J->ts[J->N + i] = s->t;
    }
    if (s->retcode == ORET_SUCCESS) {
        int fin = isfinite(s->t);
        for (int c = 0; c < n; ++c) fin = fin && isfinite(s->u[c]);
        if (!fin) s->retcode = ORET_NONFINITE;
    }
    *nacc_out = nacc;
    *nrej_out = nrej;
    return s->retcode;
}

/* ------------------------------------------------------------------------------------ */
/* Exported: Float32 restatement over an ensemble (SoA layout, same as the HIP solver)   */
/* ------------------------------------------------------------------------------------ */
/* u0[c*N + i], p[j*N + i]; outputs ts[row*N + i], us[(row*n + c)*N + i].
 * t0s/tfs may be NULL (then the shared t0/tf apply). saveat may be NULL. tstops (ntstops) may
 * be NULL; cbs holds ncb callback descriptors (discrete and continuous, each kind in order).
 * Returns 0, or a negative value for an unsupported combination. */
int deg_oracle_solve_cb(int system, int method, int adaptive, int64_t N,
                        const float *u0, const float *p, const float *t0s, const float *tfs,
                        float t0, float tf, float dt, float abstol, float reltol,
                        const float *saveat, int nsaveat, int nsave, int save_everystep,
                        int flags, int64_t max_attempts,
                        float *ts, float *us, int32_t *retcode, int32_t *naccept, int32_t *nreject,
                        int nthreads, const float *tstops, int ntstops, const ocb *cbs, int ncb) {
    if (system < 0 || system >= OSYS_COUNT) return -1;
    if (method < OM_TSIT5 || method > OM_KVAERNO5) return -1;
    if (ntstops < 0 || ncb < 0 || (ntstops > 0 && !tstops) || (ncb > 0 && !cbs)) return -1;
    /* a system without an analytic Jacobian (the dual Lorenz) is solved by the explicit methods */
    if (!OSYSTEMS[system].jac && !(flags & DEG_O_FDJAC) &&
        !(method == OM_TSIT5 || method == OM_VERN7 || method == OM_VERN9))
        return -1;
    init_tabs();
    osys Sys = OSYSTEMS[system];
    Sys.fd = (flags & DEG_O_FDJAC) != 0;
    Sys.pivlu = (flags & DEG_O_PIVLU) != 0;
    if (flags & DEG_O_AUGMENTED) {
        Sys.dualnv = 0;
        Sys.dualp = 0;
    }
    const osys *S = &Sys;
    ojob J;
    J.S = S;
    J.method = method;
    J.adaptive = adaptive;
    J.N = N;
    J.dt = dt;
    J.abstol = abstol;
    J.reltol = reltol;
    J.saveat = saveat;
    J.nsaveat = nsaveat;
    J.nsave = nsave;
    J.save_everystep = save_everystep;
    J.kahan = (flags & DEG_O_KAHAN) != 0;
    J.detpow = (flags & DEG_O_DETPOW) != 0;
    J.max_attempts = max_attempts > 0 ? max_attempts : (int64_t)1 << 62;
    J.ts = ts;
    J.us = us;
    J.tstops = tstops;
    J.ntstops = tstops ? ntstops : 0;
    J.cbs = cbs;
    J.ncb = cbs ? ncb : 0;
    int bad = 0;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) reduction(| : bad)
    for (int64_t i = 0; i < N; ++i) {
        float uu[OMAXN], pp[OMAXN];
        for (int c = 0; c < S->n; ++c) uu[c] = u0[(int64_t)c * N + i];
        for (int j = 0; j < S->np; ++j) pp[j] = p[(int64_t)j * N + i];
        float a = t0s ? t0s[i] : t0;
        float b = tfs ? tfs[i] : tf;
        int32_t na = 0, nr = 0;
        int rc = solve_traj(&J, i, uu, pp, a, b, &na, &nr);
        if (rc < 0) bad = 1;
        if (retcode) retcode[i] = rc < 0 ? -1 : rc;
        if (naccept) naccept[i] = na;
        if (nreject) nreject[i] = nr;
    }
    return bad ? -2 : 0;
}

/* The Jacobian and time gradient an implicit step would use at (u, p, t): the system's analytic
 * ones (fd = 0) or the forward differences of autodiff = false (fd = 1). */
int deg_oracle_jac(int system, int fd, const float *u, const float *p, float t, float *J, float *dT) {
    if (system < 0 || system >= OSYS_COUNT) return -1;
    osys Sys = OSYSTEMS[system];
    Sys.fd = fd != 0;
    sys_jac(&Sys, u, p, t, J);
    sys_tgrad(&Sys, u, p, t, dT);
    return 0;
}

int deg_oracle_solve(int system, int method, int adaptive, int64_t N,
                     const float *u0, const float *p, const float *t0s, const float *tfs,
                     float t0, float tf, float dt, float abstol, float reltol,
                     const float *saveat, int nsaveat, int nsave, int save_everystep,
                     int flags, int64_t max_attempts,
                     float *ts, float *us, int32_t *retcode, int32_t *naccept, int32_t *nreject,
                     int nthreads) {
    return deg_oracle_solve_cb(system, method, adaptive, N, u0, p, t0s, tfs, t0, tf, dt, abstol, reltol, saveat,
                               nsaveat, nsave, save_everystep, flags, max_attempts, ts, us, retcode, naccept,
                               nreject, nthreads, NULL, 0, NULL, 0);
}

/* ------------------------------------------------------------------------------------ */
/* FP64 Vern9 with the same PI controller, stepping exactly onto each save time          */
/* ------------------------------------------------------------------------------------ */
static int vern9_traj(orhs_fn_d f, int n, const double *u0, const double *p, double t0,
                      double abstol, double reltol, double dt0, const double *tsave, int nsave,
                      double *out, int64_t max_steps, int64_t *nsteps_out) {
    const int s = OV9_S;
    const double beta1 = 7.0 / 90.0, beta2 = 2.0 / 45.0, qmax = 10.0, qmin = 0.2, gamma = 0.9;
    double u[OMAXN], un[OMAXN], tmp[OMAXN], k[OMAXS + 1][OMAXN];
    memcpy(u, u0, sizeof(double) * n);
    double t = t0, dt = dt0, qold = 1e-4;
    int64_t steps = 0;
    int next = 0;
    while (next < nsave && tsave[next] <= t0) {
        for (int c = 0; c < n; ++c) out[(int64_t)next * n + c] = u0[c];
        ++next;
    }
    while (next < nsave) {
        double target = tsave[next];
        double h = dt;
        int hit = 0;
        if (t + h >= target) {
            h = target - t;
            hit = 1;
        }
        if (++steps > max_steps) return 2;
        f(u, p, t, k[1]);
        for (int i = 2; i <= s; ++i) {
            for (int c = 0; c < n; ++c) {
                double acc = 0.0;
                for (int j = 1; j < i; ++j)
                    if (OV9_AM[i][j]) acc += OV9_A[i][j] * k[j][c];
                tmp[c] = u[c] + h * acc;
            }
            f(tmp, p, t + OV9_C[i] * h, k[i]);
        }
        double err = 0.0;
        for (int c = 0; c < n; ++c) {
            double acc = 0.0, eacc = 0.0;
            for (int j = 1; j <= s; ++j) {
                if (OV9_BM[j]) acc += OV9_B[j] * k[j][c];
                if (OV9_BTM[j]) eacc += OV9_BT[j] * k[j][c];
            }
            un[c] = u[c] + h * acc;
            double sc = abstol + fmax(fabs(u[c]), fabs(un[c])) * reltol;
            double e = h * eacc / sc;
            err += e * e;
        }
        double EEst = sqrt(err / n);
        if (!(EEst == EEst)) return 3;
        if (EEst > 1.0) {
            double q11 = pow(EEst, beta1);
            dt = h / fmin(1.0 / qmin, q11 / gamma);
            if (dt < 1e-300) return 1;
            continue;
        }
        double q = (EEst == 0.0) ? 1.0 / qmax : pow(EEst, beta1) / pow(qold, beta2);
        q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
        qold = fmax(EEst, 1e-4);
        memcpy(u, un, sizeof(double) * n);
        if (hit) {
            t = target;
            for (int c = 0; c < n; ++c) out[(int64_t)next * n + c] = u[c];
            ++next;
            /* keep the controller's proposal for a full step, not the clipped one */
            double prop = h / q;
            dt = (prop > dt) ? prop : dt;
        } else {
            t = t + h;
            dt = h / q;
        }
    }
    if (nsteps_out) *nsteps_out = steps;
    return 0;
}

/* out[(row*n + c)*N + i] in double. tsave (shared, double) may be NULL: then one row at the
 * per-trajectory final time (tfs[i] or tf). */
int deg_oracle_vern9(int system, int64_t N, const float *u0, const float *p, const float *t0s,
                     const float *tfs, double t0, double tf, double abstol, double reltol,
                     double dt0, const double *tsave, int nsave, double *out, int32_t *status,
                     int64_t *nsteps, int nthreads) {
    if (system < 0 || system >= OSYS_COUNT) return -1;
    const osys *S = &OSYSTEMS[system];
    orhs_fn_d f = OSYSTEMS_D[system];
    const int n = S->n;
    if (nthreads <= 0) nthreads = 1;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads) reduction(| : bad)
    for (int64_t i = 0; i < N; ++i) {
        double uu[OMAXN], pp[OMAXN];
        double ts1[1];
        for (int c = 0; c < n; ++c) uu[c] = (double)u0[(int64_t)c * N + i];
        for (int j = 0; j < S->np; ++j) pp[j] = (double)p[(int64_t)j * N + i];
        double a = t0s ? (double)t0s[i] : t0;
        double b = tfs ? (double)tfs[i] : tf;
        const double *tsv = tsave;
        int ns = nsave;
        if (!tsave) {
            ts1[0] = b;
            tsv = ts1;
            ns = 1;
        }
        double *row = (double *)malloc(sizeof(double) * (size_t)ns * (size_t)n);
        if (!row) {
            bad = 1;
            continue;
        }
        int64_t st = 0;
        int rc = vern9_traj(f, n, uu, pp, a, abstol, reltol, dt0, tsv, ns, row, (int64_t)1 << 40, &st);
        if (status) status[i] = rc;
        if (nsteps) nsteps[i] = st;
        for (int r = 0; r < ns; ++r)
            for (int c = 0; c < n; ++c) out[((int64_t)r * n + c) * N + i] = row[r * n + c];
        free(row);
    }
    return bad ? -2 : 0;
}

/* ------------------------------------------------------------------------------------ */
/* Introspection used by the order-condition and helper tests                           */
/* ------------------------------------------------------------------------------------ */
float deg_oracle_det_pow(float x, float y) { return det_pow(x, y); }

/* the dual error norm on one hand-made state (host checks of its abs and max rules) */
float deg_oracle_dual_eest(int nv, int P, int n, const float *e, const float *uprev, const float *u, float abstol,
                           float reltol) {
    return dual_eest(nv, P, n, e, uprev, u, abstol, reltol);
}

/* which: 0 Tsit5, 1 Vern7, 2 Vern9. c[18], A[18*18], b[18], bt[18] in double. Returns s. */
int deg_oracle_tableau(int which, double *c, double *A, double *b, double *bt) {
    int s;
    const double *C, *AA, *B, *BT;
    if (which == 0) { s = OT5_S; C = OT5_C; AA = &OT5_A[0][0]; B = OT5_B; BT = OT5_BT; }
    else if (which == 1) { s = OV7_S; C = OV7_C; AA = &OV7_A[0][0]; B = OV7_B; BT = OV7_BT; }
    else if (which == 2) { s = OV9_S; C = OV9_C; AA = &OV9_A[0][0]; B = OV9_B; BT = OV9_BT; }
    else return -1;
    for (int i = 0; i <= OMAXS; ++i) {
        c[i] = 0.0; b[i] = 0.0; bt[i] = 0.0;
        for (int j = 0; j <= OMAXS; ++j) A[i * (OMAXS + 1) + j] = 0.0;
    }
    for (int i = 0; i <= s; ++i) {
        c[i] = C[i]; b[i] = B[i]; bt[i] = BT[i];
        for (int j = 0; j <= s; ++j) A[i * (OMAXS + 1) + j] = AA[i * (s + 1) + j];
    }
    return s;
}

/* Tsit5 dense-output weights b_i(theta), i = 1..7, in Float32 as the solver evaluates them. */
void deg_oracle_tsit5_btheta(float th, float *out7) {
    init_tabs();
    out7[0] = th * fmaf(th, fmaf(th, fmaf(th, T5R[1][4], T5R[1][3]), T5R[1][2]), T5R[1][1]);
    for (int i = 2; i <= 7; ++i) out7[i - 1] = th * (th * fmaf(th, fmaf(th, T5R[i][4], T5R[i][3]), T5R[i][2]));
}
double deg_oracle_tsit5_r(int i, int k) { return OT5_R[i][k]; }

/* Vern9 dense output: extra-stage nodes C[27], coefficients A[27*27] (rows 17..26) and the
 * polynomial coefficients R[27*10], in double; and the Float32 b_i(th) the oracle evaluates. */
void deg_oracle_vern9_dense(double *C, double *A, double *R) {
    for (int i = 0; i < 27; ++i) {
        C[i] = OV9X_C[i];
        for (int j = 0; j < 27; ++j) A[i * 27 + j] = OV9X_A[i][j];
        for (int k = 0; k < 10; ++k) R[i * 10 + k] = OV9R[i][k];
    }
}
void deg_oracle_vern9_btheta(float th, float *out27) {
    init_tabs();
    for (int i = 0; i < 27; ++i) out27[i] = 0.0f;
    for (int m = 0; m < 19; ++m) out27[V9_BSTAGES[m]] = vern9_b(V9_BSTAGES[m], th);
}
