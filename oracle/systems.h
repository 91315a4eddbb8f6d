/* Right-hand sides used by the oracle (test infrastructure only).
 *
 * Each system is written out-of-place in Float32 with the association order of the
 * reference test problems, so that the oracle and a kernel built from the same
 * expression tree perform the same IEEE operations:
 *   lorenz    test/gpu_kernel_de/gpu_ode_regression.jl:4-12, jac test/lower_level_api.jl:18-35
 *   decay     test/gpu_kernel_de/stiff_ode/gpu_ode_regression.jl:5-8 (u' = -p u)
 *   growth    test/public_interface.jl:9-11 (u' = u)
 *   quad      test/gpu_kernel_de/finite_diff.jl:6-9 (u' = -p u^2)
 *   robertson test/gpu_kernel_de/stiff_ode/gpu_ode_mass_matrix.jl (ODE form, no mass matrix)
 *   vdp       Van der Pol, u1' = u2, u2' = mu (1 - u1^2) u2 - u1
 *   tdep      u' = p t u (non-autonomous, closed form u0 exp(p t^2 / 2))
 *   rot       u1' = -p u2, u2' = p u1 (closed form rotation)
 *   ball      u1' = u2, u2' = -p (test/gpu_kernel_de/gpu_ode_continuous_callbacks.jl:6-10)
 *   rober_dae Robertson as a DAE, M = diag(1, 1, 0), third equation y1 + y2 + y3 - 1
 *             (test/gpu_kernel_de/stiff_ode/gpu_ode_mass_matrix.jl:5-27)
 *   robertson_eye  the ODE form with an explicit identity mass matrix
 *   dae2      u1' = -0.04 u1 + 1e4 u2, 0 = u1 + u2 - 1, M = diag(1, 0)
 *             (test/gpu_kernel_de/stiff_ode/gpu_ode_modelingtoolkit_dae.jl:22-44)
 * Jacobians are row-major J[r*n + c] = d f_r / d u_c; M is the row-major mass matrix or NULL
 * (no mass matrix: the reference's UniformScaling I).
 */
#ifndef DEG_ORACLE_SYSTEMS_H
#define DEG_ORACLE_SYSTEMS_H

enum { OSYS_LORENZ = 0, OSYS_DECAY = 1, OSYS_GROWTH = 2, OSYS_QUAD = 3, OSYS_ROBERTSON = 4,
       OSYS_VDP = 5, OSYS_TDEP = 6, OSYS_ROT = 7, OSYS_BALL = 8, OSYS_ROBER_DAE = 9, OSYS_ROBERTSON_EYE = 10,
       OSYS_DAE2 = 11, OSYS_LORENZ_DUAL = 12, OSYS_COUNT = 13 };

typedef void (*orhs_fn)(const float *u, const float *p, float t, float *du);
typedef void (*ojac_fn)(const float *u, const float *p, float t, float *J);

typedef struct {
    int n, np;
    orhs_fn f;
    ojac_fn jac;
    orhs_fn tgrad;
    const float *M;
    int fd; /* 1: jac and tgrad by forward differences of f (autodiff = false) */
    int dualnv, dualp; /* dual-number state: dualnv values, each with dualp partials (0: plain) */
    int pivlu;         /* 1: W by LU with partial pivoting for every n (a check of Cramer's rule) */
} osys;

static void lorenz_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    float s = p[0], r = p[1], b = p[2];
    du[0] = s * (u[1] - u[0]);
    du[1] = u[0] * (r - u[2]) - u[1];
    du[2] = u[0] * u[1] - b * u[2];
}
static void lorenz_jac(const float *u, const float *p, float t, float *J) {
    (void)t;
    float s = p[0], r = p[1], b = p[2];
    J[0] = -s;       J[1] = s;     J[2] = 0.0f;
    J[3] = r - u[2]; J[4] = -1.0f; J[5] = -u[0];
    J[6] = u[1];     J[7] = u[0];  J[8] = -b;
}
static void zero_tgrad3(const float *u, const float *p, float t, float *g) {
    (void)u; (void)p; (void)t;
    g[0] = 0.0f; g[1] = 0.0f; g[2] = 0.0f;
}
static void zero_tgrad2(const float *u, const float *p, float t, float *g) {
    (void)u; (void)p; (void)t;
    g[0] = 0.0f; g[1] = 0.0f;
}
static void zero_tgrad1(const float *u, const float *p, float t, float *g) {
    (void)u; (void)p; (void)t;
    g[0] = 0.0f;
}

static void decay_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    du[0] = -p[0] * u[0];
}
static void decay_jac(const float *u, const float *p, float t, float *J) {
    (void)u; (void)t;
    J[0] = -p[0];
}

static void growth_f(const float *u, const float *p, float t, float *du) {
    (void)p; (void)t;
    du[0] = u[0];
}
static void growth_jac(const float *u, const float *p, float t, float *J) {
    (void)u; (void)p; (void)t;
    J[0] = 1.0f;
}

static void quad_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    du[0] = -p[0] * u[0] * u[0];
}
static void quad_jac(const float *u, const float *p, float t, float *J) {
    (void)t;
    J[0] = -2.0f * p[0] * u[0];
}

static void robertson_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    float k1 = p[0], k2 = p[1], k3 = p[2];
    du[0] = -k1 * u[0] + k3 * u[1] * u[2];
    du[1] = k1 * u[0] - k2 * u[1] * u[1] - k3 * u[1] * u[2];
    du[2] = k2 * u[1] * u[1];
}
static void robertson_jac(const float *u, const float *p, float t, float *J) {
    (void)t;
    float k1 = p[0], k2 = p[1], k3 = p[2];
    J[0] = -k1; J[1] = k3 * u[2];                          J[2] = k3 * u[1];
    J[3] = k1;  J[4] = -2.0f * k2 * u[1] - k3 * u[2];      J[5] = -k3 * u[1];
    J[6] = 0.0f; J[7] = 2.0f * k2 * u[1];                  J[8] = 0.0f;
}

static void vdp_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    float mu = p[0];
    du[0] = u[1];
    du[1] = mu * (1.0f - u[0] * u[0]) * u[1] - u[0];
}
static void vdp_jac(const float *u, const float *p, float t, float *J) {
    (void)t;
    float mu = p[0];
    J[0] = 0.0f;                                J[1] = 1.0f;
    J[2] = -2.0f * mu * u[0] * u[1] - 1.0f;     J[3] = mu * (1.0f - u[0] * u[0]);
}

static void tdep_f(const float *u, const float *p, float t, float *du) {
    du[0] = p[0] * t * u[0];
}
static void tdep_jac(const float *u, const float *p, float t, float *J) {
    (void)u;
    J[0] = p[0] * t;
}
static void tdep_tgrad(const float *u, const float *p, float t, float *g) {
    (void)t;
    g[0] = p[0] * u[0];
}

static void rot_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    du[0] = -p[0] * u[1];
    du[1] = p[0] * u[0];
}
static void rot_jac(const float *u, const float *p, float t, float *J) {
    (void)u; (void)t;
    J[0] = 0.0f;  J[1] = -p[0];
    J[2] = p[0];  J[3] = 0.0f;
}

/* falling ball: u1' = u2, u2' = -p1 (test/gpu_kernel_de/gpu_ode_continuous_callbacks.jl:6-10) */
static void ball_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    du[0] = u[1];
    du[1] = -p[0];
}
static void ball_jac(const float *u, const float *p, float t, float *J) {
    (void)u; (void)p; (void)t;
    J[0] = 0.0f;  J[1] = 1.0f;
    J[2] = 0.0f;  J[3] = 0.0f;
}

/* Robertson DAE: y1' = -k1 y1 + k3 y2 y3, y2' = k1 y1 - k2 y2^2 - k3 y2 y3, 0 = y1 + y2 + y3 - 1 */
static void rober_dae_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    float k1 = p[0], k2 = p[1], k3 = p[2];
    du[0] = -k1 * u[0] + k3 * u[1] * u[2];
    du[1] = k1 * u[0] - k2 * (u[1] * u[1]) - k3 * u[1] * u[2];
    du[2] = u[0] + u[1] + u[2] - 1.0f;
}
static void rober_dae_jac(const float *u, const float *p, float t, float *J) {
    (void)t;
    float k1 = p[0], k2 = p[1], k3 = p[2];
    J[0] = -k1;  J[1] = k3 * u[2];                          J[2] = k3 * u[1];
    J[3] = k1;   J[4] = -2.0f * k2 * u[1] - k3 * u[2];      J[5] = -k3 * u[1];
    J[6] = 1.0f; J[7] = 1.0f;                               J[8] = 1.0f;
}
static void dae2_f(const float *u, const float *p, float t, float *du) {
    (void)p; (void)t;
    du[0] = -0.04f * u[0] + 1.0e4f * u[1];
    du[1] = u[0] + u[1] - 1.0f;
}
static void dae2_jac(const float *u, const float *p, float t, float *J) {
    (void)u; (void)p; (void)t;
    J[0] = -0.04f; J[1] = 1.0e4f;
    J[2] = 1.0f;   J[3] = 1.0f;
}
static const float M_DIAG10[4] = {1.0f, 0.0f, 0.0f, 0.0f};
static const float M_DIAG110[9] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
static const float M_EYE3[9] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};

/* Lorenz over dual numbers with 6 partials: state and parameters are the 3 values followed by
 * the partials of each component (u[3 + 6 i + j]). Each operation takes the form the tracer
 * emits for it: x - y -> (xv - yv, xd - yd); x * y -> (xv yv, xd yv + yd xv). */
#define OLD_P 6
static void lorenz_dual_f(const float *u, const float *p, float t, float *du) {
    (void)t;
    const float *u0d = u + 3, *u1d = u + 3 + OLD_P, *u2d = u + 3 + 2 * OLD_P;
    const float *sd = p + 3, *rd = p + 3 + OLD_P, *bd = p + 3 + 2 * OLD_P;
    const float s = p[0], r = p[1], b = p[2];
    /* du1 = s * (u1 - u0) */
    const float d1 = u[1] - u[0];
    du[0] = s * d1;
    for (int j = 0; j < OLD_P; ++j) {
        const float d1d = u1d[j] - u0d[j];
        du[3 + j] = sd[j] * d1 + d1d * s;
    }
    /* du2 = u0 * (r - u2) - u1 */
    const float e = r - u[2];
    const float m = u[0] * e;
    du[1] = m - u[1];
    for (int j = 0; j < OLD_P; ++j) {
        const float ed = rd[j] - u2d[j];
        const float md = u0d[j] * e + ed * u[0];
        du[3 + OLD_P + j] = md - u1d[j];
    }
    /* du3 = u0 * u1 - b * u2 */
    const float a = u[0] * u[1];
    const float bb = b * u[2];
    du[2] = a - bb;
    for (int j = 0; j < OLD_P; ++j) {
        const float ad = u0d[j] * u[1] + u1d[j] * u[0];
        const float bbd = bd[j] * u[2] + u2d[j] * b;
        du[3 + 2 * OLD_P + j] = ad - bbd;
    }
}

static const osys OSYSTEMS[OSYS_COUNT] = {
    {3, 3, lorenz_f, lorenz_jac, zero_tgrad3, NULL, 0, 0, 0},
    {1, 1, decay_f, decay_jac, zero_tgrad1, NULL, 0, 0, 0},
    {1, 1, growth_f, growth_jac, zero_tgrad1, NULL, 0, 0, 0},
    {1, 1, quad_f, quad_jac, zero_tgrad1, NULL, 0, 0, 0},
    {3, 3, robertson_f, robertson_jac, zero_tgrad3, NULL, 0, 0, 0},
    {2, 1, vdp_f, vdp_jac, zero_tgrad2, NULL, 0, 0, 0},
    {1, 1, tdep_f, tdep_jac, tdep_tgrad, NULL, 0, 0, 0},
    {2, 1, rot_f, rot_jac, zero_tgrad2, NULL, 0, 0, 0},
    {2, 1, ball_f, ball_jac, zero_tgrad2, NULL, 0, 0, 0},
    {3, 3, rober_dae_f, rober_dae_jac, zero_tgrad3, M_DIAG110, 0, 0, 0},
    {3, 3, robertson_f, robertson_jac, zero_tgrad3, M_EYE3, 0, 0, 0},
    {2, 0, dae2_f, dae2_jac, zero_tgrad2, M_DIAG10, 0, 0, 0},
    {21, 21, lorenz_dual_f, NULL, NULL, NULL, 0, 3, 6},
};

/* The same right-hand sides in double precision, for the FP64 oracle. */
static void lorenz_fd(const double *u, const double *p, double t, double *du) {
    (void)t;
    du[0] = p[0] * (u[1] - u[0]);
    du[1] = u[0] * (p[1] - u[2]) - u[1];
    du[2] = u[0] * u[1] - p[2] * u[2];
}
static void decay_fd(const double *u, const double *p, double t, double *du) {
    (void)t;
    du[0] = -p[0] * u[0];
}
static void growth_fd(const double *u, const double *p, double t, double *du) {
    (void)p; (void)t;
    du[0] = u[0];
}
static void quad_fd(const double *u, const double *p, double t, double *du) {
    (void)t;
    du[0] = -p[0] * u[0] * u[0];
}
static void robertson_fd(const double *u, const double *p, double t, double *du) {
    (void)t;
    du[0] = -p[0] * u[0] + p[2] * u[1] * u[2];
    du[1] = p[0] * u[0] - p[1] * u[1] * u[1] - p[2] * u[1] * u[2];
    du[2] = p[1] * u[1] * u[1];
}
static void vdp_fd(const double *u, const double *p, double t, double *du) {
    (void)t;
    du[0] = u[1];
    du[1] = p[0] * (1.0 - u[0] * u[0]) * u[1] - u[0];
}
static void tdep_fd(const double *u, const double *p, double t, double *du) {
    du[0] = p[0] * t * u[0];
}
static void rot_fd(const double *u, const double *p, double t, double *du) {
    (void)t;
    du[0] = -p[0] * u[1];
    du[1] = p[0] * u[0];
}
static void ball_fd(const double *u, const double *p, double t, double *du) {
    (void)t;
    du[0] = u[1];
    du[1] = -p[0];
}
typedef void (*orhs_fn_d)(const double *u, const double *p, double t, double *du);
static const orhs_fn_d OSYSTEMS_D[OSYS_COUNT] = {
    lorenz_fd, decay_fd, growth_fd, quad_fd, robertson_fd, vdp_fd, tdep_fd, rot_fd, ball_fd,
    robertson_fd, robertson_fd, /* the DAE's solution is the ODE form's (not integrated in FP64 here) */
    rot_fd,                     /* placeholder: the 2 x 2 DAE has no FP64 ODE form here */
    rot_fd,                     /* placeholder: the dual Lorenz is not integrated in FP64 */
};

#endif
