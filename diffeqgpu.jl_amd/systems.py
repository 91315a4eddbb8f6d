"""Right-hand sides of the benchmark and test problems, written for the tracer.

Each is the out-of-place f(u, p, t) of a reference test or benchmark problem with the
same association order (u and p are 0-based here, 1-based in the Julia sources):
  lorenz     test/gpu_kernel_de/gpu_ode_regression.jl:4-12 (jac: test/lower_level_api.jl:18-35)
  decay      test/gpu_kernel_de/stiff_ode/gpu_ode_regression.jl:5-8
  growth     test/public_interface.jl:9-11
  quad       test/gpu_kernel_de/finite_diff.jl:6-9
  robertson  Robertson kinetics (test/gpu_kernel_de/stiff_ode/gpu_ode_mass_matrix.jl, ODE form)
  vdp        Van der Pol
  tdep       u' = p t u
  rot        u1' = -p u2, u2' = p u1
  ball       u1' = u2, u2' = -p1 (test/gpu_kernel_de/gpu_ode_continuous_callbacks.jl:6-10)
  rober_dae  Robertson as a DAE with mass matrix diag(1, 1, 0) and the algebraic row
             y1 + y2 + y3 - 1 (test/gpu_kernel_de/stiff_ode/gpu_ode_mass_matrix.jl:5-27)
  robertson_eye  the ODE form with an explicit identity mass matrix (the M branches of the
             implicit methods)
  dae2       u1' = -0.04 u1 + 1e4 u2, 0 = u1 + u2 - 1, M = diag(1, 0)
             (test/gpu_kernel_de/stiff_ode/gpu_ode_modelingtoolkit_dae.jl:22-44)
and the SDE problems (drift f, diffusion g; SDE_SYSTEMS):
  gbm        f = p1 u, g = p2 u (test/gpu_kernel_de/gpu_sde_convergence.jl:8-10)
  gbm1       f = u, g = u (test/gpu_kernel_de/gpu_sde_regression.jl:13-15)
  lorenz     Lorenz drift, g = (3, 3, 3) (gpu_sde_regression.jl:48-58, parameters as p)
  nondiag    f = 1.01 u, 2 x 4 noise matrix (gpu_sde_regression.jl:102-117; SMatrix{2,4} fills
             its eight arguments column-major, so row 1 is (0.3 u1, 0.9 u1, 1.2 u2, 0.3 u2))
  ou         Ornstein-Uhlenbeck f = p1 (p2 - u), g = p3
  bm         f = 0, g = p1 (scaled Brownian motion)
"""
from .problems import DualODEFunction, ODEFunction, SDEFunction
from .tracer import SVector


def lorenz(u, p, t):
    s, r, b = p[0], p[1], p[2]
    du1 = s * (u[1] - u[0])
    du2 = u[0] * (r - u[2]) - u[1]
    du3 = u[0] * u[1] - b * u[2]
    return SVector(du1, du2, du3)


def lorenz_jac(u, p, t):
    s, r, b = p[0], p[1], p[2]
    return [[-s, s, 0.0], [r - u[2], -1.0, -u[0]], [u[1], u[0], -b]]


def zero_tgrad3(u, p, t):
    return SVector(0.0, 0.0, 0.0)


def zero_tgrad2(u, p, t):
    return SVector(0.0, 0.0)


def decay(u, p, t):
    return SVector(-p[0] * u[0])


def growth(u, p, t):
    return SVector(u[0])


def quad(u, p, t):
    return SVector(-p[0] * u[0] * u[0])


def robertson(u, p, t):
    k1, k2, k3 = p[0], p[1], p[2]
    return SVector(-k1 * u[0] + k3 * u[1] * u[2],
                   k1 * u[0] - k2 * u[1] * u[1] - k3 * u[1] * u[2],
                   k2 * u[1] * u[1])


def robertson_jac(u, p, t):
    k1, k2, k3 = p[0], p[1], p[2]
    return [[-k1, k3 * u[2], k3 * u[1]],
            [k1, -2.0 * k2 * u[1] - k3 * u[2], -k3 * u[1]],
            [0.0, 2.0 * k2 * u[1], 0.0]]


def vdp(u, p, t):
    mu = p[0]
    return SVector(u[1], mu * (1.0 - u[0] * u[0]) * u[1] - u[0])


def vdp_jac(u, p, t):
    mu = p[0]
    return [[0.0, 1.0], [-2.0 * mu * u[0] * u[1] - 1.0, mu * (1.0 - u[0] * u[0])]]


def tdep(u, p, t):
    return SVector(p[0] * t * u[0])


def rot(u, p, t):
    return SVector(-p[0] * u[1], p[0] * u[0])


def ball(u, p, t):
    return SVector(u[1], -p[0])


def ball_jac(u, p, t):
    return [[0.0, 1.0], [0.0, 0.0]]


def rober_dae(u, p, t):
    k1, k2, k3 = p[0], p[1], p[2]
    return SVector(-k1 * u[0] + k3 * u[1] * u[2],
                   k1 * u[0] - k2 * u[1] ** 2 - k3 * u[1] * u[2],
                   u[0] + u[1] + u[2] - 1.0)


def rober_dae_jac(u, p, t):
    k1, k2, k3 = p[0], p[1], p[2]
    return [[-k1, k3 * u[2], k3 * u[1]],
            [k1, -2.0 * k2 * u[1] - k3 * u[2], -k3 * u[1]],
            [1.0, 1.0, 1.0]]


def dae2(u, p, t):
    return SVector(-0.04 * u[0] + 1.0e4 * u[1], u[0] + u[1] - 1.0)


def dae2_jac(u, p, t):
    return [[-0.04, 1.0e4], [1.0, 1.0]]


MASS_DIAG110 = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 0.0))
EYE3 = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))

# (ODEFunction, n, np, oracle system id)
SYSTEMS = {
    "lorenz": (ODEFunction(lorenz, jac=lorenz_jac, tgrad=zero_tgrad3), 3, 3, 0),
    "decay": (ODEFunction(decay), 1, 1, 1),
    "growth": (ODEFunction(growth), 1, 1, 2),
    "quad": (ODEFunction(quad), 1, 1, 3),
    "robertson": (ODEFunction(robertson, jac=robertson_jac, tgrad=zero_tgrad3), 3, 3, 4),
    "vdp": (ODEFunction(vdp, jac=vdp_jac), 2, 1, 5),
    "tdep": (ODEFunction(tdep), 1, 1, 6),
    "rot": (ODEFunction(rot), 2, 1, 7),
    "ball": (ODEFunction(ball, jac=ball_jac), 2, 1, 8),
    "rober_dae": (ODEFunction(rober_dae, jac=rober_dae_jac, tgrad=zero_tgrad3, mass_matrix=MASS_DIAG110), 3, 3, 9),
    "robertson_eye": (ODEFunction(robertson, jac=robertson_jac, tgrad=zero_tgrad3, mass_matrix=EYE3), 3, 3, 10),
    "dae2": (ODEFunction(dae2, jac=dae2_jac, tgrad=zero_tgrad2, mass_matrix=((1.0, 0.0), (0.0, 0.0))), 2, 0, 11),
}
# the systems with a mass matrix, solved by the implicit methods only
MASS_SYSTEMS = ("rober_dae", "robertson_eye", "dae2")

# Lorenz with dual-number state and parameters (6 partials: d/du0 and d/dp), augmented to 21 + 21
# entries; solved by the explicit methods only
SYSTEMS["lorenz_dual"] = (DualODEFunction(ODEFunction(lorenz), 3, 3, 6), 21, 21, 12)
DUAL_SYSTEMS = ("lorenz_dual",)


def gbm_f(u, p, t):
    return SVector(p[0] * u[0])


def gbm_g(u, p, t):
    return SVector(p[1] * u[0])


def gbm1_f(u, p, t):
    return SVector(u[0])


def gbm1_g(u, p, t):
    return SVector(u[0])


def sde_lorenz_g(u, p, t):
    return SVector(3.0, 3.0, 3.0)


def nondiag_f(u, p, t):
    return SVector(1.01 * u[0], 1.01 * u[1])


def nondiag_g(u, p, t):
    du1_1 = 0.3 * u[0]
    du1_2 = 0.6 * u[0]
    du1_3 = 0.9 * u[0]
    du1_4 = 0.12 * u[0]
    du2_1 = 1.2 * u[1]
    du2_2 = 0.2 * u[1]
    du2_3 = 0.3 * u[1]
    du2_4 = 1.8 * u[1]
    # SMatrix{2,4}(du1_1, du1_2, ..., du2_4) in row-major rows
    return [[du1_1, du1_3, du2_1, du2_3], [du1_2, du1_4, du2_2, du2_4]]


def ou_f(u, p, t):
    return SVector(p[0] * (p[1] - u[0]))


def ou_g(u, p, t):
    return SVector(p[2])


def bm_f(u, p, t):
    return SVector(0.0)


def bm_g(u, p, t):
    return SVector(p[0])


# (SDEFunction, n, np, noise columns or None for diagonal noise, oracle system name)
SDE_SYSTEMS = {
    "gbm": (SDEFunction(gbm_f, gbm_g), 1, 2, None, "gbm"),
    "gbm1": (SDEFunction(gbm1_f, gbm1_g), 1, 0, None, "gbm1"),
    "lorenz": (SDEFunction(lorenz, sde_lorenz_g), 3, 3, None, "lorenz"),
    "nondiag": (SDEFunction(nondiag_f, nondiag_g), 2, 0, 4, "nondiag"),
    "ou": (SDEFunction(ou_f, ou_g), 1, 3, None, "ou"),
    "bm": (SDEFunction(bm_f, bm_g), 1, 1, None, "bm"),
}
