"""Host-side tests that need no GPU: the C ABI library (loads, exports every declared symbol,
struct layout, hiprtc JIT of every method), the RHS tracer, the saveat / range
normalisation of the front end and the API surface."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

F32 = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------------------------------
# C ABI
# ---------------------------------------------------------------------------------------
def declared_functions():
    text = open(os.path.join(ROOT, "include", "deghip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(deg_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol(deg):
    from diffeqgpu_amd import _lib

    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.EXPORTS, name
    assert L.deg_version().decode().startswith("deghip")
    assert L.deg_solve_args_size() == ctypes.sizeof(_lib.SolveArgs)


def test_solve_args_offsets_match_header(deg):
    from diffeqgpu_amd import _lib

    # pointer fields sit at 8-byte boundaries; a few fixed offsets of the C layout
    off = {f[0]: getattr(_lib.SolveArgs, f[0]).offset for f in _lib.SolveArgs._fields_}
    assert off["n_traj"] == 8 and off["u0"] == 32 and off["t0"] == 64 and off["saveat"] == 88
    assert off["ts"] == 112 and off["stream"] == 160 and off["steplog_cap"] == 208 and off["chunk"] == 216
    assert off["order"] == 224 and off["tstops"] == 232 and off["ntstops"] == 240
    assert off["seeds"] == 248 and off["seed"] == 256 and off["traj_base"] == 264
    assert off["inputs_ordered"] == 272
    assert ctypes.sizeof(_lib.SolveArgs) == 280


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("flags", [0, 2, 1, 1 | 2 | 4, 1 | 16, 1 | 64, 1 | 8 | 16])
def test_jit_compiles_every_method(deg, method, flags):
    from diffeqgpu_amd import _lib, systems

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    src, _ = f.hip_source(n, np_, method in (2, 4, 5, 6, 7))
    key = _lib.precompile(src, n, np_, method, flags)
    assert re.fullmatch(r"[0-9a-f]{16}", key)
    assert os.path.exists(os.path.join(_lib.CACHE_DIR, key + ".co"))


def test_jit_source_and_errors(deg):
    from diffeqgpu_amd import _lib, systems

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    src, _ = f.hip_source(n, np_, False)
    full = _lib.jit_source(src, n, np_, 0, 1 | 16)
    assert "#define DEG_N 3" in full and "deg_tsit5_adaptive_persistent" in full and "deg_rhs" in full
    with pytest.raises(_lib.DegError, match="state dimension"):
        _lib.precompile(src, 0, np_, 0, 1)
    with pytest.raises(_lib.DegError, match="exclude"):
        _lib.precompile(src, n, np_, 0, 1 | 4 | 64)
    with pytest.raises(_lib.DegError, match="hiprtc compile"):
        _lib.precompile("this is not C", n, np_, 0, 1)


# ---------------------------------------------------------------------------------------
# tracer
# ---------------------------------------------------------------------------------------
def test_tracer_keeps_association_order(deg):
    from diffeqgpu_amd import systems, tracer

    src, flops = tracer.emit_rhs(systems.lorenz, 3, 3)
    assert flops == 8  # F = 8 for Lorenz
    lines = [l.strip() for l in src.splitlines() if "=" in l]
    assert "const float v7 = u[1] - u[0];" in lines[0]
    assert any("p[1] - u[2]" in l for l in lines)
    # (a + b) + c is not re-associated
    src2, _ = tracer.emit_rhs(lambda u, p, t: [u[0] + u[1] + p[0], u[0]], 2, 1)
    assert "u[0] + u[1]" in src2 and "+ p[0]" in src2


def _pow64(x, y):
    """powf in float64: a negative base with a non-integer exponent is NaN, as in C."""
    if x < 0 and y != math.floor(y):
        return math.nan
    try:
        return float(x) ** float(y)
    except ZeroDivisionError:
        return math.inf


def _log64(x):
    """log in float64 with C's special values (log 0 = -inf, log of a negative number = NaN)."""
    if x == 0:
        return -math.inf
    if x < 0 or x != x:
        return math.nan
    return math.log(x)


def _eval_nodes(tr, outs, u, p, t):
    """Interpret a trace in float64 (host check of emitted expressions)."""
    vals = {}
    for x in tr.g.nodes:
        op = x.op
        if op == "const":
            v = x.val
        elif op == "u":
            v = u[x.val]
        elif op == "p":
            v = p[x.val]
        elif op == "t":
            v = t
        else:
            a = [vals[y.idx] for y in x.args]
            v = {"+": lambda: a[0] + a[1], "-": lambda: a[0] - a[1], "*": lambda: a[0] * a[1],
                 "/": lambda: a[0] / a[1], "neg": lambda: -a[0], "sin": lambda: math.sin(a[0]),
                 "cos": lambda: math.cos(a[0]), "exp": lambda: math.exp(a[0]), "log": lambda: _log64(a[0]),
                 "sqrt": lambda: math.sqrt(a[0]), "tanh": lambda: math.tanh(a[0]), "abs": lambda: abs(a[0]),
                 "pow": lambda: _pow64(a[0], a[1]),
                 "select": lambda: a[2] if eval(f"{a[0]} {x.val} {a[1]}") else a[3],
                 "cmp": lambda: bool(eval(f"{a[0]} {x.val} {a[1]}")),
                 "and": lambda: a[0] and a[1], "or": lambda: a[0] or a[1], "not": lambda: not a[0],
                 "bconst": lambda: bool(x.val), "bselect": lambda: a[1] if a[0] else a[2],
                 "sbsel": lambda: a[1] if math.copysign(1.0, a[0]) < 0 else a[2]}[op]()
        vals[x.idx] = v
    return [vals[o.idx] for o in outs]


@pytest.mark.parametrize("name", ["lorenz", "robertson", "vdp", "tdep", "rot", "quad"])
def test_tracer_jacobian_matches_analytic_and_fd(deg, name):
    from diffeqgpu_amd import systems, tracer

    f, n, np_, _ = systems.SYSTEMS[name]
    tr = tracer.trace(f.f, n, np_, nout=n)
    J, T = tracer.jacobian_nodes(tr)
    rng = np.random.default_rng(0)
    u = rng.uniform(0.2, 1.5, n)
    p = rng.uniform(0.5, 2.0, np_)
    t = 0.7
    Jv = np.array(_eval_nodes(tr, [x for row in J for x in row], u, p, t)).reshape(n, n)
    Tv = np.array(_eval_nodes(tr, T, u, p, t))
    # central differences
    h = 1e-6
    Jfd = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = h
        Jfd[:, c] = (np.array(tracer.evaluate(f.f, u + e, p, t)) - np.array(tracer.evaluate(f.f, u - e, p, t))) / (2 * h)
    Tfd = (np.array(tracer.evaluate(f.f, u, p, t + h)) - np.array(tracer.evaluate(f.f, u, p, t - h))) / (2 * h)
    np.testing.assert_allclose(Jv, Jfd, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(Tv, Tfd, rtol=1e-6, atol=1e-6)
    if f.jac is not None:
        Ja = np.array([float(x) for x in np.asarray(f.jac(list(u), list(p), t), dtype=object).reshape(-1)])
        np.testing.assert_allclose(Jv.reshape(-1), Ja, rtol=1e-12, atol=1e-12)


def test_tracer_functions_and_errors(deg):
    from diffeqgpu_amd import tracer

    m = tracer

    def f(u, p, t):
        return [m.sin(u[0]) * m.exp(-t) + m.sqrt(p[0]) - u[0] ** 3 + m.ifelse(u[0], ">", 0.0, u[0], -u[0])]

    src, _ = tracer.emit_rhs(f, 1, 1, with_jac=True)
    assert "sinf" in src and "expf" in src and "__builtin_sqrtf" in src and "deg_jac" in src and "?" in src
    with pytest.raises(tracer.TraceError):
        tracer.emit_rhs(lambda u, p, t: [u[0] if u[0] > 0 else 0.0], 1, 0)
    with pytest.raises(tracer.TraceError):
        tracer.emit_rhs(lambda u, p, t: [u[0], u[0]], 1, 0)


# ---------------------------------------------------------------------------------------
# ranges and saveat normalisation (lowerlevel_solve.jl:64-109, :271-324; conversions.jl KATs)
# ---------------------------------------------------------------------------------------
def test_julia_float_range_lengths(deg):
    from diffeqgpu_amd.ranges import colon, julia_range_length

    assert julia_range_length(0.0, 0.1, 10.0) == 101
    assert julia_range_length(0.0, 0.01, 10.0) == 1001
    assert julia_range_length(0.0, 0.1, 0.2) == 3
    assert julia_range_length(1.0, 0.1, 10.0) == 91
    assert julia_range_length(0.0, 0.3, 1.0) == 4
    r = colon(0.0, 0.1, 10.0).collect()
    assert r[0] == 0 and r[-1] == F32(10.0) and r[3] == F32(0.3)
    # elements are the correctly rounded k/10, not k * 0.1f0
    np.testing.assert_array_equal(r, np.array([F32(k / 10) for k in range(101)], F32))


def test_saveat_conversions_kats(deg):
    """test/gpu_kernel_de/conversions.jl:23-59 and public_interface.jl:35-41."""
    from diffeqgpu_amd.lowerlevel import normalize_saveat_asolve, normalize_saveat_solve
    from diffeqgpu_amd.ranges import colon

    tspan = (0.0, 10.0)
    np.testing.assert_array_equal(normalize_saveat_asolve(range(1, 11), tspan), np.arange(1, 11, dtype=F32))
    np.testing.assert_array_equal(normalize_saveat_asolve(colon(1.0, 0.1, 10.0, np.float64), tspan),
                                  colon(1.0, 0.1, 10.0).collect())
    np.testing.assert_array_equal(normalize_saveat_asolve(1.0, tspan), np.arange(11, dtype=F32))
    np.testing.assert_array_equal(normalize_saveat_asolve([1.0, 5.0, 10.0], tspan), [1.0, 5.0, 10.0])
    np.testing.assert_array_equal(normalize_saveat_asolve(0.0, tspan), [0.0, 10.0])
    sv = normalize_saveat_asolve(0.1, (0.0, 0.2))
    np.testing.assert_array_equal(sv, np.array([0.0, F32(0.1), F32(0.2)], F32))
    with pytest.raises(ValueError, match="too many save points"):
        normalize_saveat_asolve(1e-5, (0.0, 10.0))
    # the fixed-step entry point re-ranges a range through range(first, last, length)
    np.testing.assert_array_equal(normalize_saveat_solve(range(1, 11), tspan), np.arange(1, 11, dtype=F32))
    np.testing.assert_array_equal(normalize_saveat_solve(2.5, tspan), [0.0, 2.5, 5.0, 7.5, 10.0])


# ---------------------------------------------------------------------------------------
# API surface (public_interface.jl:18-24; algorithms)
# ---------------------------------------------------------------------------------------
def test_api_surface(deg):
    from diffeqgpu_amd import systems

    prob = deg.ODEProblem(systems.growth, [1.0], (0.0, 0.2), [1.0])
    cp = deg.make_prob_compatible(prob)
    assert isinstance(cp, deg.ImmutableODEProblem)
    assert isinstance(deg.EnsembleGPUKernel(deg.HIPDevice()), deg.EnsembleKernelAlgorithm)
    assert isinstance(deg.GPUTsit5(), deg.GPUODEAlgorithm)
    assert [deg.alg_order(a) for a in (deg.GPUTsit5(), deg.GPUVern7(), deg.GPURosenbrock23())] == [5, 7, 2]
    with pytest.raises(NotImplementedError):
        deg.EnsembleGPUKernel(deg.HIPDevice(), 0.5)
    assert deg.GPURosenbrock23(autodiff=False).autodiff is False and deg.GPURodas5P().autodiff is True
    assert deg.GPUKvaerno3(autodiff=False) != deg.GPUKvaerno3()
    r = deg.remake(prob, p=[2.0])
    assert r.p[0] == 2.0 and r.u0[0] == 1.0
    # device flags
    from diffeqgpu_amd import _lib

    assert deg.HIPDevice().flags(True) == _lib.F_ADAPTIVE | _lib.F_PERSISTENT | _lib.F_FASTPOW
    assert deg.HIPDevice().flags(False) == _lib.F_PARITY
    assert deg.HIPDevice(contract="off", powmode="det", persistent=False).flags(True) == 1 | 2 | 4
    with pytest.raises(ValueError):
        deg.HIPDevice(powmode="bad").flags(True)


def test_pack_host_soa(deg):
    from diffeqgpu_amd import problems, systems

    probs = [deg.ODEProblem(systems.lorenz, [1.0, 0.0, float(i)], (0.0, 1.0 + i), [10.0, 28.0, 2.0]) for i in range(4)]
    u0, p, t0s, tfs, tspan = problems.pack_host([deg.make_prob_compatible(x) for x in probs])
    assert u0.shape == (3, 4) and p.shape == (3, 4)
    np.testing.assert_array_equal(u0[2], [0, 1, 2, 3])
    np.testing.assert_array_equal(tfs, [1, 2, 3, 4])
    assert t0s is not None


def test_rng_and_workloads(deg):
    from diffeqgpu_amd import rng, workloads

    a = rng.uniform_table(7, 3, 1000)
    b = rng.uniform_table(7, 3, 1000)
    np.testing.assert_array_equal(a, b)
    assert a.dtype == np.float32 and a.min() >= 0 and a.max() < 1
    assert abs(a.mean() - 0.5) < 0.02
    # counter-based: a prefix of a larger table is the smaller table
    np.testing.assert_array_equal(rng.uniform_table(7, 3, 10), a[:, :10])
    c = workloads.config("lorenz_adaptive", N=100)
    assert c["u0"].shape == (3, 100) and c["p"].shape == (3, 100)
    assert np.all(c["p"] <= workloads.LORENZ_P[:, None])
    for name in ("lorenz_fixed", "lorenz_uneven", "robertson", "vdp"):
        c = workloads.config(name, N=64)
        assert c["u0"].shape[1] == 64


def test_autotune_cache_key_and_lookup(deg, tmp_path, monkeypatch):
    """The tuner's cache key separates configurations and a cached entry is returned as is."""
    import dataclasses

    from diffeqgpu_amd import autotune, systems

    monkeypatch.setattr(autotune, "CACHE_PATH", str(tmp_path / "autotune.json"))
    monkeypatch.setattr(autotune, "SHIPPED_PATH", str(tmp_path / "shipped.json"))
    f, n, np_, _ = systems.SYSTEMS["lorenz"]

    class Probs:
        pass

    probs = Probs()
    probs.f, probs.n, probs.np, probs.N, probs.dev = f, n, np_, 1_000_000, deg.HIPDevice()
    k1 = autotune._key(probs, deg.GPUTsit5(), True, "saveat")
    k2 = autotune._key(probs, deg.GPUVern7(), True, "saveat")
    probs2 = Probs()
    probs2.__dict__.update(probs.__dict__, dev=deg.HIPDevice(kahan=True))
    k3 = autotune._key(probs2, deg.GPUTsit5(), True, "saveat")
    probs3 = Probs()
    probs3.__dict__.update(probs.__dict__, dev=deg.HIPDevice(block_size=64, waves=5))  # tuned knobs: same key
    assert len({k1, k2, k3}) == 3 and autotune._key(probs3, deg.GPUTsit5(), True, "saveat") == k1
    autotune._store({k1: dict(config=dict(block_size=128, waves=6, refill=2, chunk=64, blocks_per_cu=3),
                              kernel_ms=1.0)})
    got = autotune.tune(probs, None, deg.GPUTsit5(), saveat=1.0)
    assert got == dataclasses.replace(deg.HIPDevice(), block_size=128, waves=6, refill=2, chunk=64, blocks_per_cu=3)
    # the shipped table is consulted when the local cache has no entry
    import json

    with open(tmp_path / "shipped.json", "w") as fh:
        json.dump({k2: dict(config=dict(block_size=64, waves=0, refill=1, chunk=4, blocks_per_cu=0))}, fh)
    assert autotune.tune(probs, None, deg.GPUVern7(), saveat=1.0).block_size == 64
    assert deg.HIPDevice(waves=6).flags(True) >> 9 == 6
    with pytest.raises(ValueError):
        deg.HIPDevice(waves=9).flags(True)
