"""Warm the code-object cache (diffeqgpu.jl_amd/_jitcache) with the solver variants the test
suite and the benchmark launch, so a GPU run does not have to JIT them. hiprtc compiles
gfx950 code objects without a GPU."""
from __future__ import annotations

import multiprocessing as mp

from . import _lib
from .algorithms import HIPDevice
from .problems import ODEFunction
from .systems import DUAL_SYSTEMS, SDE_SYSTEMS, SYSTEMS

_METHODS = (_lib.DEG_TSIT5, _lib.DEG_VERN7, _lib.DEG_RB23)
_METHODS_MORE = (_lib.DEG_VERN9, _lib.DEG_RODAS4, _lib.DEG_RODAS5P, _lib.DEG_KVAERNO3, _lib.DEG_KVAERNO5)
_IMPLICIT = (_lib.DEG_RB23, _lib.DEG_RODAS4, _lib.DEG_RODAS5P, _lib.DEG_KVAERNO3, _lib.DEG_KVAERNO5)


def variants():
    devs = [
        (HIPDevice(contract="off"), False),                                   # fixed step, parity
        (HIPDevice(contract="off", powmode="det", persistent=False), True),   # adaptive, bitwise vs oracle
        (HIPDevice(contract="off", powmode="det", persistent=True), True),
        (HIPDevice(contract="off", powmode="ocml", persistent=True), True),   # adaptive, exact ops, device powf
        (HIPDevice(powmode="ocml"), True),
        (HIPDevice(), True),                                                  # adaptive defaults (persistent)
        (HIPDevice(persistent=False), True),
        (HIPDevice(powmode="fast"), True),
        (HIPDevice(contract="off", powmode="det", persistent=True, kahan=True), True),
        (HIPDevice(kahan=True), True),
    ]
    out = []
    for name, (f, n, np_, _sid) in SYSTEMS.items():
        if f.mass_matrix is not None or name in DUAL_SYSTEMS:
            continue  # the mass-matrix and dual systems are warmed below with their methods
        for m in _METHODS:
            src, _ = f.hip_source(n, np_, m == _lib.DEG_RB23)
            for dev, adaptive in devs:
                out.append((name, src, n, np_, m, dev.flags(adaptive)))
            if name in ("lorenz", "vdp", "robertson"):
                # register caps the autotuner tries on the benchmark systems, tail compaction
                for w in (0, 4, 5, 6, 7):
                    for kahan in ((False, True) if m == _lib.DEG_VERN7 else (False,)):
                        if w:
                            out.append((name, src, n, np_, m, HIPDevice(waves=w, kahan=kahan).flags(True)))
                        out.append((name, src, n, np_, m, HIPDevice(waves=w, kahan=kahan, compact=True).flags(True)))
            if m == _lib.DEG_TSIT5 and n <= 4:
                for dev in (HIPDevice(batch_saves=True), HIPDevice(batch_saves=True, contract="off", powmode="det"),
                            HIPDevice(batch_saves=True, persistent=False)):
                    out.append((name, src, n, np_, m, dev.flags(True)))
            if m != _lib.DEG_RB23:
                # the learned proposer and the step-log writer (default adaptive flags); on the benchmark
                # systems also with the register caps the learned kernel's tuner tries and the tuned PI
                # launches that write the training logs
                base = HIPDevice().flags(True)
                for extra in (_lib.F_LEARNED, _lib.F_STEPLOG, _lib.F_LEARNED | _lib.F_STEPLOG):
                    out.append((name, src, n, np_, m, base | extra))
                if m == _lib.DEG_TSIT5 and name in ("lorenz", "vdp"):
                    for w in (4, 5, 6, 7):
                        for extra in (_lib.F_LEARNED, _lib.F_STEPLOG):
                            out.append((name, src, n, np_, m, HIPDevice(waves=w).flags(True) | extra))
    # the other methods in the variants the parity tests launch (fixed step, bitwise adaptive)
    more = [(HIPDevice(contract="off"), False),
            (HIPDevice(contract="off", powmode="det", persistent=False), True),
            (HIPDevice(contract="off", powmode="det", persistent=True), True)]
    for name, (f, n, np_, _sid) in SYSTEMS.items():
        if name in DUAL_SYSTEMS:
            # dual-number states: the explicit methods, parity and default variants
            src, _ = f.hip_source(n, np_, False)
            for m in (_lib.DEG_TSIT5, _lib.DEG_VERN7, _lib.DEG_VERN9):
                for dev, adaptive in more + [(HIPDevice(), True)]:
                    out.append((name, src, n, np_, m, dev.flags(adaptive)))
            # the implicit methods: the dual Jacobian by forward differences (autodiff = false) and by
            # forward mode (fixed step)
            fd_src, _ = f.hip_source(n, np_, True, fd=True)
            ad_src, _ = f.hip_source(n, np_, True, fd=False)
            for m in _IMPLICIT:
                for dev, adaptive in more + [(HIPDevice(), True)]:
                    out.append((name, fd_src, n, np_, m, dev.flags(adaptive)))
                out.append((name, ad_src, n, np_, m, HIPDevice(contract="off").flags(False)))
            continue
        for m in (_IMPLICIT if f.mass_matrix is not None else _METHODS_MORE):
            src, _ = f.hip_source(n, np_, m in _IMPLICIT)
            # the default path of the implicit methods on the stiff fixture
            for dev, adaptive in more + ([(HIPDevice(), True)] if name == "robertson" else []):
                out.append((name, src, n, np_, m, dev.flags(adaptive)))
    # autodiff = false: differenced Jacobian and time gradient (no user jac / tgrad)
    for name in ("vdp", "tdep", "lorenz", "quad", "dae2"):
        f0, n, np_, _sid = SYSTEMS[name]
        f = ODEFunction(f0.f, mass_matrix=f0.mass_matrix)
        src, _ = f.hip_source(n, np_, True, fd=True)
        for m in _IMPLICIT:
            for dev, adaptive in more + ([(HIPDevice(), True)] if name == "quad" else []):
                out.append((name, src, n, np_, m, dev.flags(adaptive)))
    for name, (f, n, np_, nm, _o) in SDE_SYSTEMS.items():
        src, _ = f.hip_source(n, np_, nm)
        for m in ((_lib.DEG_EM,) if nm is not None else (_lib.DEG_EM, _lib.DEG_SIEA)):
            for dev in (HIPDevice(), HIPDevice(contract="fast")):
                out.append((name, src, n, np_, m, dev.sde_flags()))
    return out


def _one(j):
    name, src, n, np_, m, flags = j
    return _lib.precompile(src, n, np_, m, flags)


def run(workers: int = 8, verbose: bool = False):
    """hiprtc serialises compiles inside one process, so the warm-up forks workers."""
    jobs = variants()
    with mp.get_context("fork").Pool(workers) as pool:
        keys = pool.map(_one, jobs, chunksize=4)
    if verbose:
        print(f"{len(keys)} solver variants in the code-object cache")
    return keys


if __name__ == "__main__":
    run(verbose=True)
