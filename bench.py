"""Benchmark: trajectories/s of the adaptive GPUTsit5 ensemble on the 1M-trajectory Lorenz
workload (abstol = reltol = 1e-6, saveat 0:1:10, 11 points x 3 states written SoA), one
MI355X.

A step is one ensemble solve of all 1,000,000 trajectories with the inputs resident in
HBM. Prints one JSON line (see the keys below). The roofline entry prices the solver kernel
against the FP32 vector peak with the algorithmic flop count 264 per Tsit5 attempt
(66n + 18 + 6F, n = 3, F = 8) plus 85 per interpolated save point (15n + 40), using the
kernel's own per-trajectory accept/reject counters; the kernel time comes from HIP events
recorded on the launch stream around every timed launch.

Options: --gpus N (only 1: one GPU), --steps K, --warmup W, --n-traj, --device-flags,
--copy-probe (one 256 MiB device copy first, to calibrate HBM counters under rocprofv3),
--no-cpu (skip the CPU baseline).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (includes dual-issue/packed FP32)
FLOPS_PER_ATTEMPT = 264    # Tsit5 adaptive attempt, Lorenz (n = 3, F = 8)
FLOPS_PER_SAVE = 85        # Tsit5 dense-output point, n = 3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-traj", type=int, default=1_000_000)
    ap.add_argument("--powmode", default="fast", choices=["ocml", "fast", "det"])
    ap.add_argument("--contract", default="fast", choices=["fast", "off"])
    ap.add_argument("--persistent", type=int, default=1)
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline(cfg, threads):
    """The oracle (CPU restatement of the same algorithm) on the host cores, same inputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    sv = np.arange(11, dtype=np.float32)
    n_sample = min(cfg["N"], 200_000)
    u0, p = cfg["u0"][:, :n_sample], cfg["p"][:, :n_sample]
    pyoracle.solve("lorenz", "tsit5", True, u0[:, :1000], p[:, :1000], (0.0, 10.0), 0.1, abstol=1e-6, reltol=1e-6,
                   saveat=sv, nthreads=threads)
    reps = 0
    t0 = time.perf_counter()
    while True:
        pyoracle.solve("lorenz", "tsit5", True, u0, p, (0.0, 10.0), 0.1, abstol=1e-6, reltol=1e-6, saveat=sv,
                       nthreads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el > 10.0 or reps >= 50:
            break
    return dict(value=reps * n_sample / el, unit="trajectories/s", cores=threads, kind="port",
                sample=f"{reps} x {n_sample} trajectories of the same workload (first {n_sample} inputs), "
                       f"oracle/deg_oracle.c Float32 restatement, OpenMP {threads} threads, {el:.1f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); trajectories are independent and "
              "the multi-GPU split is a static shard of trajectory ids (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    import __graft_entry__ as entry

    deg = entry.load_package()
    from diffeqgpu_amd import systems, workloads

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    dev_t = torch.device("cuda", 0)
    if args.copy_probe:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev_t)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    cfg = workloads.config("lorenz_adaptive", N=args.n_traj)
    N = cfg["N"]
    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    dev = deg.HIPDevice(0, contract=args.contract, powmode=args.powmode, persistent=bool(args.persistent),
                        block_size=args.block_size or None, refill=args.refill)
    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
    alg = deg.GPUTsit5()
    kw = dict(dt=0.1, abstol=1e-6, reltol=1e-6, saveat=1.0)

    # The first call JIT-loads the kernel, allocates the outputs and runs one solve; every later
    # step relaunches the same solve into the same output buffers (all outputs are recomputed).
    res = deg.vectorized_asolve(probs, prob, alg, **kw)
    for _ in range(args.warmup):
        res.relaunch()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev_t)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        evs[k][0].record(stream)
        res.relaunch()
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in evs]
    ms_per_step = wall * 1e3 / args.steps
    value = N / (wall / args.steps)

    # work actually done (identical every step: the solve is deterministic)
    attempts = int((res.naccept.to(torch.int64) + res.nreject.to(torch.int64)).sum().item())
    saves_interp = N * (res.ts.shape[0] - 1)
    flops = attempts * FLOPS_PER_ATTEMPT + saves_interp * FLOPS_PER_SAVE
    k_avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = flops / k_avg_s / 1e12
    retcodes = np.bincount(res.retcode.cpu().numpy(), minlength=4).tolist()

    # FP64-oracle endpoint error on the first 1,000 trajectories
    err = None
    fx = os.path.join(ROOT, "tests", "golden", "lorenz_adaptive_vern9.npz")
    if os.path.exists(fx) and N >= 1000:
        ref = np.load(fx)["u"]
        got = res.us_soa[:, :, :1000].cpu().numpy().astype(np.float64)
        e = np.abs(got[-1] - ref[-1]).max(axis=0)
        err = dict(median=float(np.median(e)), p99=float(np.percentile(e, 99)), max=float(e.max()),
                   subset=1000, reference="FP64 Vern9 abstol=reltol=1e-12 (tests/golden/lorenz_adaptive_vern9.npz)")

    traffic = load_traffic("lorenz_adaptive_tsit5_traffic.json")
    out = {
        "metric": "trajectories/sec at fixed FP64-oracle error; learned vs PI controller",
        "value": value,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 parameter draws p = [10, 28, 8/3] .* U^3)",
        "config": {"workload": "lorenz_adaptive_1M: Lorenz, GPUTsit5 adaptive abstol=reltol=1e-6, dt0=0.1, "
                               "tspan (0,10), saveat 0:1:10 (11 x 3 SoA), PI controller",
                   "trajectories": N, "controller": "PI", "kernel_flags": res.flags,
                   "contract": args.contract, "powmode": args.powmode, "persistent": bool(args.persistent),
                   "block_size": args.block_size or 256},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel_ms": float(np.mean(kernel_ms)), "flops_per_launch": flops,
                     "attempts_per_launch": attempts, "saves_per_launch": saves_interp},
        "fp64_oracle_error": err,
        "retcodes": retcodes,
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(cfg, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
