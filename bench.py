"""Benchmark: trajectories/s of the ensemble solvers on one MI355X.

Default workload (the headline configuration): 1,000,000 Lorenz trajectories, adaptive
GPUTsit5 abstol = reltol = 1e-6, dt0 = 0.1, tspan (0, 10), saveat 0:1:10 (11 points x 3 states
written SoA), PI controller. A step is one ensemble solve of all trajectories with the inputs
resident in HBM; the step is relaunched into the same output buffers.

Other workloads (``--workload``), measured the same way:
  lorenz_uneven  4,000,000 Lorenz, rho~U(0,60), sigma~U(5,15), beta~U(1,4), tend~U(1,50), GPUVern7
                 abstol = reltol = 1e-8, compensated t, endpoint only. One step is a replay of a
                 HIP graph holding 10 sweeps. Each sweep re-draws (sigma, rho, beta, tend) on the
                 device and solves all 4M trajectories, so a step solves 40M trajectories.
  robertson      1,000,000 Robertson kinetics, k = (0.04, 3e7, 1e4) .* exp(0.3 Z), GPURosenbrock23
                 reltol = 1e-4, abstol = 1e-8, FP64 W factorisation, 20 log-spaced save times.
  vdp            2,000,000 Van der Pol, mu ~ logU(0.1, 30), u0 ~ U(-2, 2)^2, tspan (0, 50),
                 GPUTsit5 abstol = reltol = 1e-6, endpoint only.
  sde_lorenz_em / sde_lorenz_siea
                 1,000,000 trajectories of the reference's Lorenz SDE (drift p = (10, 28, 8/3),
                 diagonal additive noise 3, u0 = (1, 0, 0)), GPUEM / GPUSIEA dt = 1/256 on (0, 10):
                 2560 steps each, endpoint only, seed 123 (Philox4x32-10 increments).

Prints one JSON line. ``roofline`` prices the solver kernel against the FP32 vector peak. The
flop count is algorithmic:
  - Tsit5 attempt: 66n + 18 + 6F; Vern7 attempt: 116n + 24 + 10F;
  - Rosenbrock23 attempt: 25n + 8 + 3F + FJ + FT in FP32, plus W = I - gamma dt J and its LU
    with three solves, n^2 + n + (2/3)n^3 + 6n^2 in FP64;
  - dense output per save point: Tsit5 15n + 40, Rosenbrock23 4n + 8.
F, FJ and FT are the traced RHS, Jacobian and time-gradient flop counts. Attempts come from the
kernel's own per-trajectory accept/reject counters. SDE steps (n states, m increments, G the
diffusion's flops): EM 4n + F + G, SIEA 45n + 2F + 3G, plus 45 flops per Box-Muller pair; the
Philox4x32-10 integer work (100 integer operations per four normals) is reported beside the
flops and not counted in them. The kernel time comes from HIP events
recorded on the launch stream around every timed step.

Besides ``value`` (inputs and outputs resident in HBM) the line reports the end-to-end rate
through the C ABI from host arrays (SoA pack, H2D, solve, D2H of ts/us/retcode) and the cold
hiprtc compile time of the workload's solver; neither is part of ``value``.

Persistent adaptive workloads (not the sweeps) fit the start-order cost model (costmodel.py) on a
200,000-trajectory held-out draw before the timed region; the autotuner then also measures its order
(the shipped table holds the MI355X result for the benchmark workloads).

Before the W warmup steps the step is relaunched, untimed, for --settle-ms (300) of GPU time: the
MI355X's clocks rise over the first ~0.1 s of sustained load, and without this the K timed steps of a
short run would partly measure the ramp (reported as config.clock_settle; --settle-ms 0 turns it off).

Options: --gpus N (1 only: one GPU), --steps K, --warmup W, --workload, --n-traj, --no-autotune,
--copy-probe (one 256 MiB device copy first, to calibrate HBM counters under rocprofv3),
--no-cpu (skip the CPU baseline), --no-extras (skip the end-to-end and compile timings).
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak
METRIC = "trajectories/sec at fixed FP64-oracle error; learned vs PI controller"

WORKLOADS = {
    "lorenz_adaptive": dict(alg="GPUTsit5", steps=20, warmup=3, golden="lorenz_adaptive_vern9.npz", cpu_sample=200_000,
                            desc="lorenz_adaptive_1M: Lorenz, GPUTsit5 adaptive abstol=reltol=1e-6, dt0=0.1, "
                                 "tspan (0,10), saveat 0:1:10 (11 x 3 SoA), PI controller"),
    "lorenz_uneven": dict(alg="GPUVern7", steps=3, warmup=1, golden="lorenz_uneven_vern9.npz", cpu_sample=20_000,
                          sweeps=10,
                          desc="lorenz_uneven_4M: Lorenz rho~U(0,60) sigma~U(5,15) beta~U(1,4) tend~U(1,50), "
                               "GPUVern7 adaptive 1e-8, Kahan t, endpoint only, 10 device-drawn sweeps per "
                               "hipGraph replay, longest span first"),
    "robertson": dict(alg="GPURosenbrock23", steps=10, warmup=2, golden="robertson_radau.npz", cpu_sample=100_000,
                      desc="robertson_1M: Robertson k=(0.04,3e7,1e4).*exp(0.3Z), GPURosenbrock23 reltol=1e-4 "
                           "abstol=1e-8, FP64 W-LU, 20 log-spaced saves"),
    "vdp": dict(alg="GPUTsit5", steps=10, warmup=2, golden="vdp_vern9.npz", cpu_sample=100_000,
                desc="vdp_2M: Van der Pol mu~logU(0.1,30) u0~U(-2,2)^2 tspan (0,50), GPUTsit5 adaptive "
                     "abstol=reltol=1e-6, endpoint only, PI controller"),
}
SDE_WORKLOADS = {
    "sde_lorenz_em": dict(alg="GPUEM", steps=10, warmup=2, cpu_sample=4_000,
                          desc="sde_lorenz_em_1M: Lorenz drift p=(10,28,8/3), diagonal additive noise 3, "
                               "u0=(1,0,0) (gpu_sde_regression.jl:43-58), GPUEM dt=1/256 tspan (0,10) = 2560 "
                               "steps, endpoint only, seed 123"),
    "sde_lorenz_siea": dict(alg="GPUSIEA", steps=10, warmup=2, cpu_sample=2_000,
                            desc="sde_lorenz_siea_1M: Lorenz drift p=(10,28,8/3), diagonal additive noise 3, "
                                 "u0=(1,0,0) (gpu_sde_regression.jl:43-58), GPUSIEA dt=1/256 tspan (0,10) = "
                                 "2560 steps, endpoint only, seed 123"),
}
WORKLOADS.update(SDE_WORKLOADS)
UNEVEN_DRAWS = [("p", 0, 5.0, 10.0), ("p", 1, 0.0, 60.0), ("p", 2, 1.0, 3.0), ("tfs", 0, 1.0, 49.0)]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="lorenz_adaptive", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--warmup", type=int, default=-1)
    ap.add_argument("--n-traj", type=int, default=0)
    ap.add_argument("--powmode", default="fast", choices=["ocml", "fast", "det"])
    ap.add_argument("--contract", default="fast", choices=["fast", "off"])
    ap.add_argument("--persistent", type=int, default=1)
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default launch configuration instead of the tuned one (the tuner's cache or "
                         "the shipped MI355X table; tuned before the timed region when neither has it)")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the end-to-end and JIT-compile timings")
    ap.add_argument("--no-order", action="store_true", help="lorenz_uneven: ascending ids instead of longest first")
    ap.add_argument("--start-order", default="auto", choices=["auto", "cost", "model", "none", "span"],
                    help="persistent grids: the tuned choice (auto), descending span / initial step estimate (cost), "
                         "descending attempt count predicted by the fitted cost model (model), index order (none), "
                         "or (lorenz_uneven) descending span (span)")
    ap.add_argument("--retune", action="store_true", help="measure the launch configuration again")
    ap.add_argument("--no-cost-model", action="store_true",
                    help="do not fit the start-order cost model on a held-out draw (no start_order='model')")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposer leg")
    ap.add_argument("--order-study", action="store_true",
                    help="also time the solve under other start orders (index, span / h0 key, cost model, tiles of "
                         "16 consecutive ids, and the true attempt counts as a bound); adds 'start_orders'")
    ap.add_argument("--mlp-every", type=int, default=2, help="learned leg: evaluate the proposer every k iterations")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--settle-ms", type=float, default=300.0,
                    help="untimed relaunches of the step, before the warmup steps, until this much GPU time has "
                         "passed (MI355X clocks rise over the first ~0.1 s of sustained load); 0: none")
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


def flop_model(method, n, fc):
    """(FP32 flops per attempt, FP64 flops per attempt, flops per save point)."""
    F, FJ, FT = fc["F"], fc["FJ"], fc["FT"]
    if method == "GPUTsit5":
        return 66 * n + 18 + 6 * F, 0, 15 * n + 40
    if method == "GPUVern7":
        return 116 * n + 24 + 10 * F, 0, 0
    return 25 * n + 8 + 3 * F + FJ + FT, n * n + n + (2 * n ** 3) // 3 + 6 * n * n, 4 * n + 8


def cached_code_objects():
    """Number of code objects in the package's kernel cache. The cache ships precompiled; a count that
    grows during a run means a kernel was compiled in this process (reported as
    config.kernels_compiled_at_run). Such kernels have measured up to 8 % slower than the same code
    object loaded from the cache, so a timed run should report 0."""
    from diffeqgpu_amd import _lib

    try:
        return sum(1 for e in os.listdir(_lib.CACHE_DIR) if e.endswith(".co"))
    except OSError:
        return 0


def settle_clocks(torch, run, stream, settle_ms, max_relaunches=2000):
    """Relaunch the step, untimed, until settle_ms of GPU time has passed (at least once when
    settle_ms > 0): the GPU's clocks rise over the first ~0.1 s of sustained load, and the timed steps
    should see the steady state. Returns what was done, for the bench line."""
    if settle_ms <= 0:
        return {"relaunches": 0, "ms": 0.0}
    a_, b_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a_.record(stream)
    run()
    b_.record(stream)
    b_.synchronize()
    n = 1
    elapsed = per = a_.elapsed_time(b_)
    while elapsed < settle_ms and n < max_relaunches:
        # batches of back-to-back relaunches sized to the time left (one synchronisation per batch)
        k = max(1, min(64, int((settle_ms - elapsed) / max(per, 1e-3))))
        for _ in range(k):
            run()
        n += k
        b_.record(stream)
        b_.synchronize()
        elapsed = a_.elapsed_time(b_)
    return {"relaunches": n, "ms": elapsed,
            "what": "untimed relaunches of the same step before the warmup steps, so that the timed steps run at "
                    "settled clocks"}


def _tile_order(tile_key, N, T=16, Q=8):
    """Start order made of tiles of T consecutive trajectory ids, tiles in descending tile_key, dealt
    round-robin over the Q queues so that each claimed chunk of T entries is one whole tile (queue q
    serves order entries q, q + Q, q + 2Q, ...). N must be a multiple of T * Q."""
    rank = np.argsort(-np.asarray(tile_key), kind="stable")
    order = np.empty(N, np.int32)
    j = np.arange(T)
    for r, tile in enumerate(rank):
        q, m = r % Q, r // Q
        order[(T * m + j) * Q + q] = tile * T + j
    return order


def order_study(deg, torch, f, cfg, prob, alg, dev, kw, reps=5, rounds=3):
    """Median kernel time (HIP events, each launch including its ordering kernels) of the same solve
    under different start orders, with the tuned launch, on the workload's first N - N % 128
    trajectories (whole tiles of 16 ids for each of the 8 queues). Every leg is set up first and then
    timed in `rounds` interleaved rounds of `reps` relaunches (median over all), so that clock drift
    over the run does not favour a leg.
      index        ids in order;
      cost_key     descending span / h0 (deg_start_order);
      cost_model   descending predicted attempts (deg_start_order_model), when a model is fitted;
      tiles_by_key tiles of 16 consecutive ids -- what coalesced save-row writes would need -- in
                   descending order of the tile's largest span / h0 key;
      tiles_by_attempts, attempts
                   the same tiles, or single trajectories, in descending order of their true attempt
                   counts from an earlier solve: bounds, since the counts are known only after solving;
      attempts_bucketed, cost_model_exact
                   the true counts through the device's bucketed sort (2047 buckets over the keys'
                   range, no order inside a bucket), and the cost model's keys in exact descending
                   order (both computed once: no ordering kernels in their launches).
    Also the Spearman rank correlation of each key with the attempt counts."""
    from diffeqgpu_amd import _lib

    N = cfg["N"] - cfg["N"] % 128
    u0, p = cfg["u0"][:, :N], cfg["p"][:, :N]
    base = dataclasses.replace(dev, start_order="none")
    st = torch.cuda.current_stream()

    def setup(d, order=None):
        probs = deg.HIPProbVector.from_arrays(f, u0, p, cfg["tspan"], d)
        o = None if order is None else torch.as_tensor(order, device="cuda")
        res = deg.vectorized_asolve(probs, prob, alg, order=o, **kw)
        torch.cuda.synchronize()
        return res

    def spearman(x, y):
        rx, ry = np.argsort(np.argsort(x)), np.argsort(np.argsort(y))
        return float(np.corrcoef(rx, ry)[0, 1])

    out = {"trajectories": N, "rounds": rounds, "relaunches_per_round": reps}
    legs = {}
    legs["index"] = setup(base)
    att = (legs["index"].naccept.to(torch.int64) + legs["index"].nreject.to(torch.int64)).cpu().numpy()
    legs["cost_key"] = setup(dataclasses.replace(base, start_order="cost"))
    key = legs["cost_key"].order_key.cpu().numpy().astype(np.float64)
    out["spearman_cost_key"] = spearman(key, att)
    if dev.cost_model:
        legs["cost_model"] = setup(dataclasses.replace(base, start_order="model"))
        mkey = legs["cost_model"].order_key.cpu().numpy().astype(np.float64)
        out["spearman_cost_model"] = spearman(mkey, att)
        legs["cost_model_exact"] = setup(base, np.argsort(-mkey, kind="stable").astype(np.int32))
    legs["tiles_by_key"] = setup(base, _tile_order(key.reshape(-1, 16).max(axis=1), N))
    legs["tiles_by_attempts"] = setup(base, _tile_order(att.reshape(-1, 16).max(axis=1), N))
    legs["attempts"] = setup(base, np.argsort(-att, kind="stable").astype(np.int32))
    akey = torch.as_tensor(att.astype(np.float32), device="cuda")
    border = torch.empty(N, dtype=torch.int32, device="cuda")
    work = torch.empty(_lib.ORDER_WORK_INTS, dtype=torch.int32, device="cuda")
    _lib.order_by_span(0, akey.data_ptr(), 0.0, 0.0, N, border.data_ptr(), work.data_ptr(), st.cuda_stream)
    torch.cuda.synchronize()
    legs["attempts_bucketed"] = setup(base, border.cpu().numpy())
    ms = {k: [] for k in legs}
    for _ in range(rounds):
        for k, res in legs.items():
            for _ in range(reps):
                a_, b_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a_.record(st)
                res.relaunch()
                b_.record(st)
                b_.synchronize()
                ms[k].append(a_.elapsed_time(b_))
    for k in legs:
        out[k] = float(np.median(ms[k]))
    out["unit"] = "ms (median kernel time)"
    return out


def solve_kw(cfg):
    kw = dict(dt=cfg["dt"], abstol=cfg["abstol"], reltol=cfg["reltol"])
    if cfg.get("saveat") is not None:
        kw["saveat"] = cfg["saveat"]
    return kw


def host_cpu():
    """(threads for the CPU baseline, online CPUs, CPU model name) of the host. The threads are what
    `nproc` reports for this process: the CPUs it may run on, limited by OMP_NUM_THREADS when that is
    set (a shared host gives each GPU job its share that way)."""
    try:
        threads = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        threads = os.cpu_count() or 1
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        omp = 0
    if omp > 0:
        threads = min(threads, omp)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, os.cpu_count() or threads, model


def oracle_fp64_error(name, cfg, pyoracle):
    """The same error statistic for the CPU restatement of the reference's algorithm (Float32, the
    reference's operation sequence with the deterministic controller power) on the first 1,000
    trajectories: the reference's own error at these tolerances."""
    method = {"GPUTsit5": "tsit5", "GPUVern7": "vern7", "GPURosenbrock23": "rb23"}[WORKLOADS[name]["alg"]]
    m = min(cfg["N"], 1000)
    kw = dict(abstol=cfg["abstol"], reltol=cfg["reltol"], kahan=bool(cfg.get("kahan")), detpow=True)
    sv = cfg.get("saveat")
    if sv is not None:
        kw["saveat"] = (np.arange(11, dtype=np.float32) if np.ndim(sv) == 0 else np.asarray(sv, np.float32))
    else:
        kw["save_everystep"] = False
    if cfg.get("tfs") is not None:
        kw["tfs"] = cfg["tfs"][:m]
    r = pyoracle.solve(cfg["system"], method, True, cfg["u0"][:, :m], cfg["p"][:, :m], cfg["tspan"], cfg["dt"], **kw)
    return fp64_error(name, r["us"])


def cpu_baseline(name, cfg, threads, sample):
    """The oracle (CPU Float32 restatement of the same algorithm) on the host cores, same inputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    method = {"GPUTsit5": "tsit5", "GPUVern7": "vern7", "GPURosenbrock23": "rb23"}[WORKLOADS[name]["alg"]]
    n_sample = min(cfg["N"], sample)
    sl = slice(0, n_sample)
    kw = dict(abstol=cfg["abstol"], reltol=cfg["reltol"], nthreads=threads, kahan=bool(cfg.get("kahan")))
    if cfg.get("saveat") is not None:
        sv = cfg["saveat"]
        kw["saveat"] = (np.arange(11, dtype=np.float32) if np.ndim(sv) == 0 else np.asarray(sv, np.float32))
    if cfg.get("tfs") is not None:
        kw["tfs"] = cfg["tfs"][sl]
    args = (cfg["system"], method, True, cfg["u0"][:, sl], cfg["p"][:, sl], cfg["tspan"], cfg["dt"])
    reps = 0
    t0 = time.perf_counter()
    while True:
        pyoracle.solve(*args, **kw)
        reps += 1
        el = time.perf_counter() - t0
        if el > 10.0 or reps >= 400:
            break
    _, nproc, model = host_cpu()
    return dict(value=reps * n_sample / el, unit="trajectories/s", cores=threads, kind="port",
                sample=f"{reps} x {n_sample} trajectories of the same workload (first {n_sample} inputs), "
                       f"oracle/deg_oracle.c Float32 restatement, OpenMP {threads} threads, {el:.1f} s",
                nproc=threads, online_cpus=nproc, cpu_model=model,
                fp64_oracle_error=oracle_fp64_error(name, cfg, pyoracle) if cfg["N"] >= 1000 else None)


def end_to_end(deg, torch, f, cfg, prob, alg, dev, kw, reps=3):
    """Host arrays in, host arrays out: SoA pack + H2D, the solve, D2H of ts/us/retcode (ms)."""
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev, tfs=cfg.get("tfs"))
        res = deg.vectorized_asolve(probs, prob, alg, **kw)
        out = (res.ts.cpu(), res.us_soa.cpu(), res.retcode.cpu())
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
        del probs, res, out
    return best


def jit_compile_seconds(deg, f, n, np_, alg, dev):
    """Cold hiprtc compile of the workload's solver (a private cache directory)."""
    import tempfile

    from diffeqgpu_amd import _lib
    from diffeqgpu_amd.algorithms import GPUODEImplicitAlgorithm

    src, _ = f.hip_source(n, np_, isinstance(alg, GPUODEImplicitAlgorithm))
    L = _lib.lib()
    with tempfile.TemporaryDirectory() as tmp:
        L.deg_set_cache_dir(tmp.encode())
        try:
            t0 = time.perf_counter()
            _lib.precompile(src, n, np_, alg.method, dev.flags(True))
            return time.perf_counter() - t0
        finally:
            L.deg_set_cache_dir(_lib.CACHE_DIR.encode())


def learned_leg(deg, torch, name, f, cfg, prob, alg, dev, kw, every, steps=10, train_n=50_000):
    """The same workload with the learned step-size proposer: logs from a held-out ensemble
    (different seed) -> GPU training -> the learned kernel in its own tuned launch configuration
    (block size, register cap, refill, chunk, start order; measured before the timed runs, as for the
    PI kernel), evaluated every `every` iterations. Reported beside the PI value, with the other
    evaluation intervals measured the same way."""
    import dataclasses

    from diffeqgpu_amd import learned, workloads

    train = workloads.config(name, N=min(cfg["N"], train_n), seed=9001)
    tprobs = deg.HIPProbVector.from_arrays(f, train["u0"], train["p"], train["tspan"], dev)
    t0 = time.perf_counter()
    rows, _ = learned.collect_steplog(tprobs, prob, alg, 12_000_000, **kw)
    expo = learned.ERR_EXPONENT[type(alg)]
    tr = learned.ProposerTrainer(rows, expo=expo, rho=0.2, seed=0)
    tr.train(1500, batch=32768, lr=3e-3)
    packed = tr.pack()
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    base = dataclasses.replace(dev, block_size=None, waves=0, refill=0, chunk=0, blocks_per_cu=0, start_order=None)
    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], base)
    ldev = deg.autotune.tune(probs, prob, alg, mlp=packed, mlp_every=every, **kw)
    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], ldev)
    per_every = {}
    for ev in sorted({every, 1, 2, 4}):
        res = deg.vectorized_asolve(probs, prob, alg, mlp=packed, mlp_every=ev, **kw)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        for _ in range(2):
            res.relaunch()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a_, b_ in evs:
            a_.record(stream)
            res.relaunch()
            b_.record(stream)
        torch.cuda.synchronize()
        k_ms = float(np.median([a_.elapsed_time(b_) for a_, b_ in evs]))
        att = int((res.naccept.to(torch.int64) + res.nreject.to(torch.int64)).sum().item())
        per_every[ev] = dict(value=cfg["N"] / (k_ms / 1e3), kernel_ms=k_ms, attempts=att,
                             rejections=int(res.nreject.sum().item()),
                             retcodes=np.bincount(res.retcode.cpu().numpy(), minlength=4).tolist(),
                             fp64_oracle_error=fp64_error(name, res.us_soa[:, :, :1000].cpu().numpy())
                             if cfg["N"] >= 1000 else None)
        del res
    out = dict(per_every[every])
    out.update(train_rows=int(rows.shape[0]), train_s=train_s, rho=0.2, mlp_every=every,
               launch=dict(block_size=ldev.block_size, waves=ldev.waves, refill=ldev.refill, chunk=ldev.chunk,
                           start_order=ldev.start_order),
               by_mlp_every={str(k): {kk: v[kk] for kk in ("kernel_ms", "attempts", "rejections", "fp64_oracle_error")}
                             for k, v in per_every.items()},
               note="value = trajectories / median kernel time (HIP events), same inputs and tolerances as the PI line")
    return out


def learned_vs_pi(deg, torch, name, every, steps=8, use_cost_model=True):
    """A second learned-vs-PI comparison on its own workload (the 2M Van der Pol ensemble): the PI
    kernel in its tuned configuration and the learned kernel on the same inputs and tolerances,
    with attempts, rejections and the FP64 error of both on the fixture's 1,000 trajectories. With
    use_cost_model both tuners may also choose the start order of a fitted cost model."""
    from diffeqgpu_amd import systems, workloads

    cfg = workloads.config(name)
    f, n, np_, _ = systems.SYSTEMS[cfg["system"]]
    alg = getattr(deg, WORKLOADS[name]["alg"])()
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
    kw = solve_kw(cfg)
    dev = deg.HIPDevice(0, contract="fast", powmode="fast", persistent=True)
    if use_cost_model:
        # both controllers may start in the order of a cost model fitted on a held-out draw
        pilot = workloads.config(name, N=200_000, seed=9001)
        pprobs = deg.HIPProbVector.from_arrays(f, pilot["u0"], pilot["p"], pilot["tspan"], dev)
        dev = dataclasses.replace(dev, cost_model=deg.costmodel.fit(pprobs, prob, alg, **kw))
        del pprobs
    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
    dev = deg.autotune.tune(probs, prob, alg, **kw)
    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
    res = deg.vectorized_asolve(probs, prob, alg, **kw)
    stream = torch.cuda.current_stream()
    for _ in range(2):
        res.relaunch()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a_, b_ in evs:
        a_.record(stream)
        res.relaunch()
        b_.record(stream)
    torch.cuda.synchronize()
    pi_ms = float(np.median([a_.elapsed_time(b_) for a_, b_ in evs]))
    att = int((res.naccept.to(torch.int64) + res.nreject.to(torch.int64)).sum().item())
    pi = dict(value=cfg["N"] / (pi_ms / 1e3), kernel_ms=pi_ms, attempts=att,
              rejections=int(res.nreject.sum().item()),
              retcodes=np.bincount(res.retcode.cpu().numpy(), minlength=4).tolist(),
              fp64_oracle_error=fp64_error(name, res.us_soa[:, :, :1000].cpu().numpy()),
              launch=dict(block_size=dev.block_size, waves=dev.waves, refill=dev.refill, chunk=dev.chunk,
                          start_order=dev.start_order))
    del res, probs
    lrn = learned_leg(deg, torch, name, f, cfg, prob, alg, dev, kw, every, steps=steps, train_n=200_000)
    lrn["speedup_vs_pi"] = pi_ms / lrn["kernel_ms"]
    lrn["attempts_vs_pi"] = lrn["attempts"] / att
    return dict(workload=WORKLOADS[name]["desc"], trajectories=cfg["N"], pi=pi, learned=lrn)


def fp64_error(name, us):
    """Endpoint error (max over states) against the FP64 fixture, first 1,000 trajectories."""
    spec = WORKLOADS[name]
    fx = os.path.join(ROOT, "tests", "golden", spec["golden"])
    if not os.path.exists(fx):
        return None
    ref = np.load(fx)["u"]
    e = np.abs(us[-1].astype(np.float64) - ref[-1]).max(axis=0)
    return dict(median=float(np.median(e)), p99=float(np.percentile(e, 99)), max=float(e.max()), subset=int(e.size),
                reference=f"tests/golden/{spec['golden']} (FP64; "
                          + ("scipy Radau rtol 1e-11" if name == "robertson" else "Vern9 abstol=reltol=1e-12") + ")")


def sde_flop_model(alg, n, m, F, G):
    """(FP32 flops per step, Philox integer ops per step) of one SDE trajectory step."""
    pairs = sum(1 + (4 * b + 2 < m) for b in range((m + 3) // 4))
    bm = 45 * pairs
    if alg == "GPUEM":
        return 4 * n + F + G + bm, 100 * ((m + 3) // 4)
    return 45 * n + 2 * F + 3 * G + bm, 100 * ((m + 3) // 4)


def main_sde(args, name):
    """The SDE workloads: one step is one ensemble solve of all trajectories (inputs in HBM)."""
    import torch

    import __graft_entry__ as entry

    deg = entry.load_package()
    from diffeqgpu_amd import systems, tracer

    co0 = cached_code_objects()
    spec = SDE_WORKLOADS[name]
    steps = args.steps or spec["steps"]
    warmup = spec["warmup"] if args.warmup < 0 else args.warmup
    N = args.n_traj or 1_000_000
    f, n, np_, m_noise, oname = systems.SDE_SYSTEMS["lorenz"]
    m = n if m_noise is None else m_noise
    p0 = np.array([10.0, 28.0, np.float32(8.0) / np.float32(3.0)], np.float32)
    u0 = np.zeros((3, N), np.float32)
    u0[0] = 1.0
    p = np.repeat(p0[:, None], N, axis=1)
    dt, tspan, seed = np.float32(1.0 / 256), (0.0, 10.0), 123
    dev = deg.HIPDevice(0, contract=args.contract if args.contract != "fast" else "off",
                        block_size=args.block_size or None)
    probs = deg.HIPProbVector.from_arrays(f, u0, p, tspan, dev, seed=seed)
    prob = deg.SDEProblem(f, None, u0[:, 0], tspan, p0, seed=seed)
    alg = getattr(deg, spec["alg"])()
    stream = torch.cuda.current_stream(torch.device("cuda", 0))
    res = deg.vectorized_solve(probs, prob, alg, dt=dt, save_everystep=False)
    run = res.relaunch
    settle = settle_clocks(torch, run, stream, args.settle_ms)
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record(stream)
        run()
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in evs]
    nsteps = int(res.naccept.to(torch.int64).sum().item())
    F = tracer.flop_counts(f.f, n, np_)["F"]
    G = f.hip_source(n, np_, m_noise)[1] - F
    fstep, istep = sde_flop_model(spec["alg"], n, m, F, G)
    k_avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = nsteps * fstep / k_avg_s / 1e12
    retcodes = np.bincount(res.retcode.cpu().numpy(), minlength=4).tolist()
    out = {
        "metric": "trajectories/sec (fixed-step SDE ensemble, endpoint output)",
        "value": N / (wall / steps),
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (the reference's SDE test problem; Philox4x32-10 Wiener increments, seed 123)",
        "config": {"workload": spec["desc"], "trajectories_per_step": N, "kernel_flags": res.flags,
                   "block_size": dev.block_size or 256, "clock_settle": settle,
                   "kernels_compiled_at_run": cached_code_objects() - co0},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel_ms": float(np.mean(kernel_ms)),
                     "sde_steps_per_launch": nsteps, "flops_per_step": fstep, "philox_int_ops_per_step": istep,
                     "int_tops": nsteps * istep / k_avg_s / 1e12},
        "retcodes": retcodes,
    }
    # the same solve with FP contraction and the hardware Box-Muller (not bitwise with the oracle)
    fdev = dataclasses.replace(dev, contract="fast")
    fres = deg.vectorized_solve(deg.HIPProbVector.from_arrays(f, u0, p, tspan, fdev, seed=seed), prob, alg, dt=dt,
                                save_everystep=False)
    for _ in range(warmup):
        fres.relaunch()
    torch.cuda.synchronize()
    fevs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a_, b_ in fevs:
        a_.record(stream)
        fres.relaunch()
        b_.record(stream)
    torch.cuda.synchronize()
    fms = float(np.mean([a_.elapsed_time(b_) for a_, b_ in fevs]))
    out["fast_normals"] = {"kernel_ms": fms, "trajectories_per_s": N / (fms / 1e3), "kernel_flags": fres.flags,
                           "what": "contract=fast: FMA contraction and the hardware log2/sqrt/sin/cos Box-Muller"}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        threads = args.cpu_threads or host_cpu()[0]
        ns = min(N, spec["cpu_sample"])
        meth = "em" if spec["alg"] == "GPUEM" else "siea"
        reps, c0 = 0, time.perf_counter()
        while True:
            pyoracle.sde_solve(oname, meth, u0[:, :ns], p[:, :ns], tspan, dt, seed=seed, save_everystep=False,
                               nthreads=threads)
            reps += 1
            el = time.perf_counter() - c0
            if el > 10.0 or reps >= 50:
                break
        _, nproc, model = host_cpu()
        out["cpu_baseline"] = dict(value=reps * ns / el, unit="trajectories/s", cores=threads, kind="port",
                                   sample=f"{reps} x {ns} trajectories of the same workload, oracle/deg_oracle_sde.c "
                                          f"Float32 restatement, OpenMP {threads} threads, {el:.1f} s",
                                   nproc=threads, online_cpus=nproc, cpu_model=model)
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); trajectories are independent and "
              "a multi-GPU run would give each GPU a contiguous range of trajectory ids", file=sys.stderr)
        sys.exit(2)
    import torch

    import __graft_entry__ as entry

    deg = entry.load_package()
    from diffeqgpu_amd import systems, tracer, workloads

    co0 = cached_code_objects()
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    name = args.workload
    if name in SDE_WORKLOADS:
        return main_sde(args, name)
    spec = WORKLOADS[name]
    steps = args.steps or spec["steps"]
    warmup = spec["warmup"] if args.warmup < 0 else args.warmup
    dev_t = torch.device("cuda", 0)
    if args.copy_probe:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev_t)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    cfg = workloads.config(name, N=args.n_traj or None)
    N = cfg["N"]
    f, n, np_, _ = systems.SYSTEMS[cfg["system"]]
    start = {"auto": None, "cost": "cost", "model": "model"}.get(args.start_order, "none")
    if start == "model" and (args.no_cost_model or not args.persistent or name == "lorenz_uneven"):
        print("--start-order model needs the fitted cost model (persistent grids, not --no-cost-model or "
              "lorenz_uneven)", file=sys.stderr)
        sys.exit(2)
    if args.no_order:
        start = "none"
    dev = deg.HIPDevice(0, contract=args.contract, powmode=args.powmode, persistent=bool(args.persistent),
                        block_size=args.block_size or None, refill=args.refill, chunk=args.chunk,
                        kahan=bool(cfg.get("kahan")), start_order=start)
    alg = getattr(deg, spec["alg"])()
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
    kw = solve_kw(cfg)
    cost_model = None
    if not args.no_cost_model and name != "lorenz_uneven" and bool(args.persistent):
        # the start-order cost model, fitted on a held-out draw of the workload (another seed), before
        # the timed region like the tuning; the tuner then also measures start_order="model"
        pilot = workloads.config(name, N=min(N, 200_000), seed=9001)
        t_fit = time.perf_counter()
        pprobs = deg.HIPProbVector.from_arrays(f, pilot["u0"], pilot["p"], pilot["tspan"], dev, tfs=pilot.get("tfs"))
        w = deg.costmodel.fit(pprobs, prob, alg, **kw)
        cost_model = dict(pilot_trajectories=pilot["N"], pilot_seed=9001, fit_s=time.perf_counter() - t_fit,
                          weights=list(w))
        dev = dataclasses.replace(dev, cost_model=w)
        del pprobs
    tuned = None
    if not args.no_autotune and not (args.block_size or args.refill or args.chunk):
        tprobs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev, tfs=cfg.get("tfs"))
        dev = deg.autotune.tune(tprobs, prob, alg, refresh=args.retune, **kw)
        if start is not None or args.start_order == "span":
            dev = dataclasses.replace(dev, start_order=start or "none")  # an explicit choice beats the tuned one
        tuned = dict(block_size=dev.block_size, waves=dev.waves, refill=dev.refill, chunk=dev.chunk,
                     blocks_per_cu=dev.blocks_per_cu, start_order=dev.start_order)
        del tprobs

    if name == "lorenz_uneven":
        sweeps = spec["sweeps"]
        probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev, tfs=cfg["tfs"])
        sw = deg.ParameterSweep(probs, prob, alg, sweeps=sweeps, seed=303, draws=UNEVEN_DRAWS,
                                longest_first=args.start_order == "span" and not args.no_order, **kw)
        stream = sw.stream
        run = sw.replay
        per_step = sweeps * N
        results = sw.results
    else:
        probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
        stream = torch.cuda.current_stream(dev_t)
        # The first call JIT-loads the kernel, allocates the outputs and runs one solve; every later
        # step relaunches the same solve into the same output buffers (all outputs are recomputed).
        res = deg.vectorized_asolve(probs, prob, alg, **kw)
        run = res.relaunch
        per_step = N
        results = [res]

    co_timed = cached_code_objects()
    settle = settle_clocks(torch, run, stream, args.settle_ms)
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record(stream)
        run()
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in evs]
    ms_per_step = wall * 1e3 / steps
    value = per_step / (wall / steps)

    # work actually done per step (identical every step: the solves are deterministic)
    fc = tracer.flop_counts(f.f, n, np_, f.jac, f.tgrad)
    fa32, fa64, fsave = flop_model(spec["alg"], n, fc)
    attempts = sum(int((r.naccept.to(torch.int64) + r.nreject.to(torch.int64)).sum().item()) for r in results)
    # save points after t0 are dense-output evaluations (a save point at t0 copies u0)
    nsave_interp = sum(N * int((r.saveat > float(np.float32(cfg["tspan"][0]))).sum().item())
                       for r in results if r.saveat is not None)
    flops = attempts * (fa32 + fa64) + nsave_interp * fsave
    k_avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = flops / k_avg_s / 1e12
    retcodes = np.bincount(np.concatenate([r.retcode.cpu().numpy() for r in results]), minlength=4).tolist()

    err = None
    if N >= 1000:
        err = fp64_error(name, results[0].us_soa[:, :, :1000].cpu().numpy())

    traffic = load_traffic({"lorenz_adaptive": "lorenz_adaptive_tsit5_traffic.json"}.get(name, ""))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if spec["alg"] != "GPURosenbrock23" else "f32 (W-LU f64)",
        "data": "synthetic (seeded SplitMix64 draws of the workload's parameter law)",
        "config": {"workload": spec["desc"], "trajectories_per_step": per_step, "controller": "PI",
                   "kernel_flags": results[0].flags, "contract": args.contract, "powmode": args.powmode,
                   "persistent": bool(args.persistent), "block_size": dev.block_size or 256,
                   "waves": dev.waves, "refill": dev.refill or 4, "chunk": dev.chunk or 16, "autotuned": tuned,
                   "start_order": ("span" if name == "lorenz_uneven" and args.start_order == "span" and not args.no_order
                                   else dev.start_order),
                   "cost_model": cost_model, "clock_settle": settle,
                   "kernels_compiled_at_run": {"before_timed_steps": co_timed - co0}},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel_ms": float(np.mean(kernel_ms)),
                     "kernel_ms_min_median_max": [float(np.min(kernel_ms)), float(np.median(kernel_ms)),
                                                  float(np.max(kernel_ms))],
                     "flops_per_step": flops,
                     "attempts_per_step": attempts, "saves_per_step": nsave_interp,
                     "flops_per_attempt": {"fp32": fa32, "fp64": fa64}, "rhs_flops": fc},
        "fp64_oracle_error": err,
        "retcodes": retcodes,
    }
    if name != "lorenz_uneven" and not args.no_extras:
        e2e_ms = end_to_end(deg, torch, f, cfg, prob, alg, dev, kw)
        out["end_to_end"] = {"ms": e2e_ms, "trajectories_per_s": N / (e2e_ms / 1e3),
                             "what": "host SoA arrays -> H2D -> solve -> D2H of ts, us, retcode (PCIe-inclusive)"}
    if not args.no_extras:
        out["jit_compile_s"] = jit_compile_seconds(deg, f, n, np_, alg, dev)
    if args.order_study and name != "lorenz_uneven" and bool(args.persistent):
        out["start_orders"] = order_study(deg, torch, f, cfg, prob, alg, dev, kw)
    if not args.no_learned and spec["alg"] in ("GPUTsit5", "GPUVern7") and name != "lorenz_uneven":
        # learned vs PI on the same GPU, inputs and tolerances; the PI line above is `value`
        lrn = learned_leg(deg, torch, name, f, cfg, prob, alg, dev, kw, args.mlp_every)
        pi_ms = float(np.mean(kernel_ms))
        lrn["speedup_vs_pi"] = pi_ms / lrn["kernel_ms"]
        lrn["attempts_vs_pi"] = lrn["attempts"] / attempts
        out["learned"] = lrn
    if not args.no_learned and name == "lorenz_adaptive":
        out["learned_c5"] = learned_vs_pi(deg, torch, "vdp", args.mlp_every,
                                          use_cost_model=not args.no_cost_model and bool(args.persistent))
    if not args.no_cpu:
        threads = args.cpu_threads or host_cpu()[0]
        out["cpu_baseline"] = cpu_baseline(name, cfg, threads, spec["cpu_sample"])
        if err is not None and out["cpu_baseline"].get("fp64_oracle_error"):
            ref_err = out["cpu_baseline"]["fp64_oracle_error"]
            out["fp64_oracle_error"]["reference_algorithm"] = {
                "median": ref_err["median"], "p99": ref_err["p99"], "max": ref_err["max"],
                "what": "the CPU restatement of the reference's Float32 algorithm on the same 1,000 trajectories"}
    out["config"]["kernels_compiled_at_run"]["whole_run"] = cached_code_objects() - co0
    print(json.dumps(out))


if __name__ == "__main__":
    main()
