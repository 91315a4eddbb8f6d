"""ctypes binding of libdeghip.so (C ABI declared in include/deghip.h).

The library is built in-tree by ``csrc/build.py``; nothing here falls back to another
implementation: if the library is missing, loading fails with an error that says how to
build it.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libdeghip.so")
CACHE_DIR = os.path.join(PKG_DIR, "_jitcache")

DEG_OK = 0
DEG_TSIT5, DEG_VERN7, DEG_RB23, DEG_VERN9, DEG_RODAS4, DEG_RODAS5P, DEG_KVAERNO3, DEG_KVAERNO5 = range(8)
DEG_EM, DEG_SIEA = 8, 9
F_ADAPTIVE = 1 << 0
F_PARITY = 1 << 1
F_DETPOW = 1 << 2
F_KAHAN = 1 << 3
F_PERSISTENT = 1 << 4
F_LEARNED = 1 << 5
F_FASTPOW = 1 << 6
F_STEPLOG = 1 << 7
F_SAVEAT_GLOBAL = 1 << 8
F_WAVES_SHIFT = 9
F_COMPACT = 1 << 12
F_BATCH_SAVES = 1 << 13
F_TSTOPS = 1 << 14
F_CALLBACKS = 1 << 15
F_ORDERED_INPUTS = 1 << 16
SAVEAT_LDS_CAP = 4096

RET_SUCCESS, RET_DTMIN, RET_MAXITERS, RET_NONFINITE, RET_TERMINATED, RET_FAILURE = 0, 1, 2, 3, 4, 5
RETCODE_NAMES = {RET_SUCCESS: "Success", RET_DTMIN: "DtLessThanMin", RET_MAXITERS: "MaxIters",
                 RET_NONFINITE: "Unstable", RET_TERMINATED: "Terminated", RET_FAILURE: "Failure"}

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint32)


class SolveArgs(ctypes.Structure):
    """Mirror of ``deg_solve_args``."""

    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("n", ctypes.c_int32),
        ("n_traj", ctypes.c_int64),
        ("stride", ctypes.c_int64),
        ("id_offset", ctypes.c_int64),
        ("u0", ctypes.c_void_p),
        ("p", ctypes.c_void_p),
        ("t0s", ctypes.c_void_p),
        ("tfs", ctypes.c_void_p),
        ("t0", ctypes.c_float),
        ("tf", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("abstol", ctypes.c_float),
        ("reltol", ctypes.c_float),
        ("saveat", ctypes.c_void_p),
        ("nsaveat", ctypes.c_int32),
        ("nsave", ctypes.c_int32),
        ("save_everystep", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("ts", ctypes.c_void_p),
        ("us", ctypes.c_void_p),
        ("retcode", ctypes.c_void_p),
        ("naccept", ctypes.c_void_p),
        ("nreject", ctypes.c_void_p),
        ("max_attempts", ctypes.c_int64),
        ("stream", ctypes.c_void_p),
        ("grid_blocks", ctypes.c_int32),
        ("refill_batch", ctypes.c_int32),
        ("elapsed_ms", ctypes.c_void_p),
        ("mlp", ctypes.c_void_p),
        ("steplog", ctypes.c_void_p),
        ("steplog_count", ctypes.c_void_p),
        ("steplog_cap", ctypes.c_int64),
        ("chunk", ctypes.c_int32),
        ("mlp_every", ctypes.c_int32),
        ("order", ctypes.c_void_p),
        ("tstops", ctypes.c_void_p),
        ("ntstops", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("seeds", ctypes.c_void_p),
        ("seed", ctypes.c_uint64),
        ("traj_base", ctypes.c_int64),
        ("inputs_ordered", ctypes.c_void_p),
    ]


EXPORTS = [
    "deg_ctx_create", "deg_ctx_destroy", "deg_ctx_num_cus", "deg_set_cache_dir", "deg_kernel_compile",
    "deg_kernel_destroy", "deg_kernel_attr", "deg_jit_precompile", "deg_jit_source", "deg_solve", "deg_sync",
    "deg_last_error", "deg_version", "deg_solve_args_size", "deg_draw_uniform", "deg_graph_begin", "deg_graph_end",
    "deg_graph_launch", "deg_graph_num_nodes", "deg_graph_destroy", "deg_mlp_packed_size", "deg_mlp_partial_cols",
    "deg_mlp_grad", "deg_mlp_adam", "deg_mlp_forward", "deg_mlp_pack", "deg_mlp_infer", "deg_order_by_span",
    "deg_draw_normal", "deg_start_order", "deg_start_order_noalloc", "deg_start_order_model",
]
STEPLOG_COLS = 12
MLP_NPARAM = 1313

_lock = threading.Lock()
_lib = None


class DegError(RuntimeError):
    pass


def _declare(lib):
    vp = ctypes.c_void_p
    lib.deg_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.deg_ctx_destroy.argtypes = [vp]
    lib.deg_ctx_num_cus.argtypes = [vp]
    lib.deg_set_cache_dir.argtypes = [ctypes.c_char_p]
    lib.deg_kernel_compile.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint32, ctypes.POINTER(vp)]
    lib.deg_kernel_destroy.argtypes = [vp]
    lib.deg_kernel_attr.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.deg_jit_precompile.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                       ctypes.c_char_p]
    lib.deg_jit_source.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                   ctypes.c_char_p, ctypes.c_int64]
    lib.deg_jit_source.restype = ctypes.c_int64
    lib.deg_solve.argtypes = [vp, ctypes.POINTER(SolveArgs)]
    lib.deg_sync.argtypes = [vp]
    lib.deg_last_error.restype = ctypes.c_char_p
    lib.deg_version.restype = ctypes.c_char_p
    lib.deg_solve_args_size.restype = ctypes.c_uint32
    lib.deg_draw_uniform.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_float, ctypes.c_float, vp]
    lib.deg_order_by_span.argtypes = [vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_int64, vp, vp, vp]
    lib.deg_start_order.argtypes = [vp, ctypes.POINTER(SolveArgs), vp, vp, vp]
    lib.deg_start_order_noalloc.argtypes = [vp, ctypes.POINTER(SolveArgs), vp]
    lib.deg_start_order_model.argtypes = [vp, ctypes.POINTER(SolveArgs), vp, ctypes.c_int32, vp, vp, vp]
    lib.deg_draw_normal.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32,
                                    ctypes.c_int32, vp]
    lib.deg_graph_begin.argtypes = [vp]
    lib.deg_graph_end.argtypes = [vp, ctypes.POINTER(vp)]
    lib.deg_graph_launch.argtypes = [vp, vp]
    lib.deg_graph_num_nodes.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.deg_graph_destroy.argtypes = [vp]
    f32, i64, i32 = ctypes.c_float, ctypes.c_int64, ctypes.c_int32
    lib.deg_mlp_packed_size.restype = ctypes.c_uint32
    lib.deg_mlp_partial_cols.restype = ctypes.c_int32
    lib.deg_mlp_grad.argtypes = [vp, vp, i64, vp, vp, f32, f32, f32, vp, i32, vp]
    lib.deg_mlp_adam.argtypes = [vp, vp, vp, vp, i32, i32, f32, f32, f32, f32, vp, vp]
    lib.deg_mlp_forward.argtypes = [vp, vp, vp, i64, i64, vp, vp]
    lib.deg_mlp_pack.argtypes = [vp, vp, f32, f32, vp, vp]
    lib.deg_mlp_infer.argtypes = [vp, vp, i64, i64, vp, vp]
    for name in ("deg_ctx_create", "deg_ctx_destroy", "deg_ctx_num_cus", "deg_set_cache_dir", "deg_kernel_compile",
                 "deg_kernel_destroy", "deg_kernel_attr", "deg_jit_precompile", "deg_solve", "deg_sync",
                 "deg_draw_uniform", "deg_graph_begin", "deg_graph_end", "deg_graph_launch", "deg_graph_num_nodes",
                 "deg_graph_destroy", "deg_mlp_grad", "deg_mlp_adam", "deg_mlp_forward", "deg_mlp_pack",
                 "deg_mlp_infer", "deg_order_by_span", "deg_draw_normal", "deg_start_order",
                 "deg_start_order_noalloc", "deg_start_order_model"):
        getattr(lib, name).restype = ctypes.c_int


def lib():
    """Load libdeghip.so once (raises DegError when it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise DegError(f"{LIB_PATH} is missing: build it with `python __graft_entry__.py` "
                               "(or diffeqgpu.jl_amd/csrc/build.py); the HIP path has no fallback")
            L = ctypes.CDLL(LIB_PATH)
            _declare(L)
            if L.deg_solve_args_size() != ctypes.sizeof(SolveArgs):
                raise DegError("deg_solve_args layout mismatch between libdeghip.so and the Python binding")
            os.makedirs(CACHE_DIR, exist_ok=True)
            L.deg_set_cache_dir(CACHE_DIR.encode())
            _lib = L
        return _lib


def check(rc: int, what: str = ""):
    if rc != DEG_OK:
        msg = lib().deg_last_error().decode(errors="replace")
        raise DegError(f"{what} failed ({rc}): {msg}")
    return rc


def precompile(rhs_src: str, n: int, np_: int, method: int, flags: int) -> str:
    key = ctypes.create_string_buffer(32)
    check(lib().deg_jit_precompile(rhs_src.encode(), n, np_, method, flags, key), "deg_jit_precompile")
    return key.value.decode()


def jit_source(rhs_src: str, n: int, np_: int, method: int, flags: int) -> str:
    L = lib()
    size = L.deg_jit_source(rhs_src.encode(), n, np_, method, flags, None, 0)
    buf = ctypes.create_string_buffer(int(size) + 1)
    L.deg_jit_source(rhs_src.encode(), n, np_, method, flags, buf, size + 1)
    return buf.value.decode()


class Context:
    """A device context of libdeghip (one per HIP device)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        check(L.deg_ctx_create(device, ctypes.byref(h)), "deg_ctx_create")
        self.handle = h
        self.device = device
        self.num_cus = L.deg_ctx_num_cus(h)
        self._kernels = {}

    def kernel(self, rhs_src: str, n: int, np_: int, method: int, flags: int):
        key = (rhs_src, n, np_, method, flags)
        k = self._kernels.get(key)
        if k is None:
            h = ctypes.c_void_p()
            check(lib().deg_kernel_compile(self.handle, rhs_src.encode(), n, np_, method, flags, ctypes.byref(h)),
                  "deg_kernel_compile")
            k = Kernel(h, n, np_, method, flags)
            self._kernels[key] = k
        return k

    def sync(self):
        check(lib().deg_sync(self.handle), "deg_sync")


class Kernel:
    def __init__(self, handle, n, np_, method, flags):
        self.handle = handle
        self.n = n
        self.np = np_
        self.method = method
        self.flags = flags

    def attr(self, which: int) -> int:
        v = ctypes.c_int()
        check(lib().deg_kernel_attr(self.handle, which, ctypes.byref(v)), "deg_kernel_attr")
        return v.value

    def solve(self, args: SolveArgs):
        args.struct_size = ctypes.sizeof(SolveArgs)
        check(lib().deg_solve(self.handle, ctypes.byref(args)), "deg_solve")

    def start_order(self, args: SolveArgs, key_ptr: int, order_ptr: int, work_ptr: int):
        """Cost-ordered start (deg_start_order) for the trajectories of ``args``."""
        args.struct_size = ctypes.sizeof(SolveArgs)
        check(lib().deg_start_order(self.handle, ctypes.byref(args), ctypes.c_void_p(key_ptr),
                                    ctypes.c_void_p(order_ptr), ctypes.c_void_p(work_ptr)), "deg_start_order")

    def start_order_noalloc(self, args: SolveArgs, order_ptr: int):
        """deg_start_order_noalloc: the same order with the library's static device scratch."""
        args.struct_size = ctypes.sizeof(SolveArgs)
        check(lib().deg_start_order_noalloc(self.handle, ctypes.byref(args), ctypes.c_void_p(order_ptr)),
              "deg_start_order_noalloc")

    def start_order_model(self, args: SolveArgs, model_ptr: int, nmodel: int, key_ptr: int, order_ptr: int,
                          work_ptr: int):
        """Cost-ordered start by a fitted cost model (deg_start_order_model)."""
        args.struct_size = ctypes.sizeof(SolveArgs)
        check(lib().deg_start_order_model(self.handle, ctypes.byref(args), ctypes.c_void_p(model_ptr), int(nmodel),
                                          ctypes.c_void_p(key_ptr), ctypes.c_void_p(order_ptr),
                                          ctypes.c_void_p(work_ptr)), "deg_start_order_model")


_contexts = {}


def context(device: int = 0) -> Context:
    with _lock:
        c = _contexts.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            _contexts[device] = c
    return c


def draw_uniform(out_ptr: int, n: int, seed: int, rows: int, row: int, lo: float, scale: float, stream_ptr,
                 id_offset: int = 0):
    """Device draws ``out[i] = lo + scale * U(seed, (id_offset + i) * rows + row)`` (rng.uniform_table)."""
    check(lib().deg_draw_uniform(ctypes.c_void_p(out_ptr), n, id_offset, seed & 0xFFFFFFFFFFFFFFFF, rows, row, lo,
                                 scale, ctypes.c_void_p(stream_ptr)), "deg_draw_uniform")


def draw_normal(out_ptr: int, n: int, seed: int, step: int, m: int, stream_ptr, id_offset: int = 0):
    """Device standard normals of the SDE kernels' generator: ``out[c * n + i]`` = normal c of step
    ``step`` of noise stream ``id_offset + i`` under ``seed`` (deg_draw_normal)."""
    check(lib().deg_draw_normal(ctypes.c_void_p(out_ptr), n, id_offset, seed & 0xFFFFFFFFFFFFFFFF, step, m,
                                ctypes.c_void_p(stream_ptr)), "deg_draw_normal")


class Graph:
    """An instantiated HIP graph of the launches made on one stream between ``begin`` and ``end``."""

    def __init__(self, handle):
        self.handle = handle

    @staticmethod
    def begin(stream_ptr):
        check(lib().deg_graph_begin(ctypes.c_void_p(stream_ptr)), "deg_graph_begin")

    @staticmethod
    def end(stream_ptr) -> "Graph":
        h = ctypes.c_void_p()
        check(lib().deg_graph_end(ctypes.c_void_p(stream_ptr), ctypes.byref(h)), "deg_graph_end")
        return Graph(h)

    def launch(self, stream_ptr):
        check(lib().deg_graph_launch(self.handle, ctypes.c_void_p(stream_ptr)), "deg_graph_launch")

    def num_nodes(self) -> int:
        c = ctypes.c_int64()
        check(lib().deg_graph_num_nodes(self.handle, ctypes.byref(c)), "deg_graph_num_nodes")
        return c.value

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.deg_graph_destroy(self.handle)
            self.handle = None


ORDER_WORK_INTS = 4096


def order_by_span(t0s_ptr, tfs_ptr, t0: float, tf: float, n: int, order_ptr: int, work_ptr: int, stream_ptr):
    """Longest-span-first permutation of 0..n-1 into order (int32), see deg_order_by_span."""
    check(lib().deg_order_by_span(ctypes.c_void_p(t0s_ptr) if t0s_ptr else None,
                                  ctypes.c_void_p(tfs_ptr) if tfs_ptr else None, t0, tf, n, ctypes.c_void_p(order_ptr),
                                  ctypes.c_void_p(work_ptr), ctypes.c_void_p(stream_ptr)), "deg_order_by_span")
