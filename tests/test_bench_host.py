"""Host logic of bench.py that runs without a GPU: the tile start order of the order study and the
one-GPU contract."""
import importlib.util
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_tile_order_serves_whole_tiles_in_key_order():
    """Queue q serves order entries q, q + 8, ...; each chunk of 16 it claims must be one tile of 16
    consecutive ids, and the first chunks of the 8 queues the 8 most expensive tiles."""
    b = _bench()
    N = 128 * 300
    key = np.random.default_rng(0).random(N // 16)
    order = b._tile_order(key, N)
    assert order.dtype == np.int32 and np.array_equal(np.sort(order), np.arange(N))
    rank = np.argsort(-key, kind="stable")
    for q in range(8):
        chunks = order[q::8].reshape(-1, 16)
        assert np.all(chunks[:, 0] % 16 == 0) and np.all(np.diff(chunks, axis=1) == 1)
        # queue q's m-th chunk is the tile of rank 8 m + q
        np.testing.assert_array_equal(chunks[:, 0] // 16, rank[q::8])


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr
