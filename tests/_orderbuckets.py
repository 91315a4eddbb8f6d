"""The start-order buckets of deg_order_by_span, restated on the host for the order tests: bucket 0
holds zero, negative and NaN keys; buckets 1..2047 split the range [bmin, bmax] of the positive keys'
float bit patterns into ranges of 2^shift patterns, with shift the smallest that fits the range into
2047 buckets."""
import numpy as np


def buckets(keys):
    k = np.ascontiguousarray(np.asarray(keys, np.float32))
    bits = k.view(np.uint32).astype(np.int64)
    pos = k > 0
    out = np.zeros(k.shape, np.int64)
    if not pos.any():
        return out
    bmin, bmax = int(bits[pos].min()), int(bits[pos].max())
    shift = 0
    while ((bmax - bmin) >> shift) > 2046:
        shift += 1
    out[pos] = 1 + np.minimum((bits[pos] - bmin) >> shift, 2046)
    return out


def assert_descending(keys, order):
    """order is a permutation of the trajectories in non-increasing bucket order."""
    order = np.asarray(order)
    assert np.array_equal(np.sort(order), np.arange(len(keys)))
    b = buckets(keys)[order]
    assert np.all(np.diff(b) <= 0)
