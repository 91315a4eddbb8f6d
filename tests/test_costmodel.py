"""The start-order cost model (diffeqgpu.jl_amd/costmodel.py) on the host: the features, the fit, the
device knob. The device key and the solve in model order are tested in test_gpu_costmodel.py."""
import numpy as np
import pytest

F32 = np.float32


def test_features_match_their_definition(deg):
    from diffeqgpu_amd import costmodel

    key = np.array([100.0, 3.0])
    u0 = np.array([[1.0, -2.0], [0.0, 0.5]])
    p = np.array([[10.0, 0.25]])
    X = costmodel.features(key, u0, p, np.array([10.0, 4.0]))
    tiny = 2.0 ** -20
    want = np.array([[1.0, np.log2(100.0), np.log2(1.0 + tiny), np.log2(tiny), np.log2(10.0 + tiny), np.log2(10.0)],
                     [1.0, np.log2(3.0), np.log2(2.0 + tiny), np.log2(0.5 + tiny), np.log2(0.25 + tiny), 2.0]])
    np.testing.assert_allclose(X, want, rtol=1e-15)


def test_fit_recovers_a_log_linear_law(deg):
    """log2 attempts = 3 + 0.1 log2 key + 0.5 log2 p0 - 0.25 log2 p2 is recovered; features that do not
    vary over the pilot (here u0 and the span) get weight 0 (the intercept absorbs them)."""
    from diffeqgpu_amd import costmodel

    r = np.random.default_rng(0)
    N = 200_000   # as large as a pilot: the constant columns' standard deviation is then 4e-12, not 0
    key = 10 ** r.uniform(2, 4, N)
    u0 = np.zeros((3, N))
    u0[0] = 1.0
    p = r.uniform(0.1, 10.0, (3, N))
    span = np.full(N, 10.0)
    y = 3 + 0.1 * np.log2(key) + 0.5 * np.log2(p[0] + 2.0 ** -20) - 0.25 * np.log2(p[2] + 2.0 ** -20)
    w = costmodel.fit_arrays(key, u0, p, span, 2.0 ** y)
    np.testing.assert_allclose(w, [3.0, 0.1, 0.0, 0.0, 0.0, 0.5, 0.0, -0.25, 0.0], atol=2e-3)
    # a wrong feature (a transposed or dropped column) would not fit: the residual is what pins it
    X = costmodel.features(key, u0, p, span)
    assert np.abs(X @ np.asarray(w) - y).max() < 1e-3


def test_model_order_needs_weights(deg):
    assert deg.HIPDevice(start_order="model", cost_model=(1.0,) * 9).cost_order(True)
    with pytest.raises(ValueError, match="cost_model"):
        deg.HIPDevice(start_order="model").cost_order(True)
    with pytest.raises(ValueError):
        deg.HIPDevice(start_order="nope").cost_order(True)
    assert not deg.HIPDevice(start_order="model", cost_model=(1.0,) * 9).cost_order(False)
