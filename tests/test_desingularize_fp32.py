"""The substep kernels form desingularize's ``a + 0.5*eps`` in float32, where
the reference evaluates that one addition in double (``0.5*eps`` is a double
literal, gpu_ocean/SWESimulators/gpu_kernels/common.cu:822-824) and rounds to
float.  This checks, for every float32 h the kernels can pass (the
desingularised branch has cutoff < h < eps, the friction term any |h| < eps),
that both forms give the same bits: exhaustively for h in [2^-17, 0.1) and on
a sample of smaller h, for the default eps = 0.1 and two other values."""

import numpy as np
import pytest


def _both(h, eps):
    eps = np.float32(eps)
    a = (h * h) / (np.float32(2.0) * eps)
    ref = (a.astype(np.float64) + 0.5 * np.float64(eps)).astype(np.float32)
    f32 = a + np.float32(0.5) * eps
    return ref, f32


@pytest.mark.parametrize("eps", [0.1, 0.05, 0.3])
def test_float32_sum_equals_double_then_round(eps):
    lo = np.float32(2.0 ** -17).view(np.uint32)
    hi = np.float32(eps).view(np.uint32)
    step = 1 << 23
    for start in range(int(lo), int(hi), step):
        bits = np.arange(start, min(start + step, int(hi)), dtype=np.uint32)
        h = bits.view(np.float32)
        ref, f32 = _both(h, eps)
        assert np.array_equal(ref.view(np.uint32), f32.view(np.uint32))
        ref, f32 = _both(-h, eps)
        assert np.array_equal(ref.view(np.uint32), f32.view(np.uint32))
    rng = np.random.default_rng(3)
    h = (rng.random(1 << 20) * 2.0 ** -17).astype(np.float32)
    h = np.concatenate([h, np.float32([0.0, 1e-30, 1e-20, 1e-10, 1e-6])])
    ref, f32 = _both(h, eps)
    assert np.array_equal(ref.view(np.uint32), f32.view(np.uint32))
