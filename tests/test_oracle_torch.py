"""Pins of the differentiable torch restatement (the reference gradient of the
HIP adjoint): it must equal the NumPy oracle's forward, and its autograd
gradient must equal central finite differences of the NumPy oracle in float64.
"""

import numpy as np
import pytest
import torch

from oracle import cdklm16_np as O
from oracle import cdklm16_torch as T
import helpers as H

BCS = [(1, 1, 1, 1), (2, 2, 2, 2), (1, 2, 1, 2), (2, 1, 2, 1)]
KINDS = {"plain": {}, "shallow": {"shallow": True}, "land": {"land": True}}


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))


def _np_rk2(p, Hi64, Hm64, q, dt, n):
    """n fixed-dt RK2 steps of the NumPy oracle in float64."""
    e, u, v = (np.array(a, dtype=np.float64) for a in q)
    for _ in range(n):
        O.periodic_bc(e, u, v, p)
        s = [a.copy() for a in (e, u, v)]
        s = O.cdklm_substep(dt, 0, e, u, v, *s, Hi64, Hm64, p, wind_t=0.0, dtype=np.float64)
        s = [np.array(a) for a in s]
        O.periodic_bc(*s, p)
        out = [a.copy() for a in (e, u, v)]
        e, u, v = (np.array(a) for a in O.cdklm_substep(dt, 1, *s, *out, Hi64, Hm64, p,
                                                           wind_t=0.0, dtype=np.float64))
    return e, u, v


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("rk", [0, 1])
def test_torch_substep_equals_numpy_oracle(bc, kind, rk):
    p, Hi, Hm, eta, hu, hv = H.random_case(21, 17, bc, 3, **KINDS[kind])
    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    q = [a.astype(np.float64) for a in (eta, hu, hv)]
    qn = [0.9 * q[0], 1.1 * q[1], 0.7 * q[2]]
    a = O.cdklm_substep(0.5, rk, *q, *qn, Hi64, Hm64, p, wind_t=0.0, dtype=np.float64)
    b = T.substep(0.5, rk, *map(_t, q), *map(_t, qn), _t(Hi64), _t(Hm64), p)
    for x, y in zip(a, b):
        np.testing.assert_allclose(y.numpy()[2:-2, 2:-2], x[2:-2, 2:-2], rtol=0, atol=1e-11)


def test_torch_dry_clamp_matches_numpy_oracle():
    """A step long enough to dry cells of the shallow strip: both restatements
    clamp the same cells."""
    p, Hi, Hm, eta, hu, hv = H.random_case(21, 17, (1, 1, 1, 1), 4, shallow=True)
    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    q = [a.astype(np.float64) for a in (eta, hu, hv)]
    q[1][:, :8] += 0.05        # push water out of the strip
    dt = 40.0
    a = O.cdklm_substep(dt, 0, *q, *q, Hi64, Hm64, p, wind_t=0.0, dtype=np.float64)
    clamped = (a[0] + Hm64)[2:-2, 2:-2] <= p.depth_cutoff * 1.0000001
    assert clamped.sum() > 0
    b = T.substep(dt, 0, *map(_t, q), *map(_t, q), _t(Hi64), _t(Hm64), p)
    for x, y in zip(a, b):
        np.testing.assert_allclose(y.numpy()[2:-2, 2:-2], x[2:-2, 2:-2], rtol=0, atol=1e-10)


FD_CASES = [((1, 1, 1, 1), "plain", 11), ((2, 2, 2, 2), "plain", 12),
            ((1, 2, 1, 2), "land", 13), ((2, 1, 2, 1), "shallow", 14)]


@pytest.mark.parametrize("bc,kind,seed", FD_CASES)
def test_torch_gradient_equals_numpy_finite_differences(bc, kind, seed):
    """d/d(eps) of a weighted sum of the state after 3 RK2 steps, along a
    random direction: torch autograd vs central differences of the NumPy
    oracle, both float64, relative error <= 1e-4.

    eps = 1e-6: the scheme is piecewise smooth (minmod limiters, wave-speed
    extrema, upwinding), and a perturbation of 1e-3 per cell crosses some of
    those switches, which makes central differences wrong by O(eps) (percent
    level here); at 1e-6 truncation and round-off are both below 1e-8."""
    nx, ny, n, dt = 14, 11, 3, 2.0
    p, Hi, Hm, eta, hu, hv = H.random_case(nx, ny, bc, seed, **KINDS[kind])
    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    rng = np.random.default_rng(100 + seed)
    q0 = [a.astype(np.float64) for a in (eta, hu, hv)]
    W = [np.zeros_like(q0[0]) for _ in range(3)]
    D = [np.zeros_like(q0[0]) for _ in range(3)]
    for k in range(3):
        W[k][2:-2, 2:-2] = rng.standard_normal((ny, nx))
        D[k][2:-2, 2:-2] = rng.standard_normal((ny, nx))
    if kind == "shallow":      # keep perturbations out of the near-dry strip
        for k in range(3):
            D[k][:, : (nx + 4) // 3 + 2] = 0.0

    def loss_np(q):
        out = _np_rk2(p, Hi64, Hm64, q, dt, n)
        return sum(float(np.sum(w[2:-2, 2:-2] * o[2:-2, 2:-2])) for w, o in zip(W, out))

    eps = 1e-6
    lp = loss_np([a + eps * d for a, d in zip(q0, D)])
    lm = loss_np([a - eps * d for a, d in zip(q0, D)])
    fd = (lp - lm) / (2 * eps)

    qt = [_t(a).requires_grad_(True) for a in q0]
    out = T.rk2_steps(*qt, _t(Hi64), _t(Hm64), p, dt, n)
    L = sum((_t(w)[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for w, o in zip(W, out))
    L.backward()
    ad = sum(float((q.grad * _t(d)).sum()) for q, d in zip(qt, D))
    assert abs(ad - fd) <= 1e-4 * abs(fd), (ad, fd)
