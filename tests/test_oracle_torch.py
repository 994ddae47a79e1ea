"""Pins of the differentiable torch restatement (the reference gradient of the
HIP adjoint): it must equal the NumPy oracle's forward, and its autograd
gradient must equal central finite differences of the NumPy oracle in float64.
"""

import contextlib

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


# ---------------------------------------------------------------------------
# float32 fidelity and branch pinning
# ---------------------------------------------------------------------------

def _t32(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("rk", [0, 1])
def test_torch_float32_substep_bit_equals_numpy_oracle(bc, kind, rk):
    """In float32 the torch restatement rounds like the oracle (constants in
    float32, the double-literal addition of desingularize in double, correctly
    rounded square roots), so the branch decisions it records are the
    oracle's."""
    p, Hi, Hm, eta, hu, hv = H.random_case(23, 19, bc, 7, **KINDS[kind])
    Hm32 = O.init_bm(Hi)
    q = [eta, hu, hv]
    O.periodic_bc(*q, p)
    qn = [(0.9 * eta).astype(np.float32), (1.1 * hu).astype(np.float32), hv.copy()]
    dt = np.float32(0.7 * O.max_dt(eta, hu, hv, Hm32, p))
    a = O.cdklm_substep(dt, rk, *q, *[x.copy() for x in qn], Hi, Hm32, p, wind_t=0.0)
    b = T.substep(float(dt), rk, *map(_t32, q), *map(_t32, qn), _t32(Hi), _t32(Hm32), p)
    for x, y in zip(a, b):
        assert y.dtype == torch.float32
        assert np.array_equal(np.asarray(x)[2:-2, 2:-2], y.numpy()[2:-2, 2:-2])


def _lake_with_eddy(bc, seed):
    """Water at rest on the west half (every slope candidate zero: minmod,
    upwind and wave-speed ties everywhere) and a random state on the east."""
    p, Hi, Hm, eta, hu, hv = H.random_case(16, 13, bc, seed)
    w = (p.nx + 4) // 2
    for a in (eta, hu, hv):
        a[:, :w] = 0.0
    return p, Hi, Hm, eta, hu, hv


def _pinned_case(bc, seed):
    p, Hi, Hm, eta, hu, hv = _lake_with_eddy(bc, seed)
    Hm32 = O.init_bm(Hi)
    q32 = [eta, hu, hv]
    qn32 = [(0.5 * eta).astype(np.float32), hu.copy(), (0.5 * hv).astype(np.float32)]
    dt = float(np.float32(0.5 * O.max_dt(eta, hu, hv, Hm32, p)))
    pins = T.BranchPins()
    with torch.no_grad(), T.pinned(pins):
        T.substep(dt, 1, *map(_t32, q32), *map(_t32, qn32), _t32(Hi), _t32(Hm32), p)
    pins.replay = True
    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    return p, dt, pins, [a.astype(np.float64) for a in q32], [a.astype(np.float64) for a in qn32], Hi64, Hm64


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2)])
def test_pinned_float64_gradient_is_the_pinned_linearisation(bc):
    """Decisions recorded in float32 and replayed in float64 make the substep
    one smooth piece: its autograd gradient equals central differences of the
    same pinned map (eps = 1e-6) to 1e-8, also on water at rest, where every
    limiter, upwind and wave-speed choice is a tie.  Unpinned, the central
    difference there is the nonlinear one-sided mixture (minmod of the
    directional slopes) and autograd's tie convention differs from it: the
    check below shows the pinning is what makes the two agree."""
    p, dt, pins, q, qn, Hi64, Hm64 = _pinned_case(bc, 31)
    rng = np.random.default_rng(32)
    W = [_t(rng.standard_normal(q[0].shape)) for _ in range(3)]
    worst_free = 0.0
    for k in range(3):
        D = [_t(rng.standard_normal(q[0].shape)) for _ in range(6)]

        def J(s, pinned):
            args = [_t(a) + s * d for a, d in zip(q + qn, D)]
            ctx = T.pinned(pins) if pinned else contextlib.nullcontext()
            with ctx:
                out = T.substep(dt, 1, *args, _t(Hi64), _t(Hm64), p)
            return sum((w[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for w, o in zip(W, out))

        for pinned in (True, False):
            x = [_t(a).requires_grad_(True) for a in q + qn]
            ctx = T.pinned(pins) if pinned else contextlib.nullcontext()
            with ctx:
                out = T.substep(dt, 1, *x, _t(Hi64), _t(Hm64), p)
            L = sum((w[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for w, o in zip(W, out))
            gr = torch.autograd.grad(L, x, allow_unused=True)
            ad = sum(float((g * d).sum()) for g, d in zip(gr, D) if g is not None)
            scale = sum(float((g * d).abs().sum()) for g, d in zip(gr, D) if g is not None)
            with torch.no_grad():
                fd = float(J(1e-6, pinned) - J(-1e-6, pinned)) / 2e-6
            if pinned:
                assert abs(ad - fd) <= 1e-8 * scale, (k, ad, fd, scale)
            else:
                worst_free = max(worst_free, abs(ad - fd) / scale)
    assert worst_free > 1e-4, worst_free


def test_pinning_changes_nothing_away_from_ties():
    """Decisions recorded in float64 at a random state (no ties) and replayed
    at that state give the unpinned value and gradient."""
    p, Hi, Hm, eta, hu, hv = H.random_case(16, 13, (2, 1, 2, 1), 5)
    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    q = [a.astype(np.float64) for a in (eta, hu, hv)]
    rng = np.random.default_rng(6)
    W = [_t(rng.standard_normal(q[0].shape)) for _ in range(3)]
    pins = T.BranchPins()
    with torch.no_grad(), T.pinned(pins):
        T.substep(1.0, 0, *map(_t, q), *map(_t, q), _t(Hi64), _t(Hm64), p)
    pins.replay = True
    res = []
    for ctx in (T.pinned(pins), contextlib.nullcontext()):
        x = [_t(a).requires_grad_(True) for a in q]
        with ctx:
            out = T.substep(1.0, 0, *x, *x, _t(Hi64), _t(Hm64), p)
        L = sum((w[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for w, o in zip(W, out))
        res.append((float(L.detach()), torch.autograd.grad(L, x)))
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1], res[1][1]):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-14)


def test_replay_must_match_the_recording():
    p, dt, pins, q, qn, Hi64, Hm64 = _pinned_case((1, 1, 1, 1), 31)
    p2, Hi2, Hm2, e2, u2, v2 = H.random_case(14, 11, (1, 1, 1, 1), 3)
    with pytest.raises(RuntimeError, match="recorded shape"):
        with T.pinned(pins):
            T.substep(1.0, 1, *map(_t32, (e2, u2, v2)), *map(_t32, (e2, u2, v2)), _t32(Hi2),
                      _t32(O.init_bm(Hi2)), p2)
    short = T.BranchPins()
    short.conds = pins.conds[:5]
    short.replay = True
    with pytest.raises(RuntimeError, match="more decisions"):
        with T.pinned(short):
            T.substep(dt, 1, *map(_t, q), *map(_t, qn), _t(Hi64), _t(Hm64), p)
    extra = T.BranchPins()
    extra.conds = pins.conds + pins.conds[:1]
    extra.replay = True
    with pytest.raises(RuntimeError, match="replayed"):
        with T.pinned(extra):
            T.substep(dt, 1, *map(_t, q), *map(_t, qn), _t(Hi64), _t(Hm64), p)
