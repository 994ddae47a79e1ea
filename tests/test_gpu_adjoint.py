"""The hand-written HIP adjoint of the fixed-dt RK2 step, called through the C
ABI and the autograd wrapper.

References: torch autograd of the float64 restatement of the oracle (the
reference gradient), linearised at the float32 forward trajectory the adjoint
differentiates, and float64 central differences of the NumPy oracle along the
same trajectory.  Tolerances: relative L2 <= 1e-5 for gradients, 1e-4 for
finite-difference directional derivatives."""

import numpy as np
import pytest
import torch

from oracle import cdklm16_np as O
from oracle import cdklm16_torch as T
import helpers as H

pytestmark = pytest.mark.gpu

BCS = [("wall", (1, 1, 1, 1)), ("periodic", (2, 2, 2, 2)),
       ("periodicNS", (2, 1, 2, 1)), ("periodicEW", (1, 2, 1, 2))]
KINDS = {"plain": {}, "shallow": {"shallow": True}, "land": {"land": True}}


def _say(*a):
    print(*a, flush=True)


def _mods():
    from SWESimulators import Common
    from SWESimulators import _gocean as G
    from SWESimulators import adjoint as A
    return Common, G, A


def _t64(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))


def _domain(p):
    Common, G, A = _mods()
    return Common.make_domain(p.nx, p.ny, p.dx, p.dy, p.g, 0.0, p.theta, 2, p.bc,
                              f=float(p.f), north_x=0.0, north_y=1.0,
                              desing_eps=p.desing_eps, flux_slope_eps=p.flux_slope_eps,
                              depth_cutoff=p.depth_cutoff, land_value=p.land_value)


def _case(nx, ny, bc, seed, kind, f=None):
    # At a wall face the mirrored ghost makes u- + u+ vanish up to rounding, so
    # the upwind choice of the tangential flux is decided by rounding; dyadic
    # depths make that rounding identical in float32 and float64.
    p, Hi, Hm, eta, hu, hv = H.random_case(nx, ny, bc, seed, dyadic=True, **KINDS[kind])
    if f is not None:
        p.f = np.float32(f)
    return p, Hi, eta, hu, hv


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _torch_vjp_substep(p, Hi, q, W, dt):
    """float64 J^T W of the rk_step-0 map (interior of W only)."""
    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    qt = [_t64(a).requires_grad_(True) for a in q]
    out = T.substep(dt, 0, *qt, *qt, _t64(Hi64), _t64(Hm64), p)
    L = sum((_t64(w)[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for w, o in zip(W, out))
    g = torch.autograd.grad(L, qt)
    return [x.numpy() for x in g]


def _weights(rng, nx, ny):
    W = [np.zeros((ny + 4, nx + 4)) for _ in range(3)]
    for w in W:
        w[2:-2, 2:-2] = rng.standard_normal((ny, nx))
    return W


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _hip_grad(p, Hi, q, W, dt, n, K):
    Common, G, A = _mods()
    d = _domain(p)
    adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=K)
    qd = [_dev(a).requires_grad_(True) for a in q]
    out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
    L = sum((_dev(w)[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for w, o in zip(W, out))
    L.backward()
    torch.cuda.synchronize()
    return [x.grad.cpu().numpy().astype(np.float64) for x in qd], [o.detach().cpu().numpy() for o in out]


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("f", [1.2e-4, 1e-2])
def test_adjoint_substep_vs_torch_fp64(bcname, bc, kind, f):
    """One substep's J^T w at a large dt (so flux and Coriolis terms are a
    large part of the product) against torch autograd in float64."""
    _say("adjoint substep", bcname, kind, f)
    Common, G, A = _mods()
    nx, ny, dt = 29, 23, 25.0
    p, Hi, eta, hu, hv = _case(nx, ny, bc, 5, kind, f=f)
    O.periodic_bc(eta, hu, hv, p)
    W = _weights(np.random.default_rng(7), nx, ny)
    ref = _torch_vjp_substep(p, Hi, [eta, hu, hv], W, dt)
    d = _domain(p)
    lib = G.load()
    q = [_dev(a) for a in (eta, hu, hv)]
    w = [_dev(a) for a in W]
    o = [torch.zeros_like(q[0]) for _ in range(3)]
    scratch = torch.empty(int(lib.gocean_adjoint_scratch_floats(d)), dtype=torch.float32, device="cuda")
    Hi_d = _dev(Hi)
    # forward output: at this dt some shallow cells are clamped dry, and the
    # clamp mask comes from the forward eta
    fwd = [torch.zeros_like(q[0]) for _ in range(3)]
    G.check(lib.gocean_cdklm_substep(d, dt, None, 0, *(G.ptr(a) for a in q), None, None, None,
                                     *(G.ptr(a) for a in fwd), G.ptr(Hi_d), G.Fields(), G.Wind(), -1,
                                     None), "forward substep")
    G.check(lib.gocean_adjoint_substep(d, None, dt, None, *(G.ptr(a) for a in q), *(G.ptr(a) for a in w),
                                       G.ptr(fwd[0]), None, None, None, None, 1.0, 0.0,
                                       *(G.ptr(a) for a in o), G.ptr(Hi_d), G.ptr(scratch), None),
            "adjoint substep")
    torch.cuda.synchronize()
    got = [a.cpu().numpy().astype(np.float64) for a in o]
    for name, g, r in zip(("eta", "hu", "hv"), got, ref):
        # ghost entries of the output adjoint are zero (folded into interior)
        assert np.all(g[:2] == 0) and np.all(g[-2:] == 0) and np.all(g[:, :2] == 0) and np.all(g[:, -2:] == 0)
        e = _rel(g[2:-2, 2:-2], r[2:-2, 2:-2])
        _say("  ", name, "rel L2", e)
        assert e <= 1e-5, (name, e)


def _trajectory(p, Hi, q0, dt, n):
    """(Q^n, Q*) of every step of the float32 forward run (the NumPy oracle in
    float32, bit-identical to the HIP forward, which another test checks)."""
    Hm = O.init_bm(Hi)
    dt32 = np.float32(dt)
    e, u, v = (np.array(a, np.float32) for a in q0)
    traj = []
    for _ in range(n):
        O.periodic_bc(e, u, v, p)
        s = [np.array(a) for a in O.cdklm_substep(dt32, 0, e, u, v, e.copy(), u.copy(), v.copy(), Hi, Hm, p)]
        O.periodic_bc(*s, p)
        nxt = [np.array(a) for a in O.cdklm_substep(dt32, 1, *s, e.copy(), u.copy(), v.copy(), Hi, Hm, p)]
        traj.append(((e, u, v), s))
        e, u, v = nxt
    return traj, (e, u, v)


def _interior(a):
    a = a.clone() if isinstance(a, torch.Tensor) else a.copy()
    a[:2] = 0
    a[-2:] = 0
    a[:, :2] = 0
    a[:, -2:] = 0
    return a


def _chain_grad(p, Hi, traj, lam_out, dt):
    """float64 reverse pass through each RK2 step, linearised at the float32
    trajectory points (torch autograd of the float64 restatement)."""
    Hi64 = _t64(H.hi64(Hi))
    Hm64 = _t64(O.init_bm(H.hi64(Hi), dtype=np.float64))
    lam = [_interior(_t64(a)) for a in lam_out]
    for qn, qs in reversed(traj):
        qn_t = [_t64(a).requires_grad_(True) for a in qn]
        qs_t = [_t64(a).requires_grad_(True) for a in qs]
        out = T.substep(dt, 1, *qs_t, *qn_t, Hi64, Hm64, p)
        L = sum((l[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for l, o in zip(lam, out))
        g = torch.autograd.grad(L, qs_t + qn_t, allow_unused=True)
        g_s = [_interior(x) for x in g[:3]]
        g_n = [x if x is not None else torch.zeros_like(lam[0]) for x in g[3:]]
        qn2 = [_t64(a).requires_grad_(True) for a in qn]
        out2 = T.substep(dt, 0, *qn2, *qn2, Hi64, Hm64, p)
        L2 = sum((a[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for a, o in zip(g_s, out2))
        g2 = torch.autograd.grad(L2, qn2)
        lam = [_interior((a + b).detach()) for a, b in zip(g2, g_n)]
    return [x.numpy() for x in lam]


def _np_substep64(p, Hi64, Hm64, rk, q_in, q_n, dt):
    q_in = [np.array(a, dtype=np.float64) for a in q_in]
    q_n = [np.array(a, dtype=np.float64) for a in q_n]
    O.periodic_bc(*q_in, p)     # the oracle reads periodic images from the ghosts
    return [np.array(a) for a in O.cdklm_substep(dt, rk, *q_in, *q_n, Hi64, Hm64, p,
                                                 wind_t=0.0, dtype=np.float64)]


def _fd_tangent(p, Hi, traj, d0, dt, eps=1e-8):
    """float64 central-difference tangent-linear run through the substeps of
    the float32 trajectory, each linearised at the trajectory's own points:
    dQ* = D1 P(Q^n) dQ^n, dQ^{n+1} = D M1(Q*, Q^n) (dQ*, dQ^n), with P / M1 the
    two substeps of the NumPy oracle in float64.  The scheme is piecewise
    smooth and even discontinuous (the tangential flux is upwinded on the
    sign of u- + u+, and jumps where that sum changes sign), so the step must
    be small enough not to cross a switch: 1e-8, still far above float64
    round-off for these magnitudes."""
    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    dq = [_interior(np.asarray(a, np.float64)) for a in d0]
    for qn, qs in traj:
        qn64 = [a.astype(np.float64) for a in qn]
        qs64 = [a.astype(np.float64) for a in qs]
        # the tangent map is linear: difference along dq / |dq|_max and scale
        # back, so every substep is perturbed by exactly eps
        sc = max(float(np.abs(a).max()) for a in dq)
        u = [a / sc for a in dq]
        pl = _np_substep64(p, Hi64, Hm64, 0, [a + eps * b for a, b in zip(qn64, u)],
                           [a + eps * b for a, b in zip(qn64, u)], dt)
        mi = _np_substep64(p, Hi64, Hm64, 0, [a - eps * b for a, b in zip(qn64, u)],
                           [a - eps * b for a, b in zip(qn64, u)], dt)
        us = [_interior((a - b) / (2 * eps)) for a, b in zip(pl, mi)]
        ss = max(float(np.abs(a).max()) for a in us)
        pl = _np_substep64(p, Hi64, Hm64, 1, [a + eps * b / ss for a, b in zip(qs64, us)],
                           [a + eps * b / ss for a, b in zip(qn64, u)], dt)
        mi = _np_substep64(p, Hi64, Hm64, 1, [a - eps * b / ss for a, b in zip(qs64, us)],
                           [a - eps * b / ss for a, b in zip(qn64, u)], dt)
        dq = [_interior((a - b) / (2 * eps) * ss * sc) for a, b in zip(pl, mi)]
    return dq


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("kind", list(KINDS))
def test_adjoint_rk2_steps_vs_fp64_chain(bcname, bc, kind):
    """Gradient of a weighted sum of the state after 7 RK2 steps (checkpoint
    every 3) with respect to the initial eta, hu, hv: HIP adjoint vs the
    float64 reverse pass linearised along the same float32 trajectory."""
    _say("adjoint rk2 steps", bcname, kind)
    nx, ny, dt, n = 26, 21, 3.0, 7
    p, Hi, eta, hu, hv = _case(nx, ny, bc, 8, kind)
    W = _weights(np.random.default_rng(9), nx, ny)
    traj, _ = _trajectory(p, Hi, (eta, hu, hv), dt, n)
    ref = _chain_grad(p, Hi, traj, W, dt)
    got, _ = _hip_grad(p, Hi, [eta, hu, hv], W, dt, n, K=3)
    for name, g, r in zip(("eta", "hu", "hv"), got, ref):
        e = _rel(g[2:-2, 2:-2], r[2:-2, 2:-2])
        _say("  ", name, "rel L2", e)
        assert e <= 1e-5, (name, e)
        assert np.all(g[:2] == 0) and np.all(g[:, :2] == 0)


def test_adjoint_dry_clamp_vs_fp64_chain():
    """Steps long enough to dry cells of the shallow strip: cells the forward
    clamps pass no adjoint, as in the float64 reverse pass."""
    _say("adjoint dry clamp")
    nx, ny, dt, n = 24, 19, 12.0, 3
    p, Hi, eta, hu, hv = _case(nx, ny, (1, 1, 1, 1), 4, "shallow")
    hu[:, :10] += 0.05
    W = _weights(np.random.default_rng(3), nx, ny)
    traj, end = _trajectory(p, Hi, (eta, hu, hv), dt, n)
    Hm = O.init_bm(Hi)
    Hreg = np.float32(0.25) * (((Hi[:-1, :-1] + Hi[1:, :-1]) + Hi[:-1, 1:]) + Hi[1:, 1:])
    n_clamped = sum(int((q[0] <= -Hreg + np.float32(p.depth_cutoff))[2:-2, 2:-2].sum())
                    for q in [t[1] for t in traj] + [end])
    _say("   clamped cells over the run", n_clamped)
    assert n_clamped > 0
    ref = _chain_grad(p, Hi, traj, W, dt)
    got, _ = _hip_grad(p, Hi, [eta, hu, hv], W, dt, n, K=2)
    for name, g, r in zip(("eta", "hu", "hv"), got, ref):
        e = _rel(g[2:-2, 2:-2], r[2:-2, 2:-2])
        _say("  ", name, "rel L2", e)
        assert e <= 1e-5, (name, e)


@pytest.mark.parametrize("bcname,bc,kind", [("wall", (1, 1, 1, 1), "plain"),
                                             ("periodic", (2, 2, 2, 2), "land"),
                                             ("periodicEW", (1, 2, 1, 2), "shallow")])
def test_adjoint_point_misfit_fd_fp64(bcname, bc, kind):
    """Gradient of the point-observation misfit J = 0.5 sum (eta(x_k) - y_k)^2
    (bilinear interpolation, 16 points) after 12 steps, checkpoint every 5:

    * along 4 seeded random directions d, <grad J, d> from the HIP adjoint
      equals <dJ/d(eta_n), dq_n> with dq_n from a float64 central-difference
      (eps = 1e-8) tangent-linear run of the NumPy oracle through the same
      float32 trajectory's substeps, to 1e-4 of sum |grad J * d|;
    * the full gradient equals the float64 reverse pass along that
      trajectory to 1e-5 relative L2;
    * against the gradient of the float64 run (its own trajectory) the
      difference is reported: float32 rounding moves the trajectory by about
      1e-5 of the signal, across kinks of the scheme (wave-speed maxima,
      limiter switches), so that gradient is a neighbouring smooth piece."""
    _say("adjoint point misfit", bcname, kind)
    Common, G, A = _mods()
    nx, ny, dt, n = 32, 27, 3.0, 12
    p, Hi, eta, hu, hv = _case(nx, ny, bc, 21, kind)
    rng = np.random.default_rng(31)
    ox = rng.uniform(1.0, nx - 1.0, 16) * p.dx
    oy = rng.uniform(1.0, ny - 1.0, 16) * p.dy
    yk = 0.05 * rng.standard_normal(16)

    d = _domain(p)
    adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=5)
    qd = [_dev(a).requires_grad_(True) for a in (eta, hu, hv)]
    out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
    A.point_misfit(out[0], ox, oy, yk, p.dx, p.dy).backward()
    torch.cuda.synchronize()
    g = [x.grad.cpu().numpy().astype(np.float64) for x in qd]

    traj, end = _trajectory(p, Hi, (eta, hu, hv), dt, n)
    e_end = _t64(end[0]).requires_grad_(True)
    A.point_misfit(e_end, ox, oy, yk, p.dx, p.dy).backward()
    lam_n = [e_end.grad.numpy(), np.zeros_like(end[0], dtype=np.float64), np.zeros_like(end[0], dtype=np.float64)]
    ref = _chain_grad(p, Hi, traj, lam_n, dt)
    for name, a, b in zip(("eta", "hu", "hv"), g, ref):
        e = _rel(a[2:-2, 2:-2], b[2:-2, 2:-2])
        _say("   vs fp64 reverse pass", name, e)
        assert e <= 1e-5, (name, e)

    for k in range(4):
        D = [_interior(rng.standard_normal(end[0].shape)) for _ in range(3)]
        dq = _fd_tangent(p, Hi, traj, D, dt)
        fd = float(np.sum(lam_n[0] * dq[0]))
        ad = sum(float(np.sum(a * b)) for a, b in zip(g, D))
        # relative to the size of the summed terms: the sum itself can cancel
        scale = sum(float(np.sum(np.abs(a * b))) for a, b in zip(g, D))
        _say("   direction", k, "fd", fd, "adjoint", ad, "rel", abs(ad - fd) / scale)
        assert abs(ad - fd) <= 1e-4 * scale, (k, ad, fd, scale)

    Hi64 = H.hi64(Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    qt = [_t64(a).requires_grad_(True) for a in (eta, hu, hv)]
    ot = T.rk2_steps(*qt, _t64(Hi64), _t64(Hm64), p, dt, n)
    A.point_misfit(ot[0], ox, oy, yk, p.dx, p.dy).backward()
    for name, a, b in zip(("eta", "hu", "hv"), g, qt):
        _say("   vs gradient of the float64 run", name, _rel(a[2:-2, 2:-2], b.grad.numpy()[2:-2, 2:-2]))


def test_adjoint_checkpoint_interval_invariance():
    """The gradient does not depend on the checkpoint interval (the recompute
    is bit-identical to the forward run)."""
    _say("adjoint checkpoint invariance")
    nx, ny, dt, n = 30, 22, 3.0, 9
    p, Hi, eta, hu, hv = _case(nx, ny, (2, 1, 2, 1), 12, "plain")
    W = _weights(np.random.default_rng(5), nx, ny)
    base, out0 = _hip_grad(p, Hi, [eta, hu, hv], W, dt, n, K=n)
    for K in (1, 2, 4):
        g, out = _hip_grad(p, Hi, [eta, hu, hv], W, dt, n, K=K)
        for a, b in zip(g, base):
            assert np.array_equal(a, b), K
        for a, b in zip(out, out0):
            assert np.array_equal(a, b), K


def test_adjoint_forward_is_the_parity_path():
    """The autograd wrapper's forward output equals the fp32 oracle run of
    the same steps bit for bit."""
    _say("adjoint forward parity")
    nx, ny, dt, n = 27, 20, 3.0, 5
    p, Hi, eta, hu, hv = _case(nx, ny, (1, 2, 1, 2), 14, "land")
    W = _weights(np.random.default_rng(1), nx, ny)
    _, out = _hip_grad(p, Hi, [eta, hu, hv], W, dt, n, K=2)
    Hm = O.init_bm(Hi)
    e, u, v = eta.copy(), hu.copy(), hv.copy()
    dt32 = np.float32(dt)
    for _ in range(n):
        O.periodic_bc(e, u, v, p)
        s = [np.array(a) for a in O.cdklm_substep(dt32, 0, e, u, v, e.copy(), u.copy(), v.copy(), Hi, Hm, p)]
        O.periodic_bc(*s, p)
        e, u, v = (np.array(a) for a in O.cdklm_substep(dt32, 1, *s, e, u, v, Hi, Hm, p))
    for a, b in zip(out, (e, u, v)):
        assert np.array_equal(a[2:-2, 2:-2], b[2:-2, 2:-2])


@pytest.mark.parametrize("bcname,bc", BCS)
def test_adjoint_ghost_cells(bcname, bc):
    """Loss on one output ghost cell: a periodic image passes its adjoint to
    the interior cell it copies; a wall-side ghost cell (left untouched by
    the forward) passes it to the same input ghost cell."""
    _say("adjoint ghost cells", bcname)
    nx, ny, dt, n = 20, 17, 3.0, 2
    p, Hi, eta, hu, hv = _case(nx, ny, bc, 2, "plain")
    Common, G, A = _mods()
    d = _domain(p)
    for (R, C) in [(0, 7), (ny + 3, 5), (6, 1), (9, nx + 2), (1, 1)]:
        for fld in range(3):
            adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=2)
            qd = [_dev(a).requires_grad_(True) for a in (eta, hu, hv)]
            out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
            out[fld][R, C].backward()
            g = [x.grad.cpu().numpy() for x in qd]
            gr, gc = R < 2 or R > ny + 1, C < 2 or C > nx + 1
            written = (not gr or bc[0] == 2) and (not gc or bc[1] == 2)
            if written:
                Ri = R + ny if R < 2 else (R - ny if R > ny + 1 else R)
                Ci = C + nx if C < 2 else (C - nx if C > nx + 1 else C)
                adj2 = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=2)
                qd2 = [_dev(a).requires_grad_(True) for a in (eta, hu, hv)]
                out2 = A.CDKLM16RK2.apply(*qd2, adj2, n, dt)
                out2[fld][Ri, Ci].backward()
                for a, b in zip(g, qd2):
                    assert np.array_equal(a, b.grad.cpu().numpy()), (R, C, fld)
            else:
                for k, a in enumerate(g):
                    expect = np.zeros_like(a)
                    if k == fld:
                        expect[R, C] = 1.0
                    assert np.array_equal(a, expect), (R, C, fld)


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("kind", ["plain", "land"])
def test_kept_records_match_recompute(bcname, bc, kind):
    """Records kept from the forward run (gocean_adjoint_record /
    gocean_adjoint_reverse) give the same bits as the checkpointed run that
    recomputes them: forward state (every cell, wall ghosts included) and
    gradient, for an uneven last segment; a second backward over the same
    records gives the same gradient again."""
    _say("kept records vs recompute", bcname, kind)
    Common, G, A = _mods()
    nx, ny, dt, n = 37, 29, 3.0, 7
    p, Hi, eta, hu, hv = _case(nx, ny, bc, 13, kind)
    rng = np.random.default_rng(5)
    W = [_dev(w) for w in _weights(rng, nx, ny)]
    d = _domain(p)
    res = {}
    for keep in (False, True):
        adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=3, keep_records=keep)
        qd = [_dev(a).requires_grad_(True) for a in (eta, hu, hv)]
        out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
        assert (adj.records is not None) == keep
        L = sum((w * o).sum() for w, o in zip(W, out))
        L.backward()
        lam = adj.backward(tuple(w.clone() for w in W))
        torch.cuda.synchronize()
        res[keep] = ([o.detach().cpu().numpy() for o in out], [x.grad.cpu().numpy() for x in qd],
                     [x.cpu().numpy() for x in lam])
    for a, b in zip(res[False][0], res[True][0]):
        assert np.array_equal(a, b)
    for a, b, c in zip(res[False][1], res[True][1], res[True][2]):
        assert np.array_equal(a, b) and np.array_equal(b, c)


def test_adjoint_rejects_out_of_scope_domains():
    Common, G, A = _mods()
    p, Hi, eta, hu, hv = _case(12, 10, (1, 1, 1, 1), 1, "plain")
    d = _domain(p)
    d.rk_order = 3
    with pytest.raises(ValueError):
        A.RK2Adjoint(d, _dev(Hi))


def test_ensemble_adjoint_matches_single_domain():
    """The batched ensemble reverse mode (fused launches over members of
    different sizes, boundaries, f, dx and dt) gives each member exactly the
    gradient of the single-domain reverse mode."""
    _say("ensemble adjoint vs single-domain adjoint")
    from SWESimulators import Common, CDKLM16Ensemble
    from test_gpu_ensemble import members
    Common_, G, A = _mods()
    ms = members(seed=3)[:3]
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.6)
    n, K = 7, 3
    ea = A.EnsembleAdjoint(ens, checkpoint_every=K)
    state0 = ens.state[0].clone()
    out = ea.forward(state0, n)
    rng = np.random.default_rng(4)
    lam = torch.zeros_like(state0)
    for k in range(ens.n):
        o, pl = int(ens.state_off[k]), ens.planes[k]
        for f in range(3):
            v = lam[o + f * pl:o + (f + 1) * pl].view(ens.ny[k] + 4, ens.pitch[k])
            v[2:-2, 2:-2] = torch.from_numpy(rng.standard_normal((ens.ny[k], ens.nx[k])).astype(np.float32))
    grad = ea.backward(lam)
    torch.cuda.synchronize()
    dts = ens.dt_host()
    for k, m in enumerate(ms):
        nx, ny = ens.nx[k], ens.ny[k]
        d = Common.make_domain(nx, ny, m["dx"], m["dy"], m["g"], 0.0, 1.3, 2, m["bc"], pitch=ens.pitch[k],
                               hi_pitch=nx + 5, f=m["f"])
        Hi_k = ens.Hi[int(ens.hi_off[k]):int(ens.hi_off[k + 1])].view(ny + 5, nx + 5)
        w = G.Wind()
        w.x0 = w.x1 = float(np.float32(m["wind_x"]))
        w.y0 = w.y1 = float(np.float32(m["wind_y"]))
        adj = A.RK2Adjoint(d, Hi_k, wind=w, checkpoint_every=K)
        o, pl = int(ens.state_off[k]), ens.planes[k]
        view = lambda t: tuple(t[o + f * pl:o + (f + 1) * pl].view(ny + 4, ens.pitch[k]) for f in range(3))
        fo = adj.forward(view(state0), n, float(dts[k]))
        gk = adj.backward(view(lam))
        torch.cuda.synchronize()
        for a, b in zip(fo, view(out)):
            assert torch.equal(a, b), ("forward", k)
        for name, a, b in zip(("eta", "hu", "hv"), gk, view(grad)):
            same = torch.equal(a, b)
            _say("   member", k, name, "bit-identical" if same else
                 f"max |d| {float((a - b).abs().max()):.3e} of {float(a.abs().max()):.3e}")
            assert same, (k, name)


def test_ensemble_kept_records_match_recompute():
    """Batched reverse mode with the records kept from the forward run
    (gocean_ensemble_adjoint_record / _reverse) against the checkpointed run
    that recomputes them: same final state (every cell of every member, and
    the ensemble's own state afterwards) and the same gradient bits, with an
    uneven last segment."""
    _say("ensemble kept records vs recompute")
    from SWESimulators import Common, CDKLM16Ensemble
    from test_gpu_ensemble import members
    Common_, G, A = _mods()
    ms = members(seed=5)[:4]
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.6)
    state0 = ens.state[0].clone()
    lam = torch.zeros_like(state0)
    rng = np.random.default_rng(8)
    for k in range(ens.n):
        o, pl = int(ens.state_off[k]), ens.planes[k]
        for f in range(3):
            v = lam[o + f * pl:o + (f + 1) * pl].view(ens.ny[k] + 4, ens.pitch[k])
            v[2:-2, 2:-2] = torch.from_numpy(rng.standard_normal((ens.ny[k], ens.nx[k])).astype(np.float32))
    res = {}
    for keep in (False, True):
        ea = A.EnsembleAdjoint(ens, checkpoint_every=3, keep_records=keep)
        out = ea.forward(state0, 8)
        assert (ea.records is not None) == keep
        g1 = ea.backward(lam)
        g2 = ea.backward(lam)
        torch.cuda.synchronize()
        res[keep] = (out.clone(), ens.state[0].clone(), g1, g2)
    for a, b in zip(res[False], res[True]):
        assert torch.equal(a, b)
    assert torch.equal(res[True][2], res[True][3])


def test_record_tile_configs_record_the_same_gradient():
    """Every tile configuration of the recording forward writes the same
    record: forward output and gradient are bit-identical whichever tile
    records, with the records kept and recomputed."""
    import torch
    from SWESimulators import autotune
    _say("record tile configurations")
    Common, G, A = _mods()
    nx, ny, dt, n = 37, 29, 3.0, 7
    p, Hi, eta, hu, hv = _case(nx, ny, (2, 1, 2, 1), 13, "plain")
    W = [_dev(w) for w in _weights(np.random.default_rng(5), nx, ny)]
    d = _domain(p)
    ref = None
    for cfg in range(len(autotune.tile_config_names())):
        for keep in (True, False):
            adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=3, keep_records=keep, record_tile_config=cfg)
            qd = [_dev(a).requires_grad_(True) for a in (eta, hu, hv)]
            out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
            sum((w * o).sum() for w, o in zip(W, out)).backward()
            torch.cuda.synchronize()
            got = [o.detach().cpu().numpy() for o in out] + [x.grad.cpu().numpy() for x in qd]
            if ref is None:
                ref = got
            for x, y in zip(ref, got):
                assert np.array_equal(x, y), (cfg, keep)
    G.check(G.load().gocean_adjoint_set_record_tile_config(-1), "reset")


@pytest.mark.parametrize("bcname,bc", [BCS[0], BCS[1], BCS[3]])
@pytest.mark.parametrize("kind", ["plain", "land"])
def test_recording_forward_interior_tiles(bcname, bc, kind):
    """A domain large enough (112 x 70) that most tiles of every recording
    tile configuration are interior tiles (every access at least one cell
    inside the interior), so the recording forward's interior body runs:
    its output equals the plain forward bit for bit, kept records equal
    recomputed ones, and the gradient equals the float64 reverse pass along
    the float32 trajectory."""
    from SWESimulators import autotune
    _say("recording forward interior tiles", bcname, kind)
    Common, G, A = _mods()
    nx, ny, dt, n = 112, 70, 3.0, 5
    p, Hi, eta, hu, hv = _case(nx, ny, bc, 17, kind)
    W = _weights(np.random.default_rng(19), nx, ny)
    d = _domain(p)
    lib = G.load()
    # the plain forward (k_substep) for the same steps
    q = [_dev(a) for a in (eta, hu, hv)]
    tmp = [x.clone() for x in q]
    G.check(lib.gocean_rk_steps(d, n, dt, None, *(G.ptr(x) for x in q), *(G.ptr(x) for x in tmp),
                                None, None, None, G.ptr(_dev(Hi)), G.Fields(), G.Wind(), -1, None), "rk")
    torch.cuda.synchronize()
    plain = [x.cpu().numpy() for x in q]
    ref = None
    for cfg in (1, 5, 6):
        for keep in (True, False):
            adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=2, keep_records=keep, record_tile_config=cfg)
            qd = [_dev(a).requires_grad_(True) for a in (eta, hu, hv)]
            out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
            sum((_dev(w) * o).sum() for w, o in zip(W, out)).backward()
            torch.cuda.synchronize()
            for a, b in zip(out, plain):
                assert np.array_equal(a.detach().cpu().numpy(), b), (cfg, keep)
            got = [x.grad.cpu().numpy().astype(np.float64) for x in qd]
            if ref is None:
                ref = got
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), (cfg, keep)
    G.check(lib.gocean_adjoint_set_record_tile_config(-1), "reset")
    traj, _ = _trajectory(p, Hi, (eta, hu, hv), dt, n)
    want = _chain_grad(p, Hi, traj, W, dt)
    for name, g, r in zip(("eta", "hu", "hv"), ref, want):
        e = _rel(g[2:-2, 2:-2], r[2:-2, 2:-2])
        _say("  ", name, "rel L2", e)
        assert e <= 1e-5, (name, e)


def test_ensemble_adjoint_interior_tiles_match_single_domain():
    """Ensemble members of 100 x 60 and 80 x 40 (interior tiles in the
    batched recording forward and reverse kernels) get exactly the
    single-domain gradient, with the records kept and recomputed."""
    _say("ensemble adjoint, members with interior tiles")
    from SWESimulators import Common, CDKLM16Ensemble
    from test_gpu_ensemble import members
    Common_, G, A = _mods()
    ms = members(seed=9, sizes=((100, 60), (80, 40), (70, 70)))
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.6)
    n, K = 6, 4
    state0 = ens.state[0].clone()
    rng = np.random.default_rng(12)
    lam = torch.zeros_like(state0)
    for k in range(ens.n):
        o, pl = int(ens.state_off[k]), ens.planes[k]
        for f in range(3):
            v = lam[o + f * pl:o + (f + 1) * pl].view(ens.ny[k] + 4, ens.pitch[k])
            v[2:-2, 2:-2] = torch.from_numpy(rng.standard_normal((ens.ny[k], ens.nx[k])).astype(np.float32))
    dts = ens.dt_host()
    for keep in (True, False):
        ea = A.EnsembleAdjoint(ens, checkpoint_every=K, keep_records=keep)
        out = ea.forward(state0, n)
        grad = ea.backward(lam)
        torch.cuda.synchronize()
        for k, m in enumerate(ms):
            nx, ny = ens.nx[k], ens.ny[k]
            d = Common.make_domain(nx, ny, m["dx"], m["dy"], m["g"], 0.0, 1.3, 2, m["bc"], pitch=ens.pitch[k],
                                   hi_pitch=nx + 5, f=m["f"])
            Hi_k = ens.Hi[int(ens.hi_off[k]):int(ens.hi_off[k + 1])].view(ny + 5, nx + 5)
            w = G.Wind()
            w.x0 = w.x1 = float(np.float32(m["wind_x"]))
            w.y0 = w.y1 = float(np.float32(m["wind_y"]))
            adj = A.RK2Adjoint(d, Hi_k, wind=w, checkpoint_every=K, keep_records=keep)
            o, pl = int(ens.state_off[k]), ens.planes[k]
            view = lambda t: tuple(t[o + f * pl:o + (f + 1) * pl].view(ny + 4, ens.pitch[k]) for f in range(3))
            fo = adj.forward(view(state0), n, float(dts[k]))
            gk = adj.backward(view(lam))
            torch.cuda.synchronize()
            for a, b in zip(fo, view(out)):
                assert torch.equal(a, b), ("forward", k, keep)
            for a, b in zip(gk, view(grad)):
                assert torch.equal(a, b), ("gradient", k, keep)
