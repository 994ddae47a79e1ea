"""The column-marching substep adjoint against the tile adjoint: both kernels
evaluate the same vector-Jacobian products with the same operations in the
same order, so the gradients of whole checkpointed RK2 runs must agree bit
for bit (values below 2^-126 taken as zero, as the parity tests do).  The
domains put strip edges on every boundary case of the marching kernel
(58 output columns per strip: nx = 56, 58, 59, 116, 117), short and tall row
strips, land, near-dry cells that the forward clamps, every boundary
condition, and the ghost-cell fold between substeps (several steps per
checkpoint segment)."""

import numpy as np
import pytest
import torch

import helpers as H

pytestmark = pytest.mark.gpu

BCS = [("wall", (1, 1, 1, 1)), ("periodic", (2, 2, 2, 2)),
       ("periodicNS", (2, 1, 2, 1)), ("periodicEW", (1, 2, 1, 2))]
SIZES = [(37, 29), (56, 21), (58, 40), (59, 33), (116, 9), (117, 70)]


def _say(*a):
    print(*a, flush=True)


def _mods():
    from SWESimulators import Common
    from SWESimulators import _gocean as G
    from SWESimulators import adjoint as A
    return Common, G, A


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _grad(kernel, p, Hi, q, W, dt, n, K, keep):
    Common, G, A = _mods()
    lib = G.load()
    G.check(lib.gocean_adjoint_set_reverse_kernel(kernel), "set reverse kernel")
    d = Common.make_domain(p.nx, p.ny, p.dx, p.dy, p.g, 0.0, p.theta, 2, p.bc, f=float(p.f),
                           north_x=0.0, north_y=1.0)
    adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=K, keep_records=keep)
    qd = [_dev(a).requires_grad_(True) for a in q]
    out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
    sum((w * o).sum() for w, o in zip(W, out)).backward()
    torch.cuda.synchronize()
    return [x.grad.cpu().numpy() for x in qd]


def _compare(p, Hi, q, dt, n, K, seed, keep=True):
    rng = np.random.default_rng(seed)
    W = [_dev(rng.standard_normal(q[0].shape)) for _ in range(3)]
    try:
        a = _grad(0, p, Hi, q, W, dt, n, K, keep)
        b = _grad(1, p, Hi, q, W, dt, n, K, keep)
    finally:
        _, G, _ = _mods()
        G.check(G.load().gocean_adjoint_set_reverse_kernel(-1), "reset")
        assert G.load().gocean_adjoint_reverse_kernel() == -1
    for name, x, y in zip(("eta", "hu", "hv"), a, b):
        assert np.isfinite(x).all() and np.abs(x).max() > 0, name
        if not H.normal_equal(x, y):
            bad = np.argwhere(x != y)
            raise AssertionError(f"{name}: tile and marching adjoints differ at {len(bad)} cells, first "
                                 f"{H.diff_report(x, y, 8)}")


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("size", SIZES)
def test_marching_adjoint_equals_tile_adjoint(bcname, bc, size):
    nx, ny = size
    _say("marching vs tile adjoint", bcname, nx, ny)
    p, Hi, Hm, eta, hu, hv = H.random_case(nx, ny, bc, 11, land=(nx + ny) % 2 == 0)
    _compare(p, Hi, [eta, hu, hv], 3.0, 4, 2, seed=nx * 100 + ny)


@pytest.mark.parametrize("bcname,bc", [BCS[0], BCS[3]])
def test_marching_adjoint_equals_tile_adjoint_near_dry(bcname, bc):
    """A shallow strip and a step long enough to clamp cells dry: the clamp
    masks of lambda and of the extra term, and the desingularised and dry
    conversion branches."""
    _say("marching vs tile adjoint, near-dry", bcname)
    p, Hi, Hm, eta, hu, hv = H.random_case(61, 27, bc, 4, shallow=True)
    hu = hu.copy()
    hu[:, :12] += 0.05
    _compare(p, Hi, [eta, hu, hv], 20.0, 3, 3, seed=3)


def test_marching_adjoint_equals_tile_adjoint_recomputed_records():
    """Records recomputed per segment instead of kept."""
    _say("marching vs tile adjoint, recomputed records")
    p, Hi, Hm, eta, hu, hv = H.random_case(117, 45, (2, 1, 2, 1), 8)
    _compare(p, Hi, [eta, hu, hv], 3.0, 5, 2, seed=9, keep=False)


def _ensemble_grad(kernel, ms, n, K, keep, lam_seed):
    from SWESimulators import CDKLM16Ensemble
    Common, G, A = _mods()
    lib = G.load()
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.6)
    state0 = ens.state[0].clone()
    rng = np.random.default_rng(lam_seed)
    lam = torch.zeros_like(state0)
    for k in range(ens.n):
        o, pl = int(ens.state_off[k]), ens.planes[k]
        for f in range(3):
            v = lam[o + f * pl:o + (f + 1) * pl].view(ens.ny[k] + 4, ens.pitch[k])
            v[2:-2, 2:-2] = torch.from_numpy(rng.standard_normal((ens.ny[k], ens.nx[k])).astype(np.float32))
    G.check(lib.gocean_adjoint_set_reverse_kernel(kernel), "set reverse kernel")
    ea = A.EnsembleAdjoint(ens, checkpoint_every=K, keep_records=keep)
    out = ea.forward(state0, n)
    grad = ea.backward(lam)
    torch.cuda.synchronize()
    return out.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.parametrize("keep", [True, False])
def test_ensemble_marching_adjoint_equals_tile_adjoint(keep):
    """The batched reverse launch over every member's strips (members of
    different sizes, boundaries, f, dx and dt in one launch) against the
    batched tile adjoint: the same final state and gradient bits."""
    from test_gpu_ensemble import members
    _say("ensemble marching vs tile adjoint, keep_records", keep)
    ms = members(seed=5, sizes=((117, 70), (58, 40), (59, 33), (37, 29)))
    try:
        fa, ga = _ensemble_grad(0, ms, 5, 2, keep, 8)
        fb, gb = _ensemble_grad(1, ms, 5, 2, keep, 8)
    finally:
        _, G, _ = _mods()
        G.check(G.load().gocean_adjoint_set_reverse_kernel(-1), "reset")
    assert np.array_equal(fa, fb)
    assert np.isfinite(ga).all() and np.abs(ga).max() > 0
    if not H.normal_equal(ga, gb):
        raise AssertionError(f"ensemble gradients differ at {int(np.sum(ga != gb))} entries, first "
                             f"{H.diff_report(ga, gb, 8)}")


def _rk2_grad(p, Hi, q, W, dt, n, K, keep, record_cfg):
    Common, G, A = _mods()
    d = Common.make_domain(p.nx, p.ny, p.dx, p.dy, p.g, 0.0, p.theta, 2, p.bc, f=float(p.f),
                           north_x=0.0, north_y=1.0)
    adj = A.RK2Adjoint(d, _dev(Hi), checkpoint_every=K, keep_records=keep, record_tile_config=record_cfg)
    qd = [_dev(a).requires_grad_(True) for a in q]
    out = A.CDKLM16RK2.apply(*qd, adj, n, dt)
    sum((w * o).sum() for w, o in zip(W, out)).backward()
    torch.cuda.synchronize()
    return [o.detach().cpu().numpy() for o in out] + [x.grad.cpu().numpy() for x in qd]


REC_SIZES = [(61, 23), (121, 50), (62, 37)]


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("size", REC_SIZES)
def test_marching_record_equals_tile_record(bcname, bc, size):
    """The column-marching recording forward (strips of 60 output columns:
    nx = 61, 62, 121 put strip edges next to the ghost columns; 16 and an
    automatic number of rows per strip) writes the branch record the
    reverse pass reads exactly as the tile recording forward: the same
    forward state and gradient bits, with records kept and recomputed."""
    nx, ny = size
    _say("marching vs tile record", bcname, nx, ny)
    _, G, _ = _mods()
    p, Hi, Hm, eta, hu, hv = H.random_case(nx, ny, bc, 21, land=(nx + ny) % 2 == 1)
    rng = np.random.default_rng(nx + ny)
    W = [_dev(rng.standard_normal(eta.shape)) for _ in range(3)]
    try:
        for keep in (True, False):
            ref = _rk2_grad(p, Hi, [eta, hu, hv], W, 3.0, 4, 2, keep, 5)
            for cfg in (8, 9):
                got = _rk2_grad(p, Hi, [eta, hu, hv], W, 3.0, 4, 2, keep, cfg)
                for k, (x, y) in enumerate(zip(ref, got)):
                    assert np.isfinite(x).all()
                    if not H.normal_equal(x, y):
                        raise AssertionError(f"record cfg {cfg} keep {keep} output {k}: "
                                             f"{int(np.sum(x != y))} entries differ, first {H.diff_report(x, y, 8)}")
    finally:
        G.check(G.load().gocean_adjoint_set_record_tile_config(-1), "reset")


def _ens_run(ms, tile, n, K, keep, seed):
    from SWESimulators import CDKLM16Ensemble
    Common, G, A = _mods()
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, tile_config=tile, courant_number=0.6)
    state0 = ens.state[0].clone()
    rng = np.random.default_rng(seed)
    lam = torch.zeros_like(state0)
    for k in range(ens.n):
        o, pl = int(ens.state_off[k]), ens.planes[k]
        for f in range(3):
            v = lam[o + f * pl:o + (f + 1) * pl].view(ens.ny[k] + 4, ens.pitch[k])
            v[2:-2, 2:-2] = torch.from_numpy(rng.standard_normal((ens.ny[k], ens.nx[k])).astype(np.float32))
    ea = A.EnsembleAdjoint(ens, checkpoint_every=K, keep_records=keep)
    out = ea.forward(state0, n)
    grad = ea.backward(lam)
    torch.cuda.synchronize()
    return out.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.parametrize("keep", [True, False])
def test_ensemble_marching_record_equals_tile_record(keep):
    """An ensemble on a marching configuration records through its strips;
    the forward state and the gradient equal those of the same ensemble on
    the 32x16 tile configuration bit for bit."""
    from test_gpu_ensemble import members
    _say("ensemble marching vs tile record, keep_records", keep)
    ms = members(seed=6, sizes=((121, 70), (61, 40), (62, 33), (37, 29)))
    fa, ga = _ens_run(ms, 1, 5, 2, keep, 3)
    for tile in (8, 9):
        fb, gb = _ens_run(ms, tile, 5, 2, keep, 3)
        assert np.array_equal(fa, fb), tile
        assert np.isfinite(ga).all() and np.abs(ga).max() > 0
        if not H.normal_equal(ga, gb):
            raise AssertionError(f"tile {tile}: gradients differ at {int(np.sum(ga != gb))} entries, first "
                                 f"{H.diff_report(ga, gb, 8)}")
