"""Localised ETKF on the batched ensemble state against the float64 oracle
(oracle/etkf_np.py letkf, the reference's ETKFOcean.LETKF, ETKFOcean.py:
583-744): boxes that wrap the periodic domain, several groups, buoys on a
lattice of cell centres (the reference's buoy layout, whose box centres take
numpy's round half to even), inflation off / fixed / adaptive, scale_w above
one (the W / max(W, 1) blend)."""

import numpy as np
import pytest

from oracle import etkf_np as E

pytestmark = pytest.mark.gpu

DX = 400.0


def _members(N, nx, ny, seed):
    rs = np.random.RandomState(seed)
    j, i = np.mgrid[0:ny + 5, 0:nx + 5]
    Hi = (40 + 4 * np.sin(i / 6.0) + 3 * np.cos(j / 5.0)).astype(np.float32)
    base = 0.2 * np.exp(-((i[:-1, :-1] - nx / 2) ** 2 + (j[:-1, :-1] - ny / 2) ** 2) / 60.0)
    out = []
    for m in range(N):
        eta = (base + 0.02 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
        hu = (1.0 + 0.5 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
        hv = (-0.5 + 0.5 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
        out.append(dict(nx=nx, ny=ny, dx=DX, dy=DX, g=9.81, f=1e-4, r=0.0, bc=(2, 2, 2, 2),
                        eta=eta, hu=hu, hv=hv, Hi=Hi))
    return out


def _lattice(nx, ny, step):
    ys, xs = np.mgrid[0:ny:step, 0:nx:step]
    return np.column_stack([(xs.ravel() + 0.5) * DX, (ys.ravel() + 0.5) * DX])


CASES = [
    # N, nx, ny, r, buoys, scale_w, inflation
    ("random", 16, 48, 40, 2.0, 7, 1.0, 1.0),
    ("random_many_groups", 24, 64, 48, 3.0, 20, 2.0, 1.3),
    ("random_adaptive", 40, 50, 50, 2.5, 15, 1.0, 0.0),
    ("lattice", 20, 60, 45, 2.0, 6, 1.0, 1.0),
    ("many_members", 300, 32, 30, 2.0, 4, 1.0, 0.0),   # members beyond the register path and one coef sweep
]


@pytest.mark.parametrize("name,N,nx,ny,r,Nd,scale_w,inflation", CASES, ids=[c[0] for c in CASES])
def test_letkf_matches_oracle(name, N, nx, ny, r, Nd, scale_w, inflation):
    import torch
    from SWESimulators import CDKLM16Ensemble, Common, ETKFOcean
    print("case", name, N, nx, ny, r, Nd, scale_w, inflation, flush=True)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, _members(N, nx, ny, seed=N), rk_order=2, courant_number=0.7)
    rs = np.random.RandomState(11)
    if name == "lattice":
        pos = _lattice(nx, ny, Nd)
    else:
        pos = np.column_stack([rs.uniform(0, nx * DX, Nd), rs.uniform(0, ny * DX, Nd)])
        pos[0] = [0.3 * DX, (ny - 0.2) * DX]            # box wraps in x and y
    torch.cuda.synchronize()
    X_f = np.stack([np.stack(ens.download(k, interior_domain_only=True)) for k in range(N)]).astype(np.float32)
    cells = np.floor(pos / DX).astype(int)
    truth = X_f[:, 1:, cells[:, 1], cells[:, 0]].mean(axis=0).T + 0.4 * rs.randn(len(pos), 2)
    obs = np.column_stack([pos, truth])
    R2 = np.array([[0.09, 0.02], [0.02, 0.16]])
    want = E.letkf(X_f, obs, pos, R2, r, DX, DX, scale_w=scale_w, inflation=inflation)
    et = ETKFOcean.ETKFOcean(ens, R2, inflation_factor=inflation, scale_w=scale_w)
    et.LETKF(obs, r_factor=r)
    torch.cuda.synchronize()
    assert et.groups == E.observation_groups(pos, r, DX, DX, nx, ny)
    if name == "random_many_groups":
        assert len(et.groups) > 1
    full = [ens.download(k) for k in range(N)]
    got = np.stack([np.stack([a[2:-2, 2:-2] for a in f]) for f in full])
    for f, fname in enumerate(("eta", "hu", "hv")):
        incr = np.max(np.abs(want[:, f] - X_f[:, f]))
        err = np.max(np.abs(got[:, f] - want[:, f]))
        print(fname, "max error", err, "max increment", incr, "ratio", err / incr, flush=True)
        assert incr > 0 and err <= 1e-5 * incr
    # cells outside every box keep the forecast bit for bit
    W = sum(E.combined_weights(pos[g], r, DX, DX, nx, ny, scale_w) for g in et.groups)
    np.testing.assert_array_equal(got[:, :, W == 0], X_f[:, :, W == 0])
    # periodic ghost cells refreshed from the analysed interior
    for k in (0, N - 1):
        for a in full[k]:
            np.testing.assert_array_equal(a[:2, 2:-2], a[-4:-2, 2:-2])
            np.testing.assert_array_equal(a[2:-2, :2], a[2:-2, -4:-2])
    # the analysis pulls the observed transports towards the observations
    # (with scale_w > 1 the reference's blend W_loc x_a + W_f x_f weighs the
    # local analysis by up to scale_w at the buoy, so this holds only up to 1)
    if scale_w <= 1.0:
        HX_a = got[:, 1:, cells[:, 1], cells[:, 0]].mean(axis=0).T
        HX_f = X_f[:, 1:, cells[:, 1], cells[:, 0]].mean(axis=0).T
        assert np.linalg.norm(HX_a - truth) < np.linalg.norm(HX_f - truth)


def test_letkf_repeated_calls_reuse_the_geometry():
    """A second analysis with the same buoys reuses the cached boxes and
    tiles and still matches the oracle applied twice."""
    import torch
    from SWESimulators import CDKLM16Ensemble, Common, ETKFOcean
    N, nx, ny, r = 12, 40, 40, 2.0
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, _members(N, nx, ny, seed=5), rk_order=2, courant_number=0.7)
    pos = _lattice(nx, ny, 9)
    torch.cuda.synchronize()
    X_f = np.stack([np.stack(ens.download(k, interior_domain_only=True)) for k in range(N)]).astype(np.float32)
    rs = np.random.RandomState(2)
    R2 = np.diag([0.1, 0.1])
    et = ETKFOcean.ETKFOcean(ens, R2)
    want = X_f
    for it in range(2):
        obs = np.column_stack([pos, 0.5 * rs.randn(len(pos), 2)])
        want = E.letkf(want, obs, pos, R2, r, DX, DX)
        et.LETKF(obs, r_factor=r)
        if it == 0:
            setup = et._letkf
    assert et._letkf is setup
    torch.cuda.synchronize()
    got = np.stack([np.stack(ens.download(k, interior_domain_only=True)) for k in range(N)])
    for f in range(3):
        incr = np.max(np.abs(want[:, f] - X_f[:, f]))
        assert np.max(np.abs(got[:, f] - want[:, f])) <= 1e-5 * incr


def test_letkf_double_jet_full_size_matches_oracle():
    """The bench's workload itself: 100 members of 500 x 300, r = 15,
    240 cell-centred buoys every 25 cells (4 groups)."""
    import torch
    from SWESimulators import CDKLM16Ensemble, Common, ETKFOcean
    import bench
    print("case double jet full size", flush=True)
    nx, ny, dx = 500, 300, 2000.0
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, bench._double_jet_members(100), courant_number=0.7)
    ys, xs = np.mgrid[0:ny:25, 0:nx:25]
    pos = np.column_stack([(xs.ravel() + 0.5) * dx, (ys.ravel() + 0.5) * dx])
    rs = np.random.RandomState(8)
    obs = np.column_stack([pos, 5.0 * rs.randn(len(pos), 2)])
    torch.cuda.synchronize()
    X_f = np.stack([np.stack(ens.download(k, interior_domain_only=True)) for k in range(100)]).astype(np.float32)
    R2 = np.diag([1.0, 1.0])
    want = E.letkf(X_f, obs, pos, R2, 15.0, dx, dx)
    et = ETKFOcean.ETKFOcean(ens, R2)
    et.LETKF(obs, r_factor=15.0)
    torch.cuda.synchronize()
    assert len(et.groups) == 4
    got = np.stack([np.stack(ens.download(k, interior_domain_only=True)) for k in range(100)])
    for f in range(3):
        incr = np.max(np.abs(want[:, f] - X_f[:, f]))
        err = np.max(np.abs(got[:, f] - want[:, f]))
        print("field", f, "max error / max increment", err / incr, flush=True)
        assert err <= 1e-5 * incr
    ens.cleanUp()
