"""Drifters and the observation operator on the GPU against the oracle
(oracle/drifters_np.py), bit for bit: single domain through the C ABI, the
GPUDrifterCollection class on the reference's BaseDrifterTest expectations,
drifters attached to a CDKLM16 simulator, and the batched ensemble forms."""

import numpy as np
import pytest

from oracle import drifters_np as D

pytestmark = pytest.mark.gpu

NX, NY, DX, DY = 40, 28, 150.0, 120.0


def _say(*a):
    print(*a, flush=True)


def _torch():
    import torch
    return torch


def _fields(rs, nx=NX, ny=NY):
    Hm = (30.0 + rs.rand(ny + 4, nx + 4)).astype(np.float32)
    eta = (0.1 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
    hu = (5 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
    hv = (5 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
    return eta, hu, hv, Hm


def _positions(rs, n, nx=NX, ny=NY, dx=DX, dy=DY):
    p = np.column_stack([rs.uniform(0, nx * dx, n), rs.uniform(0, ny * dy, n)])
    # faces, corners and the domain edges
    special = [[5 * dx, 7.5 * dy], [0.0, 0.0], [nx * dx, ny * dy], [3 * dx, 3 * dy], [0.5 * dx, (ny - 0.25) * dy]]
    return np.concatenate([p, special]).astype(np.float32)


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2), (2, 1, 2, 1)])
@pytest.mark.parametrize("wind_factor", [0.0, 0.02])
def test_drift_matches_oracle_bit_exact(bc, wind_factor):
    from SWESimulators import Common, _gocean as G
    torch = _torch()
    lib = G.load()
    _say("drift", bc, wind_factor)
    rs = np.random.RandomState(1)
    eta, hu, hv, Hm = _fields(rs)
    pos = _positions(rs, 300)
    n = pos.shape[0] - 1
    d = Common.make_domain(NX, NY, DX, DY, 9.81, 0.0, 1.3, 2, bc)
    t = [torch.from_numpy(a).cuda() for a in (eta, hu, hv, Hm)]
    P = torch.from_numpy(pos).cuda()
    w = G.Wind()
    w.t, w.x0, w.x1, w.y0, w.y1 = 0.3, 1.5, -2.0, 0.7, 3.0
    want = pos
    for step in range(4):
        G.check(lib.gocean_drift(d, n, P.data_ptr(), t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(),
                                 t[3].data_ptr(), 7.0, 1.3, w, wind_factor, 2, 2, None), "drift")
        want = D.drift(want, eta, hu, hv, Hm, NX, NY, DX, DY, 7.0, bc[0] == 2, bc[1] == 2, sensitivity=1.3,
                       wind=(1.5, -2.0, 0.7, 3.0), wind_t=0.3, wind_drift_factor=wind_factor)
        torch.cuda.synchronize()
        got = P.cpu().numpy()
        assert np.array_equal(got, want), (step, np.max(np.abs(got - want)))


def test_observe_and_enforce_bc_match_oracle():
    from SWESimulators import Common, _gocean as G
    torch = _torch()
    lib = G.load()
    rs = np.random.RandomState(2)
    eta, hu, hv, Hm = _fields(rs)
    pos = _positions(rs, 100)
    n = pos.shape[0] - 1
    d = Common.make_domain(NX, NY, DX, DY, 9.81, 0.0, 1.3, 2, (2, 2, 2, 2))
    P = torch.from_numpy(pos).cuda()
    thu, thv = torch.from_numpy(hu).cuda(), torch.from_numpy(hv).cuda()
    obs = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    G.check(lib.gocean_observe_flow(d, n, P.data_ptr(), thu.data_ptr(), thv.data_ptr(), obs.data_ptr(), 2, 2, None))
    torch.cuda.synchronize()
    assert np.array_equal(obs.cpu().numpy(), D.observe_flow(pos, hu, hv, DX, DY, n))
    out = (pos * np.float32(1.7) - np.float32(600.0)).astype(np.float32)
    P.copy_(torch.from_numpy(out))
    G.check(lib.gocean_drifters_enforce_bc(NX * DX, NY * DY, 1, 1, n, P.data_ptr(), None))
    torch.cuda.synchronize()
    assert np.array_equal(P.cpu().numpy(), D.enforce_bc(out, NX * DX, NY * DY, True, True))


def test_collection_as_reference_tests():
    """GPUDrifter_test / BaseDrifterTest expectations (:107-320) through the
    device-backed class (BoundaryConditions.get() raises in the reference,
    Common.py:858-859, so the codes are compared through codes())."""
    from SWESimulators import Common, GPUDrifterCollection
    ctx = Common.HIPContext(0)
    c = GPUDrifterCollection.GPUDrifterCollection(ctx, 6)
    want = [[1 / 6, 1 / 4], [3 / 6, 1 / 4], [5 / 6, 1 / 4], [1 / 6, 3 / 4], [3 / 6, 3 / 4], [5 / 6, 3 / 4]]
    np.testing.assert_allclose(c.getDrifterPositions(), want, atol=1e-6)
    assert c.getObservationPosition().tolist() == [0, 0]
    assert c.getObservationVariance() == 0.01 and tuple(c.getBoundaryConditions().codes()) == (1, 1, 1, 1)
    s = GPUDrifterCollection.GPUDrifterCollection(ctx, 3, 0.25, boundaryConditions=Common.BoundaryConditions(2, 2, 2, 2))
    s.setDrifterPositions(np.array([[0.9, 0.9], [0.9, 0.1], [0.1, 0.9]]))
    s.setObservationPosition(np.array([0.1, 0.1]))
    np.testing.assert_allclose(s.getDrifterPositions(), [[0.9, 0.9], [0.9, 0.1], [0.1, 0.9]], atol=1e-6)
    np.testing.assert_allclose(s.getObservationPosition(), [0.1, 0.1], atol=1e-6)
    np.testing.assert_allclose(s.getDistances(), [np.sqrt(0.08), 0.2, 0.2], atol=1e-6)
    np.testing.assert_allclose(s.getCollectionMean(), [1 - 0.1 / 3, 1 - 0.1 / 3], atol=1e-6)
    cp = s.copy()
    np.testing.assert_allclose(cp.getDrifterPositions(), s.getDrifterPositions())
    s.setDrifterPositions(np.array([[0.5, 0.5], [0.9, 0.1], [0.1, 0.9]]))
    np.testing.assert_allclose(cp.getDrifterPositions()[0], [0.9, 0.9], atol=1e-6)
    s.setDrifterPositions(np.array([[-0.25, 0.5], [1.25, 0.1], [0.1, 1.5]]))
    s.enforceBoundaryConditions()
    np.testing.assert_allclose(s.getDrifterPositions(), [[0.75, 0.5], [0.25, 0.1], [0.1, 0.5]], atol=1e-6)
    big = GPUDrifterCollection.GPUDrifterCollection(ctx, 1000, domain_size_x=10.3, domain_size_y=5.4)
    big.initializeUniform()
    p = big.getDrifterPositions()
    assert p[:, 0].min() >= 0 and p[:, 0].max() <= 10.3 and p[:, 1].min() >= 0 and p[:, 1].max() <= 5.4


def test_simulator_moves_attached_drifters():
    """CDKLM16.step with drifters: after every step the drifters move by the
    velocity of the state at the end of that step (CDKLM16.py:588,600-610)."""
    from SWESimulators import CDKLM16, Common, GPUDrifterCollection
    nx, ny, dx, dy = 48, 36, 500.0, 500.0
    j, i = np.mgrid[0:ny + 5, 0:nx + 5]
    Hi = (40 + 5 * np.sin(i / 5.0) + 3 * np.cos(j / 7.0)).astype(np.float32)
    eta = (0.3 * np.exp(-((i[:-1, :-1] - nx / 2) ** 2 + (j[:-1, :-1] - ny / 2) ** 2) / 40.0)).astype(np.float32)
    z = np.zeros((ny + 4, nx + 4), dtype=np.float32)
    ctx = Common.HIPContext(0)
    bc = Common.BoundaryConditions(2, 2, 2, 2)
    sim = CDKLM16.CDKLM16(ctx, eta, z, z, Hi, nx, ny, dx, dy, 5.0, 9.81, 1e-4, 0.0, boundary_conditions=bc)
    dr = GPUDrifterCollection.GPUDrifterCollection(ctx, 25, boundaryConditions=bc,
                                                   domain_size_x=nx * dx, domain_size_y=ny * dy)
    sim.attachDrifters(dr)
    Hm = sim.downloadBathymetry()[1]
    pos = np.concatenate([dr.getDrifterPositions(), [dr.getObservationPosition()]]).astype(np.float32)
    for step in range(5):
        sim.step(5.0)
        e, u, v = sim.download()
        pos = D.drift(pos, e, u, v, Hm, nx, ny, dx, dy, 5.0, True, True)
        got = np.concatenate([dr.getDrifterPositions(), [dr.getObservationPosition()]])
        assert np.array_equal(got, pos), step
    assert np.max(np.abs(pos[:-1] - np.concatenate([dr.getDrifterPositions()]))) == 0


def test_batched_drift_and_observe_equal_single():
    from SWESimulators import CDKLM16Ensemble, Common, _gocean as G
    import test_gpu_ensemble as TE
    torch = _torch()
    lib = G.load()
    ms = TE.members(seed=5)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.7)
    num = 37
    rs = np.random.RandomState(3)
    pos0 = np.stack([_positions(rs, num - 4, m["nx"], m["ny"], m["dx"], m["dy"]) for m in ms])
    ens.enable_drifters(num, positions=pos0, sensitivity=0.9)
    ens.step_n(3)
    torch.cuda.synchronize()
    dts = ens.dt_host()
    got = ens.drifter_positions.cpu().numpy()
    # single-domain drift of each member over a fresh ensemble run (same
    # states, same per-member dt)
    for k, m in enumerate(ms):
        P = torch.from_numpy(pos0[k]).cuda()
        w = ens.members_c[k].wind
        ens2 = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.7)
        for step in range(3):
            ens2._rk(1)
            torch.cuda.synchronize()
            e, hu, hv = ens2.views(k)
            Hm = ens2.Hm[ens2.hm_off[k]:ens2.hm_off[k + 1]]
            G.check(lib.gocean_drift(ens2.members_c[k].domain, num, P.data_ptr(), e.data_ptr(), hu.data_ptr(),
                                     hv.data_ptr(), Hm.data_ptr(), float(dts[k]), 0.9, w, 0.0, 2, 2,
                                     G.stream_handle(ens2.gpu_stream)), "drift")
        torch.cuda.synchronize()
        assert np.array_equal(P.cpu().numpy(), got[k]), k
        ens2.cleanUp()
    # observations: own positions and shared (true-state) positions
    obs_own = ens.observe().cpu().numpy().copy()
    truth = ens.drifter_positions[0].clone()
    obs_shared = ens.observe(truth).cpu().numpy().copy()
    torch.cuda.synchronize()
    for k in range(ens.n):
        e, hu, hv = (a.cpu().numpy() for a in ens.views(k))
        m = ms[k]
        assert np.array_equal(obs_own[k], D.observe_flow(got[k], hu, hv, m["dx"], m["dy"], num))
        assert np.array_equal(obs_shared[k], D.observe_flow(got[0], hu, hv, m["dx"], m["dy"], num))
