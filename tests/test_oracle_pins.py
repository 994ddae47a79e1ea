"""Pins of the vectorised oracle that do not depend on the golden files:
the literal thread-block emulation, exact closed forms and the symmetries and
invariants the scheme must keep."""

import numpy as np
import pytest

from oracle import cdklm16_np as O
from oracle.cdklm16_tiled import tiled_reference_substep
import helpers as H


_case = H.random_case


def WALL_SIDES(p):
    return any(b == O.WALL for b in p.bc)


TILED = [
    (dict(nx=11, ny=9, bc=(1, 1, 1, 1)), (3, 3)),
    (dict(nx=11, ny=9, bc=(2, 2, 2, 2)), (3, 3)),
    (dict(nx=11, ny=9, bc=(2, 1, 2, 1)), (3, 3)),
    (dict(nx=11, ny=9, bc=(1, 2, 1, 2), fvar=True), (3, 3)),
    (dict(nx=11, ny=9, bc=(1, 1, 1, 1), angvar=True, fvar=True), (3, 3)),
    (dict(nx=13, ny=10, bc=(1, 1, 1, 1), land=True), (5, 4)),
    (dict(nx=13, ny=10, bc=(1, 1, 1, 1), shallow=True, r=0.003), (5, 4)),
    (dict(nx=13, ny=10, bc=(2, 2, 2, 2), shallow=True, land=True, wind=(1e-4, 2e-5)), (4, 3)),
]


@pytest.mark.parametrize("k", range(len(TILED)))
@pytest.mark.parametrize("rk", [0, 1])
def test_vectorised_equals_literal_block_emulation(k, rk):
    cfg, block = TILED[k]
    p, Hi, Hm, eta, hu, hv = _case(seed=k, **cfg)
    # with wall sides, the block shape avoids the reference's
    # (n-1) % block == 0 quirk
    if WALL_SIDES(p):
        assert (p.nx - 1) % block[0] and (p.ny - 1) % block[1]
    eo, uo, vo = (eta * 0.9).astype(np.float32), (hu * 1.1).astype(np.float32), (hv * 0.7).astype(np.float32)
    a = O.cdklm_substep(0.5, rk, eta, hu, hv, eo, uo, vo, Hi, Hm, p, wind_t=0.3)
    b = tiled_reference_substep(0.5, rk, eta, hu, hv, eo, uo, vo, Hi, Hm, p, block=block, wind_t=0.3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_block_seam_quirk_is_confined():
    """With walls and (ny-1) % bh == 0 the reference reads the unmirrored
    ghost row in the block that ends at interior row ny; only that row (and
    the analogous column) can differ from the mirrored formulation."""
    p, Hi, Hm, eta, hu, hv = _case(13, 10, (1, 1, 1, 1), 3)
    block = (4, 3)          # (13-1) % 4 == 0 and (10-1) % 3 == 0
    a = O.cdklm_substep(0.5, 0, eta, hu, hv, eta, hu, hv, Hi, Hm, p)
    b = tiled_reference_substep(0.5, 0, eta, hu, hv, eta, hu, hv, Hi, Hm, p, block=block)
    diff = np.zeros(a[0].shape, bool)
    for x, y in zip(a, b):
        diff |= x != y
    assert diff.any()
    rows, cols = np.nonzero(diff)
    assert set(rows.tolist()) | set(cols.tolist()) <= {p.ny, p.nx} | set(range(2, p.ny + 2))
    assert all(r == p.ny or c == p.nx for r, c in zip(rows, cols))


def test_max_dt_matches_host_closed_form_exactly():
    """CDKLM16_test.py:147-150 asserts updateDt() == _getMaxTimestepHost()."""
    for seed in range(4):
        p, Hi, Hm, eta, hu, hv = _case(37, 29, (1, 1, 1, 1), seed)
        m = O.max_dt(eta, hu, hv, Hm, p)
        dt_dev = 0.8 * float(m)
        assert dt_dev == O.max_dt_host(eta, hu, hv, Hm, p, 0.8)


def _flat_case(nx, ny, bc, seed, f=0.0, dx=200.0, dy=200.0, H=60.0):
    rng = np.random.default_rng(seed)
    p = O.Params(nx=nx, ny=ny, dx=dx, dy=dy, bc=bc, f=np.float32(f))
    Hi = np.full((ny + 5, nx + 5), H, np.float32)
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    eta[2:-2, 2:-2] = (0.2 * rng.standard_normal((ny, nx))).astype(np.float32)
    hu = np.zeros_like(eta)
    hv = np.zeros_like(eta)
    hu[2:-2, 2:-2] = rng.standard_normal((ny, nx)).astype(np.float32)
    hv[2:-2, 2:-2] = rng.standard_normal((ny, nx)).astype(np.float32)
    return p, Hi, eta, hu, hv


def test_mirror_symmetry_exact():
    """x -> -x with hu -> -hu.  East/west periodic: at a wall face the two
    reconstructed velocities cancel exactly and the F.z upwinding
    (CDKLM16_kernel.cu:137, ``Qm.y + Qp.y > 0``) then takes the p side, which is
    the interior at a west wall but the ghost at an east wall, so walls in x
    are not mirror symmetric in the reference scheme."""
    p, Hi, eta, hu, hv = _flat_case(24, 19, (1, 2, 1, 2), 1)
    a = O.OracleSim(eta, hu, hv, Hi, p, dt=0.9)
    b = O.OracleSim(eta[:, ::-1].copy(), -hu[:, ::-1], hv[:, ::-1].copy(), Hi, p, dt=0.9)
    a.step_n(5)
    b.step_n(5)
    e1, u1, v1 = a.download()
    e2, u2, v2 = b.download()
    I = (slice(2, -2), slice(2, -2))
    assert np.array_equal(e1[I], e2[I][:, ::-1])
    assert np.array_equal(u1[I], -u2[I][:, ::-1])
    assert np.array_equal(v1[I], v2[I][:, ::-1])


def test_transpose_symmetry_exact():
    p, Hi, eta, hu, hv = _flat_case(21, 21, (1, 1, 1, 1), 2)
    a = O.OracleSim(eta, hu, hv, Hi, p, dt=0.9)
    b = O.OracleSim(eta.T.copy(), hv.T.copy(), hu.T.copy(), Hi.T.copy(), p, dt=0.9)
    a.step_n(5)
    b.step_n(5)
    e1, u1, v1 = a.download()
    e2, u2, v2 = b.download()
    I = (slice(2, -2), slice(2, -2))
    assert np.array_equal(e1[I], e2[I].T)
    assert np.array_equal(u1[I], v2[I].T)
    assert np.array_equal(v1[I], u2[I].T)


def test_periodic_roll_equivariance_exact():
    p, Hi, eta, hu, hv = _flat_case(20, 16, (2, 2, 2, 2), 3, f=1e-4)
    rng = np.random.default_rng(3)
    Hi = (50 + rng.random((p.ny + 5, p.nx + 5))).astype(np.float32)
    Hi[-3:, :] = Hi[2:5, :]
    Hi[:2, :] = Hi[p.ny:p.ny + 2, :]
    a = O.OracleSim(eta, hu, hv, Hi, p, dt=0.9)

    def roll(x, sy, sx, corner=False):
        y = x.copy()
        n0 = y.shape[0] - (5 if corner else 4)
        n1 = y.shape[1] - (5 if corner else 4)
        core = x[2:2 + n0, 2:2 + n1]
        y[2:2 + n0, 2:2 + n1] = np.roll(core, (sy, sx), axis=(0, 1))
        return y
    Hr = O.bathymetry_halo(Hi, p)
    Hi_r = Hr.copy()
    Hi_r[2:p.ny + 2, 2:p.nx + 2] = np.roll(Hr[2:p.ny + 2, 2:p.nx + 2], (5, 7), axis=(0, 1))
    b = O.OracleSim(roll(eta, 5, 7), roll(hu, 5, 7), roll(hv, 5, 7), Hi_r, p, dt=0.9)
    a.step_n(4)
    b.step_n(4)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(np.roll(x[2:-2, 2:-2], (5, 7), axis=(0, 1)), y[2:-2, 2:-2])


def test_lake_at_rest_is_well_balanced():
    """eta = 0, hu = hv = 0 over rough bathymetry stays at rest to round-off
    (the defining well-balanced property, CDKLM16_kernel.cu:7-11)."""
    rng = np.random.default_rng(4)
    nx, ny = 30, 26
    Hi = (20 + 15 * rng.random((ny + 5, nx + 5))).astype(np.float32)
    for bc in ((1, 1, 1, 1), (2, 2, 2, 2)):
        p = O.Params(nx=nx, ny=ny, dx=500.0, dy=500.0, bc=bc, f=np.float32(1.2e-4))
        z = np.zeros((ny + 4, nx + 4), np.float32)
        sim = O.OracleSim(z, z, z, Hi, p, dt=2.0)
        sim.step_n(20)
        e, u, v = sim.download()
        assert np.max(np.abs(e[2:-2, 2:-2])) < 1e-5
        assert np.max(np.abs(u[2:-2, 2:-2])) < 1e-3
        assert np.max(np.abs(v[2:-2, 2:-2])) < 1e-3
        # a perturbed lake moves by orders of magnitude more
        e0 = z.copy()
        e0[ny // 2 + 2, nx // 2 + 2] = 0.1
        s2 = O.OracleSim(e0, z, z, Hi, p, dt=2.0)
        s2.step_n(20)
        assert np.max(np.abs(s2.download()[1])) > 100 * max(np.max(np.abs(u)), 1e-9)


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2), (1, 2, 1, 2)])
def test_mass_conservation(bc):
    rng = np.random.default_rng(5)
    nx, ny = 32, 28
    Hi = (40 + 10 * rng.random((ny + 5, nx + 5))).astype(np.float32)
    p = O.Params(nx=nx, ny=ny, dx=400.0, dy=400.0, bc=bc, f=np.float32(1.2e-4))
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    eta[2:-2, 2:-2] = (0.3 * rng.standard_normal((ny, nx))).astype(np.float32)
    z = np.zeros_like(eta)
    sim = O.OracleSim(eta, z, z, Hi, p, dt=3.0)
    m0 = np.sum(sim.download()[0][2:-2, 2:-2].astype(np.float64))
    sim.step_n(30)
    m1 = np.sum(sim.download()[0][2:-2, 2:-2].astype(np.float64))
    assert abs(m1 - m0) < 1e-4 * nx * ny * 0.3 * 1e-3


def test_fp32_tracks_fp64():
    from SWESimulators import BathymetryAndICs as B
    c = B.config1(nx=48, ny=40)
    p = H.oracle_params(c)
    a = O.OracleSim(c.eta, c.hu, c.hv, c.Hi, p, dt=-1, dtype=np.float32)
    b = O.OracleSim(c.eta, c.hu, c.hv, c.Hi, p, dt=a.dt, dtype=np.float64)
    dt = np.float32(a.dt)
    a.step_n(20, dt)
    b.step_n(20, dt)
    for x, y in zip(a.download(), b.download()):
        assert H.max_rel(x[2:-2, 2:-2], y[2:-2, 2:-2]) < 1e-4
