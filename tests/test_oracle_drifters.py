"""Pins of the drifter oracle (oracle/drifters_np.py) and of the host-side
drifter logic (SWESimulators/BaseDrifterCollection.py), on the CPU.

The oracle's drift is checked against the textbook bilinear interpolation
(``sample_point``) through the closed form of the reference's cell lookup:
away from cell faces it samples one cell east and one cell north of the
drifter; at a multiple of dx it samples the drifter's own position.  The
host logic is checked with the reference's own expectations
(tests/dataAssimilation/BaseDrifterTest.py) on a NumPy-backed collection."""

import numpy as np
import pytest

from oracle import drifters_np as D
from SWESimulators import BaseDrifterCollection, Common

NX, NY, DX, DY = 20, 16, 100.0, 80.0


def _fields(rs, land=False):
    j, i = np.mgrid[0:NY + 4, 0:NX + 4]
    Hm = (30.0 + rs.rand(NY + 4, NX + 4)).astype(np.float32)
    eta = (0.1 * rs.randn(NY + 4, NX + 4)).astype(np.float32)
    hu = (5 * rs.randn(NY + 4, NX + 4)).astype(np.float32)
    hv = (5 * rs.randn(NY + 4, NX + 4)).astype(np.float32)
    return eta, hu, hv, Hm


def test_drift_samples_one_cell_east_and_north():
    """Velocity u = hu / (Hm + eta); the drift displacement equals the
    bilinear value at (x + dx, y + dy) off faces (driftKernels.cu:137-146)."""
    rs = np.random.RandomState(1)
    eta, hu, hv, Hm = _fields(rs)
    u = hu.astype(np.float64) / (Hm.astype(np.float64) + eta)
    v = hv.astype(np.float64) / (Hm.astype(np.float64) + eta)
    pos = np.column_stack([rs.uniform(2 * DX, (NX - 3) * DX, 40), rs.uniform(2 * DY, (NY - 3) * DY, 40)])
    pos = pos.astype(np.float32)
    dt = 3.0
    new = D.drift(pos, eta, hu, hv, Hm, NX, NY, DX, DY, dt, False, False)
    for k in range(40):
        x, y = float(pos[k, 0]), float(pos[k, 1])
        ue = D.sample_point(u, x + DX, y + DY, DX, DY)
        ve = D.sample_point(v, x + DX, y + DY, DX, DY)
        np.testing.assert_allclose(new[k, 0] - pos[k, 0], ue * dt, rtol=2e-4, atol=1e-4)
        np.testing.assert_allclose(new[k, 1] - pos[k, 1], ve * dt, rtol=2e-4, atol=1e-4)


def test_drift_at_a_face_samples_the_position_itself():
    rs = np.random.RandomState(2)
    eta, hu, hv, Hm = _fields(rs)
    u = hu.astype(np.float64) / (Hm.astype(np.float64) + eta)
    pos = np.array([[5 * DX, 4.3 * DY]], dtype=np.float32)
    new = D.drift(pos, eta, hu, hv, Hm, NX, NY, DX, DY, 1.0, False, False)
    ue = D.sample_point(u, 5 * DX, 4.3 * DY + DY, DX, DY)
    np.testing.assert_allclose(new[0, 0] - pos[0, 0], ue, rtol=1e-4, atol=1e-4)


def test_uniform_flow_wind_and_sensitivity():
    Hm = np.full((NY + 4, NX + 4), 20.0, np.float32)
    eta = np.zeros_like(Hm)
    hu = np.full_like(Hm, 20.0 * 0.5)       # u = 0.5
    hv = np.full_like(Hm, 20.0 * -0.25)     # v = -0.25
    pos = np.array([[333.0, 444.0], [1000.0, 700.0]], dtype=np.float32)
    new = D.drift(pos, eta, hu, hv, Hm, NX, NY, DX, DY, 4.0, False, False, sensitivity=2.0,
                  wind=(1.0, 3.0, -2.0, 2.0), wind_t=0.25, wind_drift_factor=0.1)
    # u = 0.5 + 0.1 (0.25*3 + 0.75*1) = 0.65, v = -0.25 + 0.1 (0.25*2 - 0.75*2) = -0.35
    # positions are float32: the displacement carries their rounding (ulp(1000) = 6e-5)
    np.testing.assert_allclose(new - pos, [[2 * 0.65 * 4, 2 * -0.35 * 4]] * 2, atol=1.3e-4)


def test_periodic_wrap_and_enforce_bc():
    Hm = np.full((NY + 4, NX + 4), 20.0, np.float32)
    eta = np.zeros_like(Hm)
    hu = np.full_like(Hm, 20.0 * 10)
    hv = np.full_like(Hm, 20.0 * -10)
    Lx, Ly = NX * DX, NY * DY
    pos = np.array([[Lx - 5.0, 3.0]], dtype=np.float32)
    new = D.drift(pos, eta, hu, hv, Hm, NX, NY, DX, DY, 1.0, True, True)
    np.testing.assert_allclose(new, [[Lx - 5.0 + 10 - Lx, 3.0 - 10 + Ly]], rtol=1e-6)
    walls = D.drift(pos, eta, hu, hv, Hm, NX, NY, DX, DY, 1.0, False, False)
    np.testing.assert_allclose(walls, [[Lx + 5.0, -7.0]], rtol=1e-6)
    w = D.enforce_bc(np.array([[-1.0, 2.0], [3.0, 9.5]]), 8.0, 9.0, True, True)
    np.testing.assert_allclose(w, [[7.0, 2.0], [3.0, 0.5]])


def test_observe_flow_floor_cell():
    rs = np.random.RandomState(3)
    _, hu, hv, _ = _fields(rs)
    pos = np.array([[250.0, 170.0], [0.0, 0.0], [NX * DX - 1, NY * DY - 1]], dtype=np.float32)
    obs = D.observe_flow(pos, hu, hv, DX, DY, 3)
    assert tuple(obs[0]) == (hu[2 + 2, 2 + 2], hv[2 + 2, 2 + 2])
    assert tuple(obs[1]) == (hu[2, 2], hv[2, 2])
    assert tuple(obs[2]) == (hu[NY + 1, NX + 1], hv[NY + 1, NX + 1])


# ---------------------------------------------------------------------------
# host logic on a NumPy-backed collection, with the reference's expectations
# ---------------------------------------------------------------------------
class _HostCollection(BaseDrifterCollection.BaseDrifterCollection):
    def __init__(self, n, observation_variance=0.01, boundaryConditions=Common.BoundaryConditions(),
                 domain_size_x=1.0, domain_size_y=1.0):
        super().__init__(n, observation_variance, boundaryConditions, domain_size_x, domain_size_y)
        self.positions = np.zeros((n + 1, 2))
        self.uniformly_distribute_drifters()

    def getDrifterPositions(self):
        return self.positions[:-1].copy()

    def getObservationPosition(self):
        return self.positions[-1].copy()

    def setDrifterPositions(self, p):
        self.positions[:-1] = p

    def setObservationPosition(self, p):
        self.positions[-1] = p

    def cleanUp(self):
        pass

    def enforceBoundaryConditions(self):
        pass


def _small(bc=(2, 2, 2, 2)):
    c = _HostCollection(3, 0.25, Common.BoundaryConditions(*bc))
    c.setDrifterPositions(np.array([[0.9, 0.9], [0.9, 0.1], [0.1, 0.9]]))
    c.setObservationPosition(np.array([0.1, 0.1]))
    return c


def test_default_placement():
    """BaseDrifterTest.test_default_constructor (:107-131)."""
    c = _HostCollection(6)
    want = [[1 / 6, 1 / 4], [3 / 6, 1 / 4], [5 / 6, 1 / 4], [1 / 6, 3 / 4], [3 / 6, 3 / 4], [5 / 6, 3 / 4]]
    np.testing.assert_allclose(c.getDrifterPositions(), want, atol=1e-6)
    assert c.getObservationPosition().tolist() == [0, 0]


def test_distances_and_innovations():
    """BaseDrifterTest.test_distances / test_innovations (:219-278)."""
    longDiag, longLine = np.sqrt(2 * 0.8 * 0.8), 0.8
    shortDiag, shortLine = np.sqrt(0.08), 0.2
    semiDiag = np.sqrt(0.2 ** 2 + 0.8 ** 2)
    cases = {(2, 2, 2, 2): ([shortDiag, shortLine, shortLine], [[.2, .2], [.2, 0], [0, .2]]),
             (1, 1, 1, 1): ([longDiag, longLine, longLine], [[-.8, -.8], [-.8, 0], [0, -.8]]),
             (1, 2, 1, 2): ([semiDiag, shortLine, longLine], [[.2, -.8], [.2, 0], [0, -.8]]),
             (2, 1, 2, 1): ([semiDiag, longLine, shortLine], [[-.8, .2], [-.8, 0], [0, .2]])}
    for bc, (dist, inn) in cases.items():
        c = _small(bc)
        np.testing.assert_allclose(c.getDistances(), dist, atol=1e-6)
        np.testing.assert_allclose(c.getInnovations(), inn, atol=1e-6)


def test_collection_mean_and_domain_size():
    """BaseDrifterTest.test_collection_mean / test_set_domain_size (:281-320)."""
    c = _small()
    np.testing.assert_allclose(c.getCollectionMean(), [1 - 0.1 / 3, 1 - 0.1 / 3], atol=1e-6)
    c.setBoundaryConditions(Common.BoundaryConditions(1, 1, 1, 1))
    np.testing.assert_allclose(c.getCollectionMean(), [1.9 / 3, 1.9 / 3], atol=1e-6)
    c.setBoundaryConditions(Common.BoundaryConditions(2, 1, 2, 1))
    np.testing.assert_allclose(c.getCollectionMean(), [1.9 / 3, 1 - 0.1 / 3], atol=1e-6)
    c = _small()
    c.setDomainSize(10.3, 5.4)
    np.testing.assert_allclose(c.getDistances(), [np.sqrt(1.28), 0.8, 0.8], atol=1e-6)


def test_weights_and_resample():
    c = _small()
    w = c.getGaussianWeight()
    assert abs(w.sum() - 1) < 1e-12 and w[1] == w[2] and w[0] < w[1]
    cw = c.getCauchyWeight(normalize=False)
    np.testing.assert_allclose(cw, 1.0 / (np.pi * 0.5 * (1 + c.getDistances() ** 2 / 0.25)))
    np.random.seed(4)
    c.resample([0, 0, 2], 0.0)
    np.testing.assert_allclose(c.getDrifterPositions(), [[0.9, 0.9], [0.9, 0.9], [0.1, 0.9]])
    c.resample([1, 1, 2], 1e-8)
    p = c.getDrifterPositions()
    assert np.allclose(p[0], [0.9, 0.9]) and np.allclose(p[1], [0.9, 0.9], atol=1e-3) and not np.array_equal(p[1], p[0])
    with pytest.raises(RuntimeError):
        c.resample([0, 1], 0.0)
