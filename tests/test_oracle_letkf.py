"""Pins of the LETKF oracle (oracle/etkf_np.py, the reference's
ETKFOcean.py:299-744) by what the mathematics fixes:

  * the Gaspari-Cohn function by its special values and smoothness at the
    knots (a wrong coefficient breaks one of them);
  * the local boxes by the coordinates they pick, near every edge;
  * the groups by a hand-worked case and by their defining property;
  * the local analysis by the global ETKF (itself pinned against the Kalman
    update in test_oracle_etkf.py): one buoy's analysis is the global ETKF of
    the box state, blended by the local weight;
  * sequential groups, disjoint boxes and inflation by closed forms."""

import numpy as np
import pytest

from oracle import etkf_np as E


def test_gaspari_cohn_special_values_and_smoothness():
    c = 3.0
    gc = lambda d: E.gaspari_cohn(np.asarray(d, dtype=np.float64), c)
    assert gc(0.0)[0] == 1.0
    assert abs(gc(c)[0] - 5.0 / 24.0) < 1e-14
    assert np.all(gc([2 * c, 2.5 * c, 10 * c]) == 0.0)
    # value and slope continuous at both knots: the second difference across
    # a knot is O(h^2) (a jump in value would make it O(1), in slope O(h))
    h = 1e-4
    for knot in (c, 2 * c):
        f = gc([knot - h, knot, knot + h])
        assert abs(f[0] - 2 * f[1] + f[2]) < 1e-6, (knot, f)
    # away from the knots the second difference is O(h^2) too (sanity of h)
    f = gc([0.5 * c - h, 0.5 * c, 0.5 * c + h])
    assert abs(f[0] - 2 * f[1] + f[2]) < 1e-6
    # monotone decreasing on (0, 2c)
    d = np.linspace(0, 2 * c, 400)
    assert np.all(np.diff(gc(d)) <= 1e-15)


@pytest.mark.parametrize("r", [2.0, 4.0, 5.5])
def test_local_weight_shape(r):
    W = E.local_weight_shape(r, 1.0, 1.0, 200, 200)
    R = int(np.ceil(1.5 * r))
    assert W.shape == (2 * R + 1, 2 * R + 1)
    assert W[R, R] == 1.0
    np.testing.assert_array_equal(W, W.T)
    np.testing.assert_array_equal(W, W[::-1, :])
    np.testing.assert_array_equal(W, W[:, ::-1])
    j, i = np.mgrid[0:2 * R + 1, 0:2 * R + 1]
    dist = np.hypot(i - R, j - R)
    assert np.all(W[dist > 1.5 * r] == 0.0)
    assert np.all(W[dist <= 1.5 * r] > 0.0)
    assert np.all(np.diff(W[R, R:]) <= 0)
    if float(r).is_integer():
        assert abs(W[R, R + int(r)] - 5.0 / 24.0) < 1e-14   # distance r dx: GC(1)
    np.testing.assert_array_equal(E.local_weight_shape(r, 1.0, 1.0, 200, 200, 2.5), 2.5 * W)


@pytest.mark.parametrize("px,py", [(20.3, 17.8), (1.2, 17.0), (38.6, 3.1), (0.4, 39.7), (39.9, 0.2)])
def test_local_indices_pick_the_periodic_box(px, py):
    nx, ny, r = 40, 42, 3.0
    R = int(np.ceil(1.5 * r))
    n = 2 * R + 1
    mask, xroll, yroll = E.local_indices(np.array([px, py]), r, 1.0, 1.0, nx, ny)
    assert mask.sum() == n * n
    J, I = np.mgrid[0:ny, 0:nx]
    box_x = np.roll(np.roll(I[mask].reshape(n, n), -yroll, 0), -xroll, 1)
    box_y = np.roll(np.roll(J[mask].reshape(n, n), -yroll, 0), -xroll, 1)
    cx, cy = int(np.round(px)), int(np.round(py))
    np.testing.assert_array_equal(box_x, np.tile((cx - R + np.arange(n)) % nx, (n, 1)))
    np.testing.assert_array_equal(box_y, np.tile(((cy - R + np.arange(n)) % ny)[:, None], (1, n)))


def test_local_box_centre_rounds_half_to_even():
    for p, c in ((10.5, 10), (11.5, 12), (10.49, 10), (10.51, 11)):
        mask, _, _ = E.local_indices(np.array([p, 20.0]), 1.0, 1.0, 1.0, 50, 50)
        cols = np.where(mask.any(axis=0))[0]
        assert cols.min() == c - 2 and cols.max() == c + 2


def test_groups_hand_case_and_separation():
    # buoys 0 and 1 are close, 2 is far: 1 moves (the column of the first
    # closest entry, row-major, is (0, 1) -> buoy 1)
    assert E.observation_groups(np.array([[0., 0.], [1., 0.], [30., 30.]]), 4, 1.0, 1.0, 64, 64) == [[0, 2], [1]]
    rs = np.random.RandomState(3)
    nx, ny, r = 120, 90, 3.0
    P = np.column_stack([rs.uniform(0, nx, 60), rs.uniform(0, ny, 60)])
    groups = E.observation_groups(P, r, 1.0, 1.0, nx, ny)
    assert sorted(k for g in groups for k in g) == list(range(60))
    assert len(groups) > 1
    for g in groups:
        for a in g:
            for b in g:
                if a != b:
                    d = np.abs(P[a] - P[b])
                    d = np.minimum(d, np.array([nx, ny]) - d)
                    assert np.hypot(*d) >= 2 * 1.5 * r


def _ensemble(N, nx, ny, seed):
    rs = np.random.RandomState(seed)
    X = rs.randn(N, 3, ny, nx).astype(np.float32)
    X[:, 0] = 0.1 * X[:, 0] + 2.0
    X[:, 1:] += np.float32(0.5)
    return X


@pytest.mark.parametrize("scale_w", [1.0, 2.0])
def test_one_buoy_is_the_global_etkf_of_its_box(scale_w):
    N, nx, ny, r = 12, 30, 26, 2.0
    X = _ensemble(N, nx, ny, 0)
    pos = np.array([[28.7, 1.4]])                       # box wraps in x and y
    obs = np.array([[pos[0, 0], pos[0, 1], 0.9, -0.3]])
    R2 = np.array([[0.2, 0.05], [0.05, 0.3]])
    got = E.letkf(X, obs, pos, R2, r, 1.0, 1.0, scale_w=scale_w)
    mask, _, _ = E.local_indices(pos[0], r, 1.0, 1.0, nx, ny)
    W = E.combined_weights(pos, r, 1.0, 1.0, nx, ny, scale_w)
    box = X[:, :, mask].astype(np.float64)              # (N, 3, n)
    Xcol = box.reshape(N, -1).T
    ix, iy = int(np.floor(pos[0, 0])), int(np.floor(pos[0, 1]))
    HX = np.stack([X[:, 1, iy, ix], X[:, 2, iy, ix]]).astype(np.float64)
    Xa = E.analysis(Xcol, HX, obs[0, 2:], R2).T.reshape(N, 3, -1)
    w = W[mask]
    want = (1 - w / np.maximum(w, 1)) * box + w * Xa
    np.testing.assert_allclose(got[:, :, mask], want, rtol=2e-6, atol=2e-6)
    np.testing.assert_array_equal(got[:, :, ~mask], X[:, :, ~mask])
    assert np.all(W[~mask] == 0)


def test_groups_apply_in_sequence_and_disjoint_boxes_commute():
    N, nx, ny, r = 10, 40, 40, 2.0
    X = _ensemble(N, nx, ny, 1)
    R2 = np.diag([0.3, 0.3])
    near = np.array([[10.2, 10.7, 0.4, 0.1], [11.9, 10.1, 0.2, 0.6]])   # one group each
    a = E.letkf(X, near, near[:, :2], R2, r, 1.0, 1.0)
    b = E.letkf(E.letkf(X, near[:1], near[:1, :2], R2, r, 1.0, 1.0), near[1:], near[1:, :2], R2, r, 1.0, 1.0)
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6)
    far = np.array([[5.2, 5.7, 0.4, 0.1], [25.9, 30.1, 0.2, 0.6]])      # one group, disjoint boxes
    assert len(E.observation_groups(far[:, :2], r, 1.0, 1.0, nx, ny)) == 1
    a = E.letkf(X, far, far[:, :2], R2, r, 1.0, 1.0)
    b = E.letkf(E.letkf(X, far[1:], far[1:, :2], R2, r, 1.0, 1.0), far[:1], far[:1, :2], R2, r, 1.0, 1.0)
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("inflation", [1.3, 0.0])
def test_inflation_scales_transport_anomalies_when_the_buoy_sees_no_spread(inflation):
    """With every member equal at the buoy cell, HX' = 0: the gain is zero,
    eta keeps its anomalies and hu, hv anomalies grow by the inflation factor
    (adaptive: sqrt(1 + D^T R^-1 D / (N-2)))."""
    N, nx, ny, r = 9, 24, 24, 2.0
    X = _ensemble(N, nx, ny, 2)
    pos = np.array([[12.5, 11.5]])
    ix, iy = 12, 11
    X[:, 1:, iy, ix] = X[0, 1:, iy, ix]
    obs = np.array([[pos[0, 0], pos[0, 1], 1.1, -0.4]])
    R2 = np.array([[0.5, 0.1], [0.1, 0.4]])
    got = E.letkf(X, obs, pos, R2, r, 1.0, 1.0)
    if inflation:
        got = E.letkf(X, obs, pos, R2, r, 1.0, 1.0, inflation=inflation)
        factor = inflation
    else:
        got = E.letkf(X, obs, pos, R2, r, 1.0, 1.0, inflation=0.0)
        D = obs[0, 2:] - X[0, 1:, iy, ix].astype(np.float64)
        factor = np.sqrt(1 + D @ np.linalg.inv(R2) @ D / (N - 2))
    c = (iy, 12)                                        # box centre cell: weight 1
    m = X[:, :, c[0], c[1]].astype(np.float64).mean(axis=0)
    anom = X[:, :, c[0], c[1]] - m
    np.testing.assert_allclose(got[:, 0, c[0], c[1]], X[:, 0, c[0], c[1]], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(got[:, 1:, c[0], c[1]], m[1:] + factor * anom[:, 1:], rtol=1e-5, atol=1e-5)
