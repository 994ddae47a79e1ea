"""Pins of the ETKF oracle (oracle/etkf_np.py) against the textbook Kalman
update on the ensemble covariance, computed here through the state-space
formulas (a different route from the ETKF's ensemble-space algebra):

  P_f = X' X'^T / (N-1),  K = P_f H^T (H P_f H^T + R)^-1,
  mean_a = mean_f + K (y - H mean_f),  P_a = (I - K H) P_f,

with a linear observation operator H so that HX_f = H X_f exactly."""

import numpy as np
import pytest

from oracle import etkf_np as E


def _case(seed, n=30, N=12, Nd=4, r=(0.3, 0.5, 0.1)):
    rs = np.random.RandomState(seed)
    X = rs.randn(n, N) + rs.randn(n, 1) * 3
    H = np.zeros((2 * Nd, n))
    idx = rs.choice(n, size=2 * Nd, replace=False)
    H[np.arange(2 * Nd), idx] = 1.0                 # observe 2 Nd state entries
    H += 0.1 * rs.randn(2 * Nd, n) * (rs.rand(2 * Nd, n) < 0.2)
    R2 = np.array([[r[0], r[2]], [r[2], r[1]]])
    y = H @ X.mean(axis=1) + rs.randn(2 * Nd)
    return X, H, R2, y


def _full_R(R2, Nd):
    return np.linalg.inv(E.construct_rinv(R2, Nd))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mean_and_covariance_equal_the_kalman_update(seed):
    X, H, R2, y = _case(seed)
    N, Nd = X.shape[1], H.shape[0] // 2
    Xa = E.analysis(X, H @ X, y, R2)
    m = X.mean(axis=1)
    Xp = X - m[:, None]
    Pf = Xp @ Xp.T / (N - 1)
    R = _full_R(R2, Nd)
    K = Pf @ H.T @ np.linalg.inv(H @ Pf @ H.T + R)
    np.testing.assert_allclose(Xa.mean(axis=1), m + K @ (y - H @ m), rtol=1e-10, atol=1e-10)
    Xap = Xa - Xa.mean(axis=1)[:, None]
    np.testing.assert_allclose(Xap @ Xap.T / (N - 1), (np.eye(len(m)) - K @ H) @ Pf, rtol=1e-8, atol=1e-10)


def test_symmetric_square_root_keeps_zero_mean_and_mixes_members_minimally():
    """The analysis anomalies X' W (W symmetric, W 1 = 1) keep zero mean, and
    with R -> infinity the analysis is the forecast."""
    X, H, R2, y = _case(3)
    Xa = E.analysis(X, H @ X, y, R2)
    assert abs((Xa - Xa.mean(axis=1)[:, None]).sum(axis=1)).max() < 1e-9
    Xa_inf = E.analysis(X, H @ X, y, R2 * 1e12)
    np.testing.assert_allclose(Xa_inf, X, rtol=1e-6, atol=1e-6)


def test_rinv_block_layout():
    """R pairs (hu_l, hv_l) at rows l and N_d + l (ETKFOcean.py:127-141)."""
    R2 = np.array([[2.0, 0.5], [0.5, 1.0]])
    R = np.linalg.inv(E.construct_rinv(R2, 3))
    assert R[1, 1] == 2.0 and R[4, 4] == 1.0 and R[1, 4] == 0.5 and R[4, 1] == 0.5
    assert R[0, 1] == 0.0 and R[0, 4] == 0.0


@pytest.mark.parametrize("seed", [0, 4])
def test_enkf_mean_is_the_kalman_update_without_observation_noise(seed):
    """With Y_pert = 0 every member takes K (y - H x_e), K = P_f H^T (H P_f
    H^T + R)^-1 on the ensemble covariance: the mean moves by the Kalman
    update, and the anomalies become (I - K H) X'."""
    X, H, R2, y = _case(seed)
    N, Nd = X.shape[1], H.shape[0] // 2
    HX = H @ X
    inn = y[:, None] - HX
    Xa = E.enkf(X, HX, inn, np.zeros_like(inn), R2)
    m = X.mean(axis=1)
    Xp = X - m[:, None]
    Pf = Xp @ Xp.T / (N - 1)
    R = _full_R(R2, Nd)
    K = Pf @ H.T @ np.linalg.inv(H @ Pf @ H.T + R)
    np.testing.assert_allclose(Xa.mean(axis=1), m + K @ (y - H @ m), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(Xa - Xa.mean(axis=1)[:, None], (np.eye(len(m)) - K @ H) @ Xp,
                               rtol=1e-8, atol=1e-10)


def test_enkf_inflation_scales_the_prior_covariance():
    """inflation f: the gain is that of the covariance f^2 P_f."""
    X, H, R2, y = _case(2)
    N, Nd = X.shape[1], H.shape[0] // 2
    HX = H @ X
    inn = y[:, None] - HX
    f = 1.4
    Xa = E.enkf(X, HX, inn, np.zeros_like(inn), R2, inflation=f)
    Xp = X - X.mean(axis=1)[:, None]
    Pf = f * f * Xp @ Xp.T / (N - 1)
    K = Pf @ H.T @ np.linalg.inv(H @ Pf @ H.T + _full_R(R2, Nd))
    np.testing.assert_allclose(Xa, X + K @ inn, rtol=1e-9, atol=1e-9)
