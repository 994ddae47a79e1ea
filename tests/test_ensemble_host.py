"""Host helpers of the batched ensemble (SWESimulators/CDKLM16Ensemble.py)
that mirror the reference's ensemble: the Gaussian particle weights
(BaseOceanStateEnsemble.getGaussianWeight, :216-261), checked by hand-worked
values and properties."""

import numpy as np

from SWESimulators.CDKLM16Ensemble import gaussian_weights


def test_gaussian_weights_hand_case():
    # two particles, one drifter, R = diag(1, 4): log w = -(du^2 + dv^2/4)/2
    inn = np.array([[[1.0, 2.0]], [[0.0, 0.0]]])
    R = np.diag([1.0, 4.0])
    w = gaussian_weights(inn, R)
    a = np.exp(-0.5 * (1.0 + 1.0))
    np.testing.assert_allclose(w, [a / (a + 1.0), 1.0 / (a + 1.0)], rtol=1e-15)
    raw = gaussian_weights(inn, R, normalize=False)
    norm = 1.0 / ((2 * np.pi) * np.sqrt(4.0))
    np.testing.assert_allclose(raw, [norm * a, norm], rtol=1e-15)


def test_gaussian_weights_properties():
    rs = np.random.RandomState(0)
    inn = rs.randn(7, 3, 2) * 50          # large innovations: the max shift keeps it finite
    R = np.array([[0.5, 0.1], [0.1, 0.3]])
    w = gaussian_weights(inn, R)
    assert abs(w.sum() - 1.0) < 1e-12 and np.all(w >= 0) and np.isfinite(w).all()
    # R_scale s is the covariance s R
    np.testing.assert_allclose(gaussian_weights(inn, R, R_scale=3.0), gaussian_weights(inn, 3.0 * R), rtol=1e-12)
    # scalar innovations use the variance R[0, 0]
    d = np.array([0.0, 1.0])
    np.testing.assert_allclose(gaussian_weights(d, R, normalize=False),
                               np.exp(-d ** 2 / (2 * 0.5)) / np.sqrt(2 * np.pi * 0.5), rtol=1e-15)
