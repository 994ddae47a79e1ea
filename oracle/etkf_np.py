"""NumPy restatement of gpu-ocean's global ETKF analysis
(gpu_ocean/SWESimulators/ETKFOcean.py:74-301), in float64.

TEST INFRASTRUCTURE ONLY: only ``tests/`` import this module; the product's
analysis (SWESimulators/ETKFOcean.py) never calls it.

Layout as the reference: X_f is (n_state, N_e) with columns
[eta.ravel(), hu.ravel(), hv.ravel()] of each member (ghost cells included),
HX_f is (2 N_d, N_e) with rows [hu_1..hu_D, hv_1..hv_D], y is (2 N_d,).
"""

from __future__ import annotations

import numpy as np


def construct_rinv(R2, n_drifters):
    """R for 2 N_d observations from the 2x2 per-drifter covariance, inverted
    (ETKFOcean.py:127-141)."""
    Nd = n_drifters
    R = np.zeros((2 * Nd, 2 * Nd))
    for l in range(Nd):
        R[l, l] = R2[0, 0]
        R[Nd + l, Nd + l] = R2[1, 1]
        R[l, Nd + l] = R2[0, 1]
        R[Nd + l, l] = R2[1, 0]
    return np.linalg.inv(R)


def analysis(X_f, HX_f, y, R2):
    """X_a (n_state, N_e) of the global ETKF (ETKFOcean.py:74-284)."""
    X_f = np.asarray(X_f, dtype=np.float64)
    HX_f = np.asarray(HX_f, dtype=np.float64)
    N = X_f.shape[1]
    Nd = HX_f.shape[0] // 2
    X_mean = X_f.mean(axis=1)
    X_pert = X_f - X_mean[:, None]
    HX_mean = HX_f.mean(axis=1)
    HX_pert = HX_f - HX_mean[:, None]
    Rinv = construct_rinv(np.asarray(R2, dtype=np.float64), Nd)
    D = np.asarray(y, dtype=np.float64) - HX_mean
    P = np.linalg.inv((N - 1) * np.eye(N) + HX_pert.T @ Rinv @ HX_pert)
    K = X_pert @ (P @ (HX_pert.T @ Rinv))
    Xa_mean = X_mean + K @ D
    sigma, V = np.linalg.eigh((N - 1) * P)
    Xa_pert = X_pert @ (V @ np.diag(np.sqrt(np.real(sigma))) @ V.T)
    return Xa_pert + Xa_mean[:, None]
