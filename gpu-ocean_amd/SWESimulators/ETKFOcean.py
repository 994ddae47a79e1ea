"""Global ETKF analysis on the batched ensemble state
(gpu_ocean/SWESimulators/ETKFOcean.py:34-301).

The reference downloads every member, assembles X_f (n_state x N_e) in
float64 on the host, and uploads every member again
(ETKFOcean.py:143-191, 286-301).  Here the ensemble state stays where
CDKLM16Ensemble keeps it -- N members of three (ny+4) x (nx+4) planes back
to back, i.e. an [N, 3, ny+4, nx+4] float32 tensor in HBM:

  * H X_f: the batched observation kernel (observeUnderlyingFlow for every
    member at the true state's drifter positions, one launch);
  * the ensemble-space algebra (2 N_d observations, N x N matrices) in
    float64 on the host, as the reference does (ETKFOcean.py:232-284);
  * the state update X_a = mean + (X_f - mean) T with T = W + w 1^T, one
    library GEMM of [N x N]^T by [N x n_state] on the device, written back in
    place.  W is the symmetric square root of (N-1) P and w = P HX'^T R^-1 D,
    so that X' w = K D of the reference.
"""

from __future__ import annotations

import numpy as np

from . import Common


class ETKFOcean(object):

    def __init__(self, ensemble, observation_cov, inflation_factor=1.0, scale_w=1.0):
        """ensemble: a CDKLM16Ensemble whose members share nx, ny (and whose
        drifters are enabled, for the observation kernel); observation_cov:
        the 2x2 covariance of one (hu, hv) observation."""
        self.ensemble = ensemble
        self.N_e = ensemble.n
        assert len(set(zip(ensemble.nx, ensemble.ny))) == 1, "the ETKF needs members of one size"
        self.n_i = ensemble.ny[0] + 4
        self.n_j = ensemble.nx[0] + 4
        self.ghost_cells_x = self.ghost_cells_y = 2
        self.plane = ensemble.planes[0]
        assert all(int(ensemble.state_off[k]) == 3 * self.plane * k for k in range(self.N_e))
        self.R = np.asarray(observation_cov, dtype=np.float64).reshape(2, 2)
        self.inflation_factor = inflation_factor
        self.scale_w = scale_w
        self.last = {}

    def state_matrix(self):
        """[N_e, 3 (ny+4)(nx+4)] view of the ensemble's current state."""
        return self.ensemble.state[0][:self.N_e * 3 * self.plane].view(self.N_e, 3 * self.plane)

    def _constructRinv(self, n_drifters):
        Nd = n_drifters
        R = np.zeros((2 * Nd, 2 * Nd))
        i = np.arange(Nd)
        R[i, i] = self.R[0, 0]
        R[Nd + i, Nd + i] = self.R[1, 1]
        R[i, Nd + i] = self.R[0, 1]
        R[Nd + i, i] = self.R[1, 0]
        return np.linalg.inv(R)

    def ETKF(self, observations, obs_positions=None):
        """Analysis of the whole ensemble in place.

        observations: (N_d, 4) [x, y, hu, hv] of the true state
        (observeTrueState, BaseOceanStateEnsemble.py:179-190), on the host.
        obs_positions: device (N_d, 2) or (N_d+1, 2) positions to observe the
        members at; default: the x, y columns of ``observations``."""
        torch = Common._torch()
        ens = self.ensemble
        st = ens.gpu_stream
        obs = np.asarray(observations, dtype=np.float64)
        Nd = obs.shape[0]
        assert Nd == ens.num_drifters, "observations must match the ensemble's drifter count"
        N = self.N_e
        dev = ens.state[0].device
        with torch.cuda.stream(st):
            if obs_positions is None:
                obs_positions = torch.from_numpy(obs[:, :2].astype(np.float32)).to(dev)
            HX = ens.observe(obs_positions)                       # [N, Nd, 2] on the device
            HX_f = HX[:, :Nd, :].permute(2, 1, 0).reshape(2 * Nd, N).to("cpu").numpy().astype(np.float64)
            X = self.state_matrix()
            X_mean = X.mean(dim=0, keepdim=True)
            X_pert = X - X_mean
        # ensemble-space algebra (ETKFOcean.py:232-284), float64
        HX_mean = HX_f.mean(axis=1)
        HX_pert = HX_f - HX_mean[:, None]
        Rinv = self._constructRinv(Nd)
        y = np.concatenate([obs[:, 2], obs[:, 3]])
        D = y - HX_mean
        P = np.linalg.inv((N - 1) * np.eye(N) + HX_pert.T @ Rinv @ HX_pert)
        w = P @ (HX_pert.T @ (Rinv @ D))
        sigma, V = np.linalg.eigh((N - 1) * P)
        W = V @ np.diag(np.sqrt(np.real(sigma))) @ V.T
        T = W + w[:, None]
        self.last = {"P": P, "w": w, "W": W, "D": D}
        with torch.cuda.stream(st):
            Tt = torch.from_numpy(np.ascontiguousarray(T.T).astype(np.float32)).to(dev)
            X_a = torch.addmm(X_mean.expand(N, -1), Tt, X_pert)
            X.copy_(X_a)
        return self.last
