"""Stochastic ensemble Kalman filter on the batched ensemble state
(EnKFOcean.EnKF, gpu_ocean/SWESimulators/EnKFOcean.py:64-296).

The reference downloads every member into an (n_state x N) float64 matrix,
forms the ensemble-space algebra with perturbed observations, and uploads
every member again.  Here the state stays in HBM as the [N, 3 (ny+4)(nx+4)]
matrix CDKLM16Ensemble keeps:

  * H X_f: the batched observation kernel at the true drifter positions;
  * HPHT = infl^2/(N-1) HX' HX'^T, F = HPHT + R, the perturbed innovations
    D = d + Y and E = HX'^T F^-1 D (2 N_d and N x N matrices) in float64 on
    the host, as the reference does (:118-237);
  * X_a = X_f + infl^2/(N-1) X' E as one library GEMM on the device,
    written back in place (:289-293).
"""

from __future__ import annotations

import numpy as np

from . import Common


class EnKFOcean(object):

    def __init__(self, ensemble, observation_cov, inflation_factor=1.0, seed=None):
        """ensemble: a CDKLM16Ensemble whose members share nx, ny, with
        drifters enabled (for the observation kernel); observation_cov: the
        2 x 2 covariance of one (hu, hv) observation; seed: of the generator
        drawing the observation perturbations (the reference uses numpy's
        global generator)."""
        self.ensemble = ensemble
        self.N_e = ensemble.n
        assert len(set(zip(ensemble.nx, ensemble.ny))) == 1, "the EnKF needs members of one size"
        self.n_i = ensemble.ny[0] + 4
        self.n_j = ensemble.nx[0] + 4
        self.plane = ensemble.planes[0]
        assert all(int(ensemble.state_off[k]) == 3 * self.plane * k for k in range(self.N_e))
        self.R = np.asarray(observation_cov, dtype=np.float64).reshape(2, 2)
        self.inflation_factor = inflation_factor
        self.rng = np.random.default_rng(seed)
        self.last = {}

    def _giveR(self, n_drifters):
        """R for 2 N_d observations, rows [hu_1..hu_D, hv_1..hv_D] (:118-130)."""
        Nd = n_drifters
        R = np.zeros((2 * Nd, 2 * Nd))
        i = np.arange(Nd)
        R[i, i] = self.R[0, 0]
        R[Nd + i, Nd + i] = self.R[1, 1]
        R[i, Nd + i] = self.R[0, 1]
        R[Nd + i, i] = self.R[1, 0]
        return R

    def EnKF(self, observations, innovations=None, Y_pert=None):
        """Analysis of the whole ensemble in place.

        observations: (N_d, 4) [x, y, hu, hv] of the true drifters.
        innovations: (N, N_d, 2) y - H x of each member (getInnovations);
        default: from the members' observations at the true positions.
        Y_pert: (2 N_d, N) observation perturbations; default: drawn from
        N(0, R) (:212-219)."""
        torch = Common._torch()
        ens = self.ensemble
        st = ens.gpu_stream
        obs = np.asarray(observations, dtype=np.float64).reshape(-1, 4)
        Nd = obs.shape[0]
        N = self.N_e
        dev = ens.state[0].device
        with torch.cuda.stream(st):
            pos = torch.from_numpy(obs[:, :2].astype(np.float32)).to(dev)
            HX = ens.observe(pos)[:, :Nd, :].to("cpu").numpy().astype(np.float64)     # (N, Nd, 2)
        HX_f = HX.transpose(2, 1, 0).reshape(2 * Nd, N)
        if innovations is None:
            innovations = obs[None, :, 2:4] - HX
        d = np.asarray(innovations, dtype=np.float64).reshape(N, Nd, 2).transpose(2, 1, 0).reshape(2 * Nd, N)
        R = self._giveR(Nd)
        if Y_pert is None:
            Y_pert = self.rng.multivariate_normal(np.zeros(2 * Nd), R, N).T
        c = self.inflation_factor ** 2 / (N - 1)
        HXp = HX_f - np.nanmean(HX_f, axis=1)[:, None]
        F = c * HXp @ HXp.T + R
        E = HXp.T @ np.linalg.solve(F, d + np.asarray(Y_pert, dtype=np.float64))
        self.last = {"E": E, "F": F}
        with torch.cuda.stream(st):
            X = ens.state[0][:N * 3 * self.plane].view(N, 3 * self.plane)
            X_pert = X - X.mean(dim=0, keepdim=True)
            M = torch.from_numpy(np.ascontiguousarray((c * E).T).astype(np.float32)).to(dev)
            X.addmm_(M, X_pert)
        return self.last
