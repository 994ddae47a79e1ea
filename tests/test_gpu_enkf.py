"""Stochastic EnKF on the batched ensemble state against the float64 oracle
(oracle/etkf_np.py enkf, the reference's EnKFOcean.EnKF, EnKFOcean.py:64-296),
with the observations taken by the batched observation kernel and checked
against the drifter oracle, and fixed observation perturbations."""

import numpy as np
import pytest

from oracle import drifters_np as DR
from oracle import etkf_np as E

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,Nd,infl", [(16, 6, 1.0), (40, 25, 1.2)])
def test_enkf_matches_oracle(N, Nd, infl):
    import torch
    from SWESimulators import CDKLM16Ensemble, Common, EnKFOcean
    from test_gpu_etkf import _members
    print("case enkf", N, Nd, infl, flush=True)
    nx, ny = 48, 40
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, _members(N, nx, ny, seed=N + 1), rk_order=2, courant_number=0.7)
    ens.step_n(2)
    rs = np.random.RandomState(17)
    pos = np.column_stack([rs.uniform(0, nx * 400.0, Nd), rs.uniform(0, ny * 400.0, Nd)]).astype(np.float32)
    ens.enable_drifters(Nd)
    torch.cuda.synchronize()
    states = [ens.download(k) for k in range(N)]
    X_f = np.stack([np.concatenate([a.ravel() for a in s]) for s in states], axis=1).astype(np.float64)
    HX_f = np.zeros((2 * Nd, N))
    for e, (eta, hu, hv) in enumerate(states):
        o = DR.observe_flow(pos, hu, hv, 400.0, 400.0, Nd)
        HX_f[:Nd, e], HX_f[Nd:, e] = o[:, 0], o[:, 1]
    truth = HX_f.mean(axis=1) + 0.3 * rs.randn(2 * Nd)
    obs = np.column_stack([pos[:, 0], pos[:, 1], truth[:Nd], truth[Nd:]])
    R2 = np.array([[0.04, 0.01], [0.01, 0.09]])
    R = np.linalg.inv(E.construct_rinv(R2, Nd))
    Y = rs.multivariate_normal(np.zeros(2 * Nd), R, N).T
    want = E.enkf(X_f, HX_f, truth[:, None] - HX_f, Y, R2, inflation=infl)
    ek = EnKFOcean.EnKFOcean(ens, R2, inflation_factor=infl)
    ek.EnKF(obs, Y_pert=Y)
    torch.cuda.synchronize()
    got = np.stack([np.concatenate([a.ravel() for a in ens.download(k)]) for k in range(N)], axis=1)
    incr = want - X_f
    P = (ny + 4) * (nx + 4)
    for f, name in enumerate(("eta", "hu", "hv")):
        sl = slice(f * P, (f + 1) * P)
        err = np.max(np.abs(got[sl] - want[sl])) / np.max(np.abs(incr[sl]))
        print(name, "max error / max increment", err, flush=True)
        assert err <= 1e-5
