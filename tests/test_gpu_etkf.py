"""Global ETKF analysis on the batched ensemble state against the float64
oracle (oracle/etkf_np.py, the reference's ETKFOcean.py:74-301), with the
observations taken by the batched observation kernel and checked against the
drifter oracle."""

import numpy as np
import pytest

from oracle import drifters_np as DR
from oracle import etkf_np as E

pytestmark = pytest.mark.gpu


def _members(N, nx, ny, seed):
    rs = np.random.RandomState(seed)
    j, i = np.mgrid[0:ny + 5, 0:nx + 5]
    Hi = (40 + 4 * np.sin(i / 6.0) + 3 * np.cos(j / 5.0)).astype(np.float32)
    base = 0.2 * np.exp(-((i[:-1, :-1] - nx / 2) ** 2 + (j[:-1, :-1] - ny / 2) ** 2) / 60.0)
    out = []
    for m in range(N):
        eta = (base + 0.02 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
        hu = (0.5 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
        hv = (0.5 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
        out.append(dict(nx=nx, ny=ny, dx=400.0, dy=400.0, g=9.81, f=1e-4, r=0.0, bc=(2, 2, 2, 2),
                        eta=eta, hu=hu, hv=hv, Hi=Hi))
    return out


@pytest.mark.parametrize("N,Nd", [(16, 6), (40, 25)])
def test_etkf_matches_oracle(N, Nd):
    import torch
    from SWESimulators import CDKLM16Ensemble, Common, ETKFOcean
    nx, ny = 48, 40
    ms = _members(N, nx, ny, seed=N)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.7)
    ens.step_n(2)
    rs = np.random.RandomState(7)
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
    want = E.analysis(X_f, HX_f, truth, R2)
    et = ETKFOcean.ETKFOcean(ens, R2)
    et.ETKF(obs)
    torch.cuda.synchronize()
    got = np.stack([np.concatenate([a.ravel() for a in ens.download(k)]) for k in range(N)], axis=1)
    incr = want - X_f.mean(axis=1)[:, None]
    P = (ny + 4) * (nx + 4)
    for f, name in enumerate(("eta", "hu", "hv")):
        sl = slice(f * P, (f + 1) * P)
        err = np.max(np.abs(got[sl] - want[sl])) / np.max(np.abs(incr[sl]))
        print(name, "max error / max analysis anomaly", err, flush=True)
        assert err <= 1e-5
    # the analysis pulls the observed transports towards the observations
    HX_a = np.zeros((2 * Nd, N))
    for e in range(N):
        _, hu, hv = ens.download(e)
        o = DR.observe_flow(pos, hu, hv, 400.0, 400.0, Nd)
        HX_a[:Nd, e], HX_a[Nd:, e] = o[:, 0], o[:, 1]
    assert np.linalg.norm(HX_a.mean(axis=1) - truth) < np.linalg.norm(HX_f.mean(axis=1) - truth)
