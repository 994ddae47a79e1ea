"""Two-stage IEWPF on the batched ensemble state against the oracle
(oracle/iewpf_np.py, the reference's IEWPFOcean.iewpf_2stage,
IEWPFOcean.py:195-329), and the batched local-SVD kernel against the
reference's golden data (iewpfRefData/pre-, postLocalSVDtoGlobal.dat)."""

import os

import numpy as np
import pytest

from oracle import iewpf_np as I

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DT, F, G, H, DX = 0.05, 0.05, 9.81, 10.0, 4.0
Q0 = 0.5 * DT * F / (G * H)


def _ensemble(N, n, factor, seed):
    from SWESimulators import CDKLM16Ensemble, Common
    rs = np.random.RandomState(seed)
    nx, ny = (n, n) if np.isscalar(n) else n
    j, i = np.mgrid[0:ny + 4, 0:nx + 4]
    members = []
    for _ in range(N):
        eta = (0.01 * np.sin(2 * np.pi * i / nx) * np.cos(2 * np.pi * j / ny) + 1e-3 * rs.randn(ny + 4, nx + 4))
        members.append(dict(nx=nx, ny=ny, dx=DX, dy=DX, g=G, f=F, r=0.0, bc=(2, 2, 2, 2),
                            eta=eta.astype(np.float32), hu=(0.01 * rs.randn(ny + 4, nx + 4)).astype(np.float32),
                            hv=(0.01 * rs.randn(ny + 4, nx + 4)).astype(np.float32),
                            Hi=np.full((ny + 5, nx + 5), H, dtype=np.float32)))
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, members, rk_order=2, courant_number=0.7)
    ens.enable_model_error(soar_q0=Q0, interpolation_factor=factor, seed=seed)
    return ens


def test_local_svd_kernel_equals_reference_data():
    import torch
    from SWESimulators import IEWPFOcean
    print("case local svd golden", flush=True)
    ens = _ensemble(3, 40, 1, seed=1)
    ens.enable_drifters(3)
    iw = IEWPFOcean.IEWPFOcean(ens, np.eye(2) * 0.01 ** 2)
    svd = np.loadtxt(os.path.join(GOLD, "iewpf_localSVD.txt")).astype(np.float32)
    pre = np.loadtxt(os.path.join(GOLD, "iewpf_preLocalSVDtoGlobal.txt")).astype(np.float32)
    post = np.loadtxt(os.path.join(GOLD, "iewpf_postLocalSVDtoGlobal.txt"))
    st = ens.gpu_stream
    with torch.cuda.stream(st):
        for k in range(3):
            ens.noise_views(k)[0].copy_(torch.from_numpy(pre))
        d_svd = torch.from_numpy(svd).to(ens.state[0].device)
        cells = torch.tensor([[30, 30]], dtype=torch.int32, device=ens.state[0].device)
    from SWESimulators import _gocean as Gc
    Gc.check(iw.lib.gocean_iewpf_local_svd(ens.noise_random.data_ptr(), iw.offsets.data_ptr(), 3, 40, 40,
                                           d_svd.data_ptr(), cells.data_ptr(), 1, Gc.stream_handle(st)),
             "gocean_iewpf_local_svd")
    torch.cuda.synchronize()
    for k in range(3):
        got = ens.noise_views(k)[0].to("cpu").numpy().astype(np.float64)
        assert np.linalg.norm(got - post) / np.max(post) < 1e-6


CASES = [
    # name, N, n, factor, drifter positions (m)
    ("factor1", 6, 40, 1, [[120.0, 120.0], [40.0, 120.0], [80.0, 40.0]]),
    ("factor1_wrap_overlap", 5, 40, 1, [[2.0, 150.0], [10.0, 158.0], [90.0, 3.0]]),
    ("factor3", 5, 45, 3, [[30.0, 160.0], [100.0, 50.0], [170.0, 170.0], [5.0, 5.0]]),
    ("factor5_bench_domain", 4, (500, 300), 5, "random64"),   # the bench's domain and drifters, 4 particles
]


@pytest.mark.parametrize("name,N,n,factor,pos", CASES, ids=[c[0] for c in CASES])
def test_iewpf_2stage_matches_oracle(name, N, n, factor, pos):
    import torch
    from SWESimulators import IEWPFOcean
    print("case", name, flush=True)
    nx, ny = (n, n) if np.isscalar(n) else n
    if isinstance(pos, str):
        prs = np.random.RandomState(64)
        pos = np.column_stack([prs.uniform(0, nx * DX, 64), prs.uniform(0, ny * DX, 64)])
    pos = np.array(pos)
    Nd = len(pos)
    ens = _ensemble(N, n, factor, seed=N + factor)
    ens.enable_drifters(Nd)
    R = np.eye(2) * 0.01 ** 2
    iw = IEWPFOcean.IEWPFOcean(ens, R)
    torch.cuda.synchronize()
    before = np.stack([np.stack(ens.download(k)) for k in range(N)]).astype(np.float64)
    gamma, nu_norm = iw.samplePerpendicular()
    torch.cuda.synchronize()
    xi = [ens.noise_views(k)[0].to("cpu").numpy().copy() for k in range(N)]
    cnx, cny = nx // factor, ny // factor
    offs = ens.noise_offsets[:, 1]
    nu = [iw.nu[int(o):int(o) + cnx * cny].view(cny, cnx).to("cpu").numpy().copy() for o in offs]
    for k in range(N):                                  # the sample is perpendicular, norms kept
        assert abs(np.sum(xi[k].astype(np.float64) * nu[k])) < 1e-4 * np.sum(xi[k].astype(np.float64) ** 2)
    rs = np.random.RandomState(3)
    inn = 0.02 * rs.randn(N, Nd, 2)
    obs = np.column_stack([pos, np.zeros((Nd, 2))])
    diag = iw.iewpf_2stage(obs, innovations=inn, sample=False)
    torch.cuda.synchronize()
    after = np.stack([np.stack(ens.download(k)) for k in range(N)]).astype(np.float64)
    p = I.Params(nx=nx, ny=ny, dx=DX, dy=DX, factor=factor, q0=Q0, L=0.75 * np.float32(DX * factor),
                 f=F, g=G, H=H, R=R)
    Hi = np.full((ny + 5, nx + 5), H)
    want, wd = I.iewpf_2stage(p, iw.S_host, iw.localSVD_host, before, Hi, pos, inn, xi, nu)
    for key in ("phi", "gamma", "nu_norm", "alpha"):
        np.testing.assert_allclose(diag[key], wd[key], rtol=1e-5, err_msg=key)
    assert abs(diag["beta"] - wd["beta"]) < 1e-6 * max(1.0, abs(wd["beta"]))
    inner = (slice(None), slice(None), slice(2, ny + 2), slice(2, nx + 2))
    incr = want[inner] - before[inner]
    err = np.abs(after[inner] - want[inner])
    for f, fname in enumerate(("eta", "hu", "hv")):
        scale = np.abs(incr[:, f]).max()
        ratio = err[:, f].max() / scale
        print(fname, "max error / max increment", ratio, flush=True)
        assert scale > 0 and ratio < 2e-5
    # periodic ghost cells refreshed
    for k in (0, N - 1):
        for a in after[k]:
            np.testing.assert_array_equal(a[:2, 2:-2], a[-4:-2, 2:-2])
            np.testing.assert_array_equal(a[2:-2, :2], a[2:-2, -4:-2])


def test_iewpf_default_innovations_and_resampling_run():
    """The default path (innovations from the particles' observations at the
    true drifters, fresh samples) runs twice and moves the observed
    transports towards the observations on average."""
    import torch
    from SWESimulators import IEWPFOcean
    N, n = 8, 40
    ens = _ensemble(N, n, 1, seed=9)
    pos = np.array([[60.0, 60.0], [100.0, 140.0]])
    ens.enable_drifters(len(pos))
    iw = IEWPFOcean.IEWPFOcean(ens, np.eye(2) * 0.01 ** 2)
    HX = iw.observeParticles(pos)
    y = HX.mean(axis=0) + 0.01
    obs = np.column_stack([pos, y])
    for _ in range(2):
        d = iw.iewpf_2stage(obs)
        assert np.all(np.isfinite(d["alpha"])) and d["beta"] >= 0
    torch.cuda.synchronize()
    HX_a = iw.observeParticles(pos)
    assert np.linalg.norm(HX_a.mean(axis=0) - y) < np.linalg.norm(HX.mean(axis=0) - y)


@pytest.mark.parametrize("factor,n", [(1, 40), (3, 45)])
def test_iewpf_one_stage_matches_oracle(factor, n):
    """One-stage IEWPF (IEWPFOcean.iewpf, IEWPFOcean.py:331-394) against the
    oracle, given the device's own xi."""
    import torch
    from SWESimulators import IEWPFOcean
    print("case one stage", factor, n, flush=True)
    N = 5
    pos = np.array([[30.0, 150.0], [100.0, 50.0], [150.0, 120.0]])
    ens = _ensemble(N, n, factor, seed=21 + factor)
    ens.enable_drifters(len(pos))
    R = np.eye(2) * 0.01 ** 2
    iw = IEWPFOcean.IEWPFOcean(ens, R)
    torch.cuda.synchronize()
    before = np.stack([np.stack(ens.download(k)) for k in range(N)]).astype(np.float64)
    iw.sampleXi()
    torch.cuda.synchronize()
    xi = [ens.noise_views(k)[0].to("cpu").numpy().copy() for k in range(N)]
    inn = 0.02 * np.random.RandomState(6).randn(N, len(pos), 2)
    diag = iw.iewpf(np.column_stack([pos, np.zeros((len(pos), 2))]), innovations=inn, sample=False)
    torch.cuda.synchronize()
    after = np.stack([np.stack(ens.download(k)) for k in range(N)]).astype(np.float64)
    p = I.Params(nx=n, ny=n, dx=DX, dy=DX, factor=factor, q0=Q0, L=0.75 * np.float32(DX * factor),
                 f=F, g=G, H=H, R=R)
    want, wd = I.iewpf_1stage(p, iw.S_host, iw.localSVD_host, before, np.full((n + 5, n + 5), H), pos, inn, xi)
    for key in ("phi", "gamma", "alpha"):
        np.testing.assert_allclose(diag[key], wd[key], rtol=1e-5, err_msg=key)
    inner = (slice(None), slice(None), slice(2, n + 2), slice(2, n + 2))
    incr = want[inner] - before[inner]
    err = np.abs(after[inner] - want[inner])
    for f, fname in enumerate(("eta", "hu", "hv")):
        ratio = err[:, f].max() / np.abs(incr[:, f]).max()
        print(fname, "max error / max increment", ratio, flush=True)
        assert ratio < 2e-5


@pytest.mark.parametrize("N", [37, 133])
def test_kalman_window_gather_equals_dense_sum(N):
    """gocean_iewpf_kalman_window against the dense sum of every drifter's
    window (float64 on the host): windows that wrap both ways, overlap each
    other and cover a whole axis, a partial last segment, member counts not a
    multiple of the 16 members one workgroup takes, and ghost cells
    untouched."""
    import torch
    from SWESimulators import Common, IEWPFOcean as P, _gocean as Gc
    print("case kalman window gather", N, flush=True)
    Common.HIPContext(0)
    lib = Gc.load()
    nx, ny = 150, 37
    plane = (ny + 4) * (nx + 4)
    rs = np.random.RandomState(12)
    Nd = 7
    origins = np.column_stack([rs.randint(0, ny, Nd), rs.randint(0, nx, Nd)]).astype(np.int32)
    origins[0] = (ny - 3, nx - 10)
    sizes = np.column_stack([rs.randint(1, 25, Nd), rs.randint(1, 80, Nd)])
    sizes[0] = (9, 30)
    sizes[2] = (ny, 12)
    wh, ww = int(sizes[:, 0].max()), int(sizes[:, 1].max())
    basis = np.zeros((Nd, 5, 3, wh, ww), np.float32)
    for d in range(Nd):
        basis[d, :, :, :sizes[d, 0], :sizes[d, 1]] = rs.randn(5, 3, sizes[d, 0], sizes[d, 1])
    coef = rs.randn(N, Nd, 5).astype(np.float32)
    X0 = rs.randn(N, 3, ny + 4, nx + 4).astype(np.float32)
    want = X0.astype(np.float64)
    for d in range(Nd):
        rows = 2 + (origins[d, 0] + np.arange(wh)) % ny
        cols = 2 + (origins[d, 1] + np.arange(ww)) % nx
        term = np.einsum("mk,kfij->mfij", coef[:, d].astype(np.float64), basis[d].astype(np.float64))
        want[:, :, rows[:, None], cols[None, :]] += term
    tiles, off, lst = P.kalman_window_tiles(origins, sizes, nx, ny)
    assert tiles[:, 0].max() < ny and tiles[:, 1].max() * P.SEGMENT < nx and lst.max() < Nd
    dcoef = P.kalman_window_coefficients(coef)
    assert dcoef.shape == (Nd, 5, -(-N // 16) * 16)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in
         dict(X=X0, basis=basis, origins=origins, coef=dcoef, tiles=tiles.astype(np.int32),
              off=off.astype(np.int32), lst=lst.astype(np.int32)).items()}
    Gc.check(lib.gocean_iewpf_kalman_window(t["X"].data_ptr(), N, nx, ny, 3 * plane, t["basis"].data_ptr(), Nd,
                                            wh, ww, t["origins"].data_ptr(), t["coef"].data_ptr(),
                                            dcoef.shape[2], t["tiles"].data_ptr(), len(tiles), t["off"].data_ptr(),
                                            t["lst"].data_ptr(), None), "gocean_iewpf_kalman_window")
    torch.cuda.synchronize()
    got = t["X"].to("cpu").numpy().astype(np.float64)
    inner = (slice(None), slice(None), slice(2, ny + 2), slice(2, nx + 2))
    scale = np.abs(want[inner] - X0[inner]).max()
    assert np.abs(got[inner] - want[inner]).max() < 1e-6 * scale
    ghost = np.ones((ny + 4, nx + 4), bool)
    ghost[2:ny + 2, 2:nx + 2] = False
    np.testing.assert_array_equal(got[:, :, ghost], X0[:, :, ghost])
