"""Wind stress on the GPU against the oracle (oracle/cdklm16_np.py).

* gridded fields -- the reference's bilinear wind textures
  (gpu_ocean/SWESimulators/gpu_kernels/common.cu:616-650, read at
  CDKLM16_kernel.cu:827-828) -- in the 2-D tile kernel and the column-marching
  kernel, bit-exact, for fields coarser than, as fine as and thinner than the
  grid, with a time weight between two samples;
* wind time series through the class's step loop (Simulator.update_wind_stress,
  Simulator.py:145-223: bracketing samples and weight from the time at the
  start of each step), uniform and gridded, against the oracle's substeps
  driven by the same weights;
* the entry points that take uniform wind only reject a grid."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import cdklm16_np as O
import helpers as H

pytestmark = pytest.mark.gpu

BCS = [("wall", (1, 1, 1, 1)), ("periodic", (2, 2, 2, 2)), ("periodicEW", (1, 2, 1, 2))]


def _say(*a):
    print(*a, flush=True)


def _fields(rng, wy, wx, scale=2e-4):
    return tuple((scale * rng.standard_normal((wy, wx))).astype(np.float32) for _ in range(4))


class _Inp:
    pass


def _inp(p, Hi, eta, hu, hv):
    i = _Inp()
    i.nx, i.ny, i.dx, i.dy, i.g, i.r = p.nx, p.ny, p.dx, p.dy, p.g, p.r
    i.bc, i.f, i.wind_x, i.wind_y = p.bc, float(p.f), 0.0, 0.0
    i.Hi, i.eta, i.hu, i.hv = Hi, eta, hu, hv
    return i


@pytest.mark.parametrize("tile", [-1, 8])
@pytest.mark.parametrize("shape", ["coarse", "same", "thin"])
@pytest.mark.parametrize("bcname,bc", BCS)
def test_gridded_wind_substep_bitexact(tile, shape, bcname, bc):
    _say("gridded wind substep", tile, shape, bcname)
    nx, ny = 71, 45
    p, Hi, Hm, eta, hu, hv = H.random_case(nx, ny, bc, 31)
    wy, wx = {"coarse": (7, 9), "same": (ny, nx), "thin": (1, 13)}[shape]
    fields = _fields(np.random.default_rng(5), wy, wx)
    wt = 0.375
    D = H.DeviceDomain(_inp(p, Hi, eta, hu, hv))
    D.bathymetry_init()
    grid = torch.from_numpy(np.stack(fields)).cuda()
    D.wind.t = wt
    D.wind.grid = grid.data_ptr()
    D.wind.grid_nx, D.wind.grid_ny = wx, wy
    p.wind_grid = fields
    src = D.state(eta, hu, hv)
    for rk in (0, 1):
        qn = [(0.9 * eta).astype(np.float32), hu, hv]
        dst = D.state(*qn)
        D.substep(0.8, rk, src, dst, tile=tile)
        ref = O.cdklm_substep(0.8, rk, eta, hu, hv, *qn, Hi, Hm, p, wind_t=wt)
        for g, r in zip(D.host(dst), ref):
            assert np.array_equal(g[2:-2, 2:-2], r[2:-2, 2:-2]), (rk, int(np.sum(g[2:-2, 2:-2] != r[2:-2, 2:-2])))
    # the wind really enters: without it the result differs
    dst = D.state(eta, hu, hv)
    D.wind.grid = None
    D.substep(0.8, 0, src, dst, tile=tile)
    ref = O.cdklm_substep(0.8, 0, eta, hu, hv, eta, hu, hv, Hi, Hm, p, wind_t=wt)
    assert not np.array_equal(D.host(dst)[1][2:-2, 2:-2], ref[1][2:-2, 2:-2])


def _interp(ts, t):
    """Bracketing sample indices and weight at time t (Simulator.py:151-160,
    :218-222)."""
    k = int(np.searchsorted(ts, t))
    t0 = max(0, k - 1)
    t1 = min(len(ts) - 1, k)
    w = max(0.0, min(1.0, (t - ts[t0]) / max(1.0e-10, ts[t1] - ts[t0])))
    return t0, t1, w


@pytest.mark.parametrize("gridded", [False, True])
@pytest.mark.parametrize("tile", [-1, 8])
def test_class_wind_time_series(gridded, tile):
    """CDKLM16.step(30.0) in two calls (12 steps of 2.5 s) over a wind series
    with samples at t = 0, 20 and 60 s: every step takes the bracketing
    samples and weight of its start time; the oracle runs the same substeps
    with those weights.  Bit-exact, and different from the run with the first
    sample held."""
    from SWESimulators import CDKLM16, Common, WindStress
    _say("class wind series", "gridded" if gridded else "uniform", tile)
    nx, ny, dt = 48, 40, 2.5
    p, Hi, Hm, eta, hu, hv = H.random_case(nx, ny, (2, 1, 2, 1), 41)
    rng = np.random.default_rng(8)
    shape = (6, 5) if gridded else (1, 1)
    ts = [0.0, 20.0, 60.0]
    X = [(3e-4 * rng.standard_normal(shape)).astype(np.float32) for _ in ts]
    Y = [(3e-4 * rng.standard_normal(shape)).astype(np.float32) for _ in ts]
    ws = WindStress.WindStress(t=ts, X=X, Y=Y)
    ctx = Common.HIPContext(0)
    sim = CDKLM16.CDKLM16(ctx, eta, hu, hv, Hi, nx, ny, p.dx, p.dy, dt, p.g, float(p.f), 0.0,
                          boundary_conditions=Common.BoundaryConditions(*p.bc), wind_stress=ws,
                          tile_config=tile)
    sim.step(15.0)
    sim.step(15.0)
    assert sim.num_iterations == 12
    got = sim.download()

    # oracle: the class's loop with the weights of each step's start time
    Hio = O.bathymetry_halo(Hi, p)
    Hmo = O.init_bm(Hio)
    h0 = [np.maximum(eta, -Hmo).astype(np.float32), hu.copy(), hv.copy()]
    h1 = [a.copy() for a in h0]
    O.periodic_bc(*h0, p)
    O.periodic_bc(*h1, p)
    t = 0.0
    for _ in range(12):
        k0, k1, w = _interp(ts, t)
        if gridded:
            p.wind_grid = (X[k0], X[k1], Y[k0], Y[k1])
        else:
            p.wind_x = (float(X[k0][0, 0]), float(X[k1][0, 0]))
            p.wind_y = (float(Y[k0][0, 0]), float(Y[k1][0, 0]))
        wt = np.float32(w)
        h1 = [np.array(a) for a in O.cdklm_substep(np.float32(dt), 0, *h0, *h1, Hio, Hmo, p, wind_t=wt)]
        O.periodic_bc(*h1, p)
        h0 = [np.array(a) for a in O.cdklm_substep(np.float32(dt), 1, *h1, *h0, Hio, Hmo, p, wind_t=wt)]
        O.periodic_bc(*h0, p)
        t += dt
    for g, r, name in zip(got, h0, ("eta", "hu", "hv")):
        assert np.array_equal(g[2:-2, 2:-2], r[2:-2, 2:-2]), name
    # holding the first sample gives a different state
    held = CDKLM16.CDKLM16(ctx, eta, hu, hv, Hi, nx, ny, p.dx, p.dy, dt, p.g, float(p.f), 0.0,
                           boundary_conditions=Common.BoundaryConditions(*p.bc),
                           wind_stress=WindStress.WindStress(t=[0.0], X=[X[0]], Y=[Y[0]]), tile_config=tile)
    held.step(30.0)
    assert not np.array_equal(held.download()[1][2:-2, 2:-2], got[1][2:-2, 2:-2])
    sim.cleanUp()
    held.cleanUp()


def test_uniform_wind_only_entry_points_reject_a_grid():
    """Ensembles, the adjoint and wind-driven drifters take uniform wind: a
    grid makes them return an error instead of ignoring it."""
    from SWESimulators import Common, _gocean as G
    lib = G.load()
    p, Hi, Hm, eta, hu, hv = H.random_case(20, 16, (1, 1, 1, 1), 3)
    d = Common.make_domain(20, 16, p.dx, p.dy, p.g, 0.0, 1.3, 2, p.bc, f=float(p.f))
    grid = torch.zeros(4 * 6, dtype=torch.float32, device="cuda")
    m = (G.Member * 1)()
    m[0].domain = d
    m[0].field_offset = -1
    m[0].wind.grid = grid.data_ptr()
    m[0].wind.grid_nx, m[0].wind.grid_ny = 3, 2
    h = ctypes.c_void_p()
    assert lib.gocean_ensemble_create(m, 1, -1, ctypes.byref(h)) != 0
    assert b"uniform wind" in lib.gocean_last_error()
    w = G.Wind()
    w.grid = grid.data_ptr()
    w.grid_nx, w.grid_ny = 3, 2
    q = [torch.from_numpy(a).cuda() for a in (eta, hu, hv)]
    Hd = torch.from_numpy(Hi).cuda()
    scratch = torch.zeros(int(lib.gocean_adjoint_scratch_floats(d)), dtype=torch.float32, device="cuda")
    lam = [torch.zeros_like(q[0]) for _ in range(3)]
    out = [torch.zeros_like(q[0]) for _ in range(3)]
    e = lib.gocean_adjoint_substep(d, ctypes.byref(w), 1.0, None, *(G.ptr(a) for a in q),
                                   *(G.ptr(a) for a in lam), None, None, None, None, None, 1.0, 0.0,
                                   *(G.ptr(a) for a in out), G.ptr(Hd), G.ptr(scratch), None)
    assert e != 0 and b"uniform wind" in lib.gocean_last_error()
    w.grid_nx = 0
    e = lib.gocean_cdklm_substep(d, 1.0, None, 0, *(G.ptr(a) for a in q), None, None, None,
                                 *(G.ptr(a) for a in out), G.ptr(Hd), G.Fields(), ctypes.byref(w), -1, None)
    assert e != 0 and b"wind grid" in lib.gocean_last_error()
