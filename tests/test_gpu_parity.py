"""Parity of the HIP kernels (called through the C ABI) with the CPU oracle.

The HIP path keeps the reference's operand order with IEEE division and square
root and no FMA contraction, so every comparison here is bit-exact (``==``,
which also equates +0 and -0).  Cases print their name before touching the
GPU.
"""

import numpy as np
import pytest

from oracle import cdklm16_np as O
import helpers as H

pytestmark = pytest.mark.gpu

rng_global = np.random.default_rng(12345)


def _say(*a):
    print(*a, flush=True)


def small_case(nx, ny, bc, seed, land=False, shallow=False, r=0.0, wind=(0.0, 0.0), f=1.2e-4):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny + 5, 0:nx + 5]
    Hi = (40 + 12 * np.sin(x / 5.0) + 9 * np.cos(y / 3.5) + rng.random((ny + 5, nx + 5))).astype(np.float32)
    if shallow:
        w = (nx + 5) // 3
        Hi[:, :w] = (0.05 + 0.03 * rng.random((ny + 5, w))).astype(np.float32)
    if land:
        Hi[ny // 2:ny // 2 + 4, nx // 2:nx // 2 + 3] = O.LAND_VALUE
    p = O.Params(nx=nx, ny=ny, dx=300.0, dy=250.0, bc=bc, f=f, r=r,
                 wind_x=(wind[0], wind[0]), wind_y=(wind[1], wind[1]))
    Hif = O.bathymetry_halo(Hi, p)
    Hm = O.init_bm(Hif)
    eta = (0.4 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    eta = np.maximum(eta, -Hm)
    if shallow:
        w = (nx + 4) // 3
        eta[:, :w] = (-Hm[:, :w] + 0.12 * rng.random((ny + 4, w))).astype(np.float32)
    hu = (3 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    hv = (3 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    if shallow:
        hu[:, :w] *= 0.01
        hv[:, :w] *= 0.01

    class Inp:
        pass
    inp = Inp()
    inp.nx, inp.ny, inp.dx, inp.dy, inp.g, inp.r = nx, ny, 300.0, 250.0, 9.81, r
    inp.bc, inp.f, inp.wind_x, inp.wind_y = bc, f, wind[0], wind[1]
    inp.Hi, inp.eta, inp.hu, inp.hv = Hi, eta, hu, hv
    return inp, p, Hif, Hm


BCS = [("wall", (1, 1, 1, 1)), ("periodic", (2, 2, 2, 2)),
       ("periodicNS", (2, 1, 2, 1)), ("periodicEW", (1, 2, 1, 2))]


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("land", [False, True])
def test_bathymetry_init_bitexact(bcname, bc, land):
    _say("bathymetry_init", bcname, "land" if land else "")
    inp, p, Hif, Hm = small_case(45, 38, bc, 1, land=land)
    D = H.DeviceDomain(inp, pitch_extra=3)
    D.bathymetry_init()
    Hi_d = D.Hi[:, :inp.nx + 5].cpu().numpy()
    Hm_d = D.Hm[:, :inp.nx + 4].cpu().numpy()
    assert np.array_equal(Hi_d, Hif)
    assert np.array_equal(Hm_d, Hm)


FEATURES = [
    dict(),
    dict(land=True),
    dict(shallow=True, r=0.003),
    dict(wind=(2e-4, -1e-4)),
    dict(shallow=True, land=True),
]


# tile configurations: the default 2-D tile and the column-marching strips
TILES = [-1, 8]


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("feat", range(len(FEATURES)))
@pytest.mark.parametrize("rk", [0, 1])
@pytest.mark.parametrize("tile", TILES)
def test_substep_bitexact(bcname, bc, feat, rk, tile):
    kw = FEATURES[feat]
    _say("substep", bcname, kw, "rk", rk, "tile", tile)
    inp, p, Hif, Hm = small_case(53, 41, bc, 10 + feat, **kw)
    D = H.DeviceDomain(inp, pitch_extra=5 if feat % 2 else 0)
    D.bathymetry_init()
    src = D.state(inp.eta, inp.hu, inp.hv)
    qn_host = [(0.9 * inp.eta).astype(np.float32), (1.1 * inp.hu).astype(np.float32),
               (0.7 * inp.hv).astype(np.float32)]
    dst = D.state(*qn_host)
    O.periodic_bc(inp.eta, inp.hu, inp.hv, p)
    src = D.state(inp.eta, inp.hu, inp.hv)
    D.substep(0.7, rk, src, dst, tile=tile)
    got = D.host(dst)
    ref = O.cdklm_substep(0.7, rk, inp.eta, inp.hu, inp.hv, *qn_host, Hif, Hm, p)
    for g, r_, name in zip(got, ref, ("eta", "hu", "hv")):
        gi, ri = g[2:-2, 2:-2], r_[2:-2, 2:-2]
        nbad = int(np.sum(gi != ri))
        assert nbad == 0, f"{name}: {nbad} cells differ, max |d| = {np.max(np.abs(gi - ri))}"


@pytest.mark.parametrize("tile", range(10))
@pytest.mark.parametrize("bcname,bc", [BCS[0], BCS[1], BCS[3]])
def test_tile_configs_bitexact(tile, bcname, bc):
    _say("tile config", tile, bcname)
    inp, p, Hif, Hm = small_case(97, 75, bc, 3, land=True)
    D = H.DeviceDomain(inp)
    D.bathymetry_init()
    O.periodic_bc(inp.eta, inp.hu, inp.hv, p)
    src = D.state(inp.eta, inp.hu, inp.hv)
    dst = D.state(inp.eta, inp.hu, inp.hv)
    D.substep(0.5, 0, src, dst, tile=tile)
    ref = O.cdklm_substep(0.5, 0, inp.eta, inp.hu, inp.hv, inp.eta, inp.hu, inp.hv, Hif, Hm, p)
    for g, r_ in zip(D.host(dst), ref):
        assert np.array_equal(g[2:-2, 2:-2], r_[2:-2, 2:-2])


@pytest.mark.parametrize("bcname,bc", BCS)
def test_field_mode_bitexact(bcname, bc):
    """Per-cell Coriolis parameter and north vector (the texture path)."""
    _say("field mode", bcname)
    inp, p, Hif, Hm = small_case(47, 39, bc, 7)
    rng = np.random.default_rng(99)
    shape = (inp.ny + 4, inp.nx + 4)
    fF = (1e-4 + 3e-5 * rng.random(shape)).astype(np.float32)
    ang = (0.4 * rng.random(shape)).astype(np.float32)
    nxF = np.sin(ang.astype(np.float64)).astype(np.float32)
    nyF = np.cos(ang.astype(np.float64)).astype(np.float32)
    p.f, p.north_x, p.north_y = fF, nxF, nyF
    D = H.DeviceDomain(inp, fields=(fF, nxF, nyF))
    D.bathymetry_init()
    O.periodic_bc(inp.eta, inp.hu, inp.hv, p)
    src = D.state(inp.eta, inp.hu, inp.hv)
    for rk in (0, 1):
        dst = D.state(inp.eta * 0.5, inp.hu, inp.hv)
        D.substep(0.6, rk, src, dst)
        ref = O.cdklm_substep(0.6, rk, inp.eta, inp.hu, inp.hv, (inp.eta * 0.5).astype(np.float32),
                              inp.hu, inp.hv, Hif, Hm, p)
        for g, r_ in zip(D.host(dst), ref):
            assert np.array_equal(g[2:-2, 2:-2], r_[2:-2, 2:-2])


@pytest.mark.parametrize("bcname,bc", BCS)
@pytest.mark.parametrize("rk_order", [1, 2, 3])
@pytest.mark.parametrize("tile", TILES)
def test_rk_steps_full_arrays(bcname, bc, rk_order, tile):
    """Whole padded arrays (ghost cells included) after several RK steps
    against the oracle host loop."""
    _say("rk_steps", bcname, "order", rk_order, "tile", tile)
    inp, p, Hif, Hm = small_case(40, 33, bc, 21, wind=(1e-4, 5e-5))
    p.rk_order = rk_order
    sim = O.OracleSim(inp.eta, inp.hu, inp.hv, inp.Hi, p, dt=0.8)
    D = H.DeviceDomain(inp, rk_order=rk_order)
    D.bathymetry_init()
    h0 = D.state(*sim.h0)
    D.periodic_bc(h0)
    h1 = D.state(*sim.h1)
    h2 = D.state(*sim.h1) if rk_order == 3 else None
    n = 6
    for _ in range(n):
        sim.step(np.float32(0.8))
    D.rk_steps(n, np.float32(0.8), h0, h1, h2, tile=tile)
    got = D.host(h0)
    for g, r_, name in zip(got, sim.h0, ("eta", "hu", "hv")):
        if bc == (1, 1, 1, 1) or rk_order != 2:
            g, r_ = g[2:-2, 2:-2], r_[2:-2, 2:-2]
        assert np.array_equal(g, r_), f"{name}: max |d| {np.max(np.abs(g - r_))}"


def test_max_dt_exact():
    _say("max_dt")
    for bcname, bc in BCS:
        inp, p, Hif, Hm = small_case(61, 47, bc, 5, land=True)
        D = H.DeviceDomain(inp)
        D.bathymetry_init()
        h = D.state(inp.eta, inp.hu, inp.hv)
        m = D.max_dt(h, courant=0.7)
        ref = O.max_dt(inp.eta, inp.hu, inp.hv, Hm, p)
        assert m[0] == ref
        assert m[1] == np.float32(np.float32(0.7) * ref)


def test_config1_one_step_and_200():
    """C1 (256^2, flat 60 m, walls, Gaussian bump + noise): one RK2 step at
    max relative error <= 1e-5 (observed 0: bit-exact), and 200 steps."""
    from SWESimulators import BathymetryAndICs as B
    _say("config1 full size")
    c1 = B.config1()
    p = H.oracle_params(c1)
    sim = O.OracleSim(c1.eta, c1.hu, c1.hv, c1.Hi, p, dt=-1)
    D = H.DeviceDomain(c1)
    D.bathymetry_init()
    h0 = D.state(*sim.h0)
    h1 = D.state(*sim.h1)
    m = D.max_dt(h0, courant=0.7)
    dt = np.float32(np.float32(0.7) * np.float32(O.max_dt(*sim.h0, sim.Hm, p)))
    assert m[1] == dt
    sim.step_n(1, dt)
    D.rk_steps(1, dt, h0, h1)
    got = D.host(h0)
    for g, r_ in zip(got, sim.h0):
        assert H.max_rel(g[2:-2, 2:-2], r_[2:-2, 2:-2]) <= 1e-5
        assert np.array_equal(g, r_)
    sim.step_n(199, dt)
    D.rk_steps(199, dt, h0, h1)
    for g, r_ in zip(D.host(h0), sim.h0):
        assert np.array_equal(g, r_)


def test_device_dt_pointer():
    """dt read from device memory (graph-capturable) equals host dt."""
    import torch
    _say("device dt")
    inp, p, Hif, Hm = small_case(33, 30, (1, 2, 1, 2), 8)
    D = H.DeviceDomain(inp)
    D.bathymetry_init()
    a0 = D.state(inp.eta, inp.hu, inp.hv)
    a1 = D.state(inp.eta, inp.hu, inp.hv)
    b0 = D.state(inp.eta, inp.hu, inp.hv)
    b1 = D.state(inp.eta, inp.hu, inp.hv)
    dt_dev = torch.tensor([0.55], dtype=torch.float32, device="cuda")
    D.rk_steps(3, 0.55, a0, a1)
    D.rk_steps(3, 123.0, b0, b1, dt_dev=dt_dev)
    for x, y in zip(D.host(a0), D.host(b0)):
        assert np.array_equal(x, y)


def test_config1_fp64_error_growth():
    """C1 on the GPU against the float64 oracle: max |d| / max |ref| per
    field after 1, 10, 50, 100 and 200 RK2 steps, printed; bounded by 1e-5
    for eta after one step and 1e-3 for every field at 200 steps."""
    from SWESimulators import BathymetryAndICs as B
    _say("config1 fp64 error growth")
    c1 = B.config1()
    p = H.oracle_params(c1)
    sim = O.OracleSim(c1.eta, c1.hu, c1.hv, c1.Hi, p, dt=-1)
    D = H.DeviceDomain(c1)
    D.bathymetry_init()
    h0 = D.state(*sim.h0)
    h1 = D.state(*sim.h1)
    dt = np.float32(np.float32(0.7) * np.float32(O.max_dt(*sim.h0, sim.Hm, p)))
    Hi64 = H.hi64(sim.Hi)
    Hm64 = O.init_bm(Hi64, dtype=np.float64)
    q = [a.astype(np.float64) for a in sim.h0]
    done = 0
    for target in (1, 10, 50, 100, 200):
        D.rk_steps(target - done, dt, h0, h1)
        for _ in range(target - done):
            e, u, v = q
            s = [np.array(a) for a in O.cdklm_substep(float(dt), 0, e, u, v, e.copy(), u.copy(), v.copy(),
                                                      Hi64, Hm64, p, wind_t=0.0, dtype=np.float64)]
            q = [np.array(a) for a in O.cdklm_substep(float(dt), 1, *s, e.copy(), u.copy(), v.copy(),
                                                      Hi64, Hm64, p, wind_t=0.0, dtype=np.float64)]
        done = target
        errs = [H.max_rel(g[2:-2, 2:-2], r_[2:-2, 2:-2]) for g, r_ in zip(D.host(h0), q)]
        _say("   steps", target, "max rel error vs fp64 (eta, hu, hv):", ["%.2e" % e for e in errs])
        if target == 1:
            assert errs[0] <= 1e-5
        if target == 200:
            assert max(errs) <= 1e-3
