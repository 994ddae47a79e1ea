"""Model error on the GPU (gocean_noise_*) against the oracle
(oracle/ocean_noise_np.py): LCG seeds bit-exact, uniform numbers bit-exact,
normal numbers and perturbations to the stated tolerances; the
OceanStateNoise class on the reference's own test setups; the simulator's
per-step perturbation; the batched form against the single-domain one."""

import ctypes

import numpy as np
import pytest

from oracle import ocean_noise_np as O

pytestmark = pytest.mark.gpu

NX, NY, DX, DY, F, G_ = 30, 40, 7.0, 7.0, 0.02, 9.81
BCS = [(True, True), (False, False), (True, False), (False, True)]
# Box-Muller in float32 (logf, sqrtf, cospif) against float64: |n| <= 6 here
NORMAL_ATOL = 4e-6
# perturbations: max error relative to the largest value of the field, the
# reference tests' 5-6 decimals after scaling by the maximum
REL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _torch():
    import torch
    return torch


def _lib():
    from SWESimulators import _gocean as G
    return G, G.load()


def _dev(a, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _desc(nx, ny, factor=1, pns=True, pew=True, q0=None, L=None, f=F, beta=0.0, north=(0.0, 1.0),
          dx=DX, dy=DY):
    G, _ = _lib()
    d = G.Noise()
    d.nx, d.ny, d.pitch, d.hi_pitch = nx, ny, nx + 4, nx + 5
    d.dx, d.dy = dx, dy
    d.interpolation_factor = factor
    d.periodic_ns, d.periodic_ew = int(pns), int(pew)
    cd = float(np.float32(dx * factor))
    d.soar_q0 = float(np.float32(dx / 1e5 if q0 is None else q0))
    d.soar_L = float(np.float32(0.75 * cd if L is None else L))
    d.g, d.f, d.beta = G_, f, beta
    d.north_x, d.north_y = north
    d.land_value = float(np.float32(1e20))
    return d


def _relerr(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


def _almost(a, b, places):
    """assert2DListAlmostEqual of the reference's tests (testUtils.py):
    round(a - b, places) == 0 everywhere."""
    d = np.round(np.asarray(a, np.float64) - np.asarray(b, np.float64), places)
    assert np.all(d == 0), float(np.max(np.abs(np.asarray(a, np.float64) - b)))


def _scale(a):
    """Scale of a perturbation field for the 5-decimal comparisons: its
    largest magnitude.  The reference's tests divide by the (signed) maximum;
    for a zero-mean random field that maximum can be many times smaller than
    the largest magnitude, which turns the 5-decimal bar into a test of
    float32 rounding for some draws (the generators are seeded from the clock,
    as in the reference)."""
    return float(np.max(np.abs(a)))


def _seeds(rs, shape):
    return (rs.rand(*shape) * O.LCG_DENOM).astype(np.uint64)


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nx,ny,pns,pew", [(30, 40, True, True), (30, 40, False, False),
                                           (45, 30, False, True), (400, 400, True, True)])
def test_lcg_generators_match_oracle(nx, ny, pns, pew):
    G, lib = _lib()
    torch = _torch()
    _say("lcg", nx, ny, pns, pew)
    cnx, cny, rnx, rny, snx, sny, *_ = O.noise_sizes(nx, ny, 1, pns, pew)
    d = _desc(nx, ny, 1, pns, pew)
    rs = np.random.RandomState(1)
    seed0 = _seeds(rs, (sny, snx))
    seed = _dev(seed0.astype(np.int64))
    rnd = torch.full((rny, rnx), -7.0, dtype=torch.float32, device="cuda")
    s = seed0
    for it in range(3):
        kind = G.RNG_LCG_NORMAL if it != 1 else G.RNG_LCG_UNIFORM
        G.check(lib.gocean_noise_generate(d, kind, seed.data_ptr(), rnd.data_ptr(), 0, None, None))
        torch.cuda.synchronize()
        s_new, want = O.lcg_generate(s, rnx, normal=(kind == G.RNG_LCG_NORMAL))
        got_seed = seed.cpu().numpy().astype(np.uint64)
        assert np.array_equal(got_seed, s_new), "seeds must be bit-exact"
        got = rnd.cpu().numpy()
        full = (rnx // 2) * 2
        if kind == G.RNG_LCG_UNIFORM:
            assert np.array_equal(got[:, :full], want[:, :full].astype(np.float32))
        else:
            err = np.max(np.abs(got[:, :full] - want[:, :full]))
            _say("  normal max abs err", err)
            assert err <= NORMAL_ATOL
            # the CPU twin (float64 uniforms), as the reference's compare_random
            _, twin = O.cpu_twin_lcg_normal(s, rnx)
            assert np.max(np.abs(got[:, :full] - twin[:, :full])) <= NORMAL_ATOL
        if rnx % 2:
            assert np.all(got[:, -1] == -7.0), "odd rand_nx: last column untouched"
        s = s_new
    if nx == 400:
        got = rnd.cpu().numpy()
        assert abs(got.mean()) < 0.01 and abs(got.var() - 1) < 0.01


def test_lcg_large_seed_takes_the_wrapping_path():
    """Seeds >= 2^31 (not produced by the generator, but accepted) follow the
    uint64 wrap-around of the reference's expression."""
    G, lib = _lib()
    torch = _torch()
    d = _desc(4, 2)
    seed0 = np.array([[2 ** 40 + 5, 2 ** 63 + 11], [3, 2 ** 31]], dtype=np.uint64)
    seed = _dev(seed0.view(np.int64))
    rnd = torch.zeros((2, 4), dtype=torch.float32, device="cuda")
    G.check(lib.gocean_noise_generate(d, G.RNG_LCG_UNIFORM, seed.data_ptr(), rnd.data_ptr(), 0, None, None))
    torch.cuda.synchronize()
    m = 2 ** 64
    want = []
    for v in seed0.ravel():
        v = int(v)
        for _ in range(2):
            v = ((v * 1103515245 + 12345) % m) % 2147483647
        want.append(v)
    assert [int(x) for x in seed.cpu().numpy().view(np.uint64).ravel()] == want


@pytest.mark.parametrize("nx,ny", [(30, 40), (37, 23)])
def test_philox_generators_match_oracle(nx, ny):
    G, lib = _lib()
    torch = _torch()
    _say("philox", nx, ny)
    cnx, cny, rnx, rny, *_ = O.noise_sizes(nx, ny, 1, False, True)
    d = _desc(nx, ny, 1, False, True)
    rnd = torch.zeros((rny, rnx), dtype=torch.float32, device="cuda")
    counter = torch.tensor([5], dtype=torch.int64, device="cuda")
    key = 0x1234_5678_9ABC_DEF0
    G.check(lib.gocean_noise_generate(d, G.RNG_PHILOX_UNIFORM, None, rnd.data_ptr(), key,
                                      counter.data_ptr(), None))
    torch.cuda.synchronize()
    assert int(counter.item()) == 6
    want = O.philox_generate(rnx, rny, key, 5, normal=False)
    assert np.array_equal(rnd.cpu().numpy(), want.astype(np.float32))
    G.check(lib.gocean_noise_generate(d, G.RNG_PHILOX_NORMAL, None, rnd.data_ptr(), key,
                                      counter.data_ptr(), None))
    torch.cuda.synchronize()
    want = O.philox_generate(rnx, rny, key, 6, normal=True)
    err = np.max(np.abs(rnd.cpu().numpy() - want))
    _say("  normal max abs err", err)
    assert err <= NORMAL_ATOL and int(counter.item()) == 7


# ---------------------------------------------------------------------------
# SOAR and balance
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("factor", [1, 3])
@pytest.mark.parametrize("pns,pew", BCS)
def test_soar_matches_oracle(factor, pns, pew):
    G, lib = _lib()
    torch = _torch()
    nx, ny = (NX, NY) if factor == 1 else (33, 27)
    _say("soar", factor, pns, pew)
    cnx, cny, rnx, rny, *_ = O.noise_sizes(nx, ny, factor, pns, pew)
    d = _desc(nx, ny, factor, pns, pew)
    rs = np.random.RandomState(2)
    xi = rs.randn(rny, rnx).astype(np.float32)
    nu = rs.randn(rny, rnx).astype(np.float32)
    coarse = torch.zeros((cny + 4, cnx + 4), dtype=torch.float32, device="cuda")
    q0 = np.float32(d.soar_q0 * 0.7)
    xi_d, nu_d = _dev(xi), _dev(nu)
    G.check(lib.gocean_noise_soar(d, float(q0), 1.0, xi_d.data_ptr(), coarse.data_ptr(), 0, None))
    G.check(lib.gocean_noise_soar(d, float(q0), 0.5, nu_d.data_ptr(), coarse.data_ptr(), 1, None))
    torch.cuda.synchronize()
    cd = float(np.float32(DX * factor))
    want = O.soar_coarse(xi, cnx, cny, cd, cd, float(q0), d.soar_L, 1.0, pns, pew)
    want = O.soar_coarse(nu, cnx, cny, cd, cd, float(q0), d.soar_L, 0.5, pns, pew, coarse=want, additive=True)
    e = _relerr(coarse.cpu().numpy(), want)
    _say("  rel err", e)
    assert e <= 1e-6


def _state(rs, nx, ny, land=()):
    H = (5.0 + 0.5 * rs.rand(ny + 5, nx + 5)).astype(np.float32)
    for (j, i) in land:
        H[j, i] = np.float32(1e20)
    eta = (0.05 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
    hu = (0.3 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
    hv = (0.3 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
    return H, eta, hu, hv


def _run_balance(d, coarse, H, eta, hu, hv, fields=None, off=(0, 0)):
    G, lib = _lib()
    torch = _torch()
    t = [_dev(a) for a in (coarse, eta, hu, hv, H)]
    G.check(lib.gocean_noise_balance(d, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(),
                                     t[4].data_ptr(), fields, off[0], off[1], None), "balance")
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in t[1:4]]


@pytest.mark.parametrize("case", ["plain", "beta", "land", "fields", "f0field"])
def test_geostrophic_balance_matches_oracle(case):
    G, lib = _lib()
    _say("geostrophic balance", case)
    nx, ny = 45, 38
    rs = np.random.RandomState(3)
    land = [(7, 9), (20, 30), (2, 2)] if case == "land" else []
    H, eta, hu, hv = _state(rs, nx, ny, land)
    coarse = (1e-3 * rs.randn(ny + 4, nx + 4)).astype(np.float32)
    beta = 3e-6 if case in ("beta", "fields") else 0.0
    f = 0.0 if case == "f0field" else F
    d = _desc(nx, ny, 1, f=f, beta=beta)
    fields, north, ffield, keep = None, (0.0, 1.0), None, []
    if case in ("fields", "f0field"):
        ang = 0.3 * rs.rand(ny + 4, nx + 4)
        nxf = np.sin(ang).astype(np.float32)
        nyf = np.cos(ang).astype(np.float32)
        ff = (F + 1e-3 * rs.rand(ny + 4, nx + 4)).astype(np.float32)
        keep = [_dev(a) for a in (ff, nxf, nyf)]
        fields = G.Fields()
        fields.f, fields.north_x, fields.north_y = (t.data_ptr() for t in keep)
        north, ffield = (nxf, nyf), ff
    got = _run_balance(d, coarse, H, eta, hu, hv, fields)
    want = O.geostrophic_balance(coarse, eta, hu, hv, H, nx, ny, DX, DY, G_, f, beta, north, ffield)
    for name, a, b, b0 in zip(("eta", "hu", "hv"), got, want, (eta, hu, hv)):
        e = _relerr(a - b0, b - b0)   # error relative to the perturbation
        _say("  ", name, "rel err", e)
        assert e <= REL
        assert np.array_equal(a[:2], b0[:2]) and np.array_equal(a[:, -2:], b0[:, -2:]), "ghosts untouched"
    if land:
        for (j, i) in land:
            for (jj, ii) in ((j - 1, i - 1), (j - 1, i), (j, i - 1), (j, i)):
                if 2 <= jj < ny + 2 and 2 <= ii < nx + 2:
                    assert got[0][jj, ii] == 0 and got[1][jj, ii] == 0 and got[2][jj, ii] == 0


@pytest.mark.parametrize("factor,off", [(3, (0, 0)), (5, (1, -1)), (5, (-2, 2)), (7, (3, 0))])
@pytest.mark.parametrize("land", [False, True])
def test_bicubic_balance_matches_oracle(factor, off, land):
    G, lib = _lib()
    _say("bicubic", factor, off, land)
    nx, ny = 14 * factor, 9 * factor
    rs = np.random.RandomState(4 + factor)
    lc = [(5, 6), (ny, nx // 2)] if land else []
    H, eta, hu, hv = _state(rs, nx, ny, lc)
    cnx, cny = nx // factor, ny // factor
    coarse = (1e-3 * rs.randn(cny + 4, cnx + 4)).astype(np.float32)
    d = _desc(nx, ny, factor, beta=1e-6)
    got = _run_balance(d, coarse, H, eta, hu, hv, None, off)
    want = O.bicubic_balance(coarse, eta, hu, hv, H, nx, ny, DX, DY, G_, F, 1e-6, factor=factor,
                             offset_i=off[0], offset_j=off[1])
    for name, a, b, b0 in zip(("eta", "hu", "hv"), got, want, (eta, hu, hv)):
        e = _relerr(a - b0, b - b0)
        _say("  ", name, "rel err", e)
        assert e <= REL
    if land:
        assert got[0][4, 5] == eta[4, 5] and got[1][5, 6] == hu[5, 6]


# ---------------------------------------------------------------------------
# the reference's test setups through the OceanStateNoise class
# (tests/stochastic/OceanStateNoise_test.py:115-281)
# ---------------------------------------------------------------------------
def _noise_obj(nx, ny, pns, pew, factor=1, use_lcg=True):
    from SWESimulators import Common, OceanStateNoise
    ctx = Common.HIPContext(0)
    stream = ctx.new_stream()
    bc = Common.BoundaryConditions(2 if pns else 1, 2 if pew else 1, 2 if pns else 1, 2 if pew else 1)
    return OceanStateNoise.OceanStateNoise(ctx, stream, nx, ny, DX, DY, bc, False,
                                           interpolation_factor=factor, use_lcg=use_lcg), stream


@pytest.mark.parametrize("factor", [1, 3, 5, 7])
@pytest.mark.parametrize("pns,pew", BCS)
def test_perturb_ocean_as_reference_tests(factor, pns, pew):
    """perturb_ocean / interpolate_perturbation_ocean: the device result
    against the CPU twins with eta = 0, H = 5, f = 0.02, beta = 0, compared
    after scaling by the largest magnitude at 5 decimals (the reference's bar,
    see _scale); the coarse buffer equals the aligned fine cells."""
    from SWESimulators import Common
    nx, ny = (NX, NY) if factor == 1 else (NX * factor, NY * factor) if factor == 3 else (10 * factor, 8 * factor)
    _say("perturb ocean (reference setup)", factor, pns, pew, nx, ny)
    noise, stream = _noise_obj(nx, ny, pns, pew, factor)
    z = np.zeros((ny + 4, nx + 4), dtype=np.float32)
    HCPU = np.ones((ny + 5, nx + 5), dtype=np.float32) * 5
    eta = Common.DeviceArray2D(stream, nx, ny, 2, 2, z)
    hu = Common.DeviceArray2D(stream, nx, ny, 2, 2, z)
    hv = Common.DeviceArray2D(stream, nx, ny, 2, 2, z)
    H = Common.DeviceArray2D(stream, nx + 1, ny + 1, 2, 2, HCPU)
    noise.perturbOceanState(eta, hu, hv, H, F, 0.0, G_, ghost_cells_x=2, ghost_cells_y=2)
    xi = noise.getRandomNumbers()
    de, dhu, dhv = O.cpu_twin_perturbation(xi, HCPU, nx, ny, DX, DY, factor, float(noise.soar_q0),
                                           float(noise.soar_L), F, 0.0, G_, pns, pew)
    got_e = eta.download(stream)[2:-2, 2:-2]
    got_u = hu.download(stream)[2:-2, 2:-2]
    got_v = hv.download(stream)[2:-2, 2:-2]
    me, mu = _scale(de), _scale(dhu)
    _almost(got_e / me, de[2:-2, 2:-2] / me, 5)
    _almost(got_u / mu, dhu / mu, 5)
    _almost(got_v / mu, dhv / mu, 5)
    if factor > 1:
        coarse = noise.getCoarseBuffer()[2:-2, 2:-2] / me
        fc = (factor - 1) // 2
        _almost(coarse, (got_e / me)[fc::factor, fc::factor], 5)


def test_interpolation_offset_as_reference_test():
    """interpolate_perturbation_ocean_offset (OceanStateNoise_test.py:283-355):
    factor 5, alignment with cells 6, 7, 8 shifts eta by 1, 0, -1."""
    from SWESimulators import Common
    factor, nx, ny = 5, NX, NY
    noise, stream = _noise_obj(nx, ny, True, True, factor)
    z = np.zeros((ny + 4, nx + 4), dtype=np.float32)
    HCPU = np.ones((ny + 5, nx + 5), dtype=np.float32) * 5
    H = Common.DeviceArray2D(stream, nx + 1, ny + 1, 2, 2, HCPU)
    noise.generateNormalDistribution()
    xi = noise.getRandomNumbers()
    de, _, _ = O.cpu_twin_perturbation(xi, HCPU, nx, ny, DX, DY, factor, float(noise.soar_q0),
                                       float(noise.soar_L), F, 0.0, G_, True, True)
    etaCPU = de[2:-2, 2:-2]
    mx = _scale(etaCPU)
    for align, off in zip((6, 7, 8), (1, 0, -1)):
        eta = Common.DeviceArray2D(stream, nx, ny, 2, 2, z)
        hu = Common.DeviceArray2D(stream, nx, ny, 2, 2, z)
        hv = Common.DeviceArray2D(stream, nx, ny, 2, 2, z)
        noise.perturbOceanState(eta, hu, hv, H, F, 0.0, G_, ghost_cells_x=2, ghost_cells_y=2,
                                update_random_field=False, align_with_cell_i=align, align_with_cell_j=align)
        got = eta.download(stream)[2:-2, 2:-2]
        skew = np.roll(etaCPU, (-off, -off), axis=(0, 1))
        _almost(got / mx, skew / mx, 5)


def test_perpendicular_pair_and_norms():
    """RandomNumbers_test.py:116-157, 190-201."""
    noise, stream = _noise_obj(400, 400, True, True, 1, use_lcg=False)
    gamma = noise.getRandomNorm()
    assert gamma == 0.0
    noise.generateNormalDistribution()
    xi = noise.getRandomNumbers().astype(np.float64)
    gamma = noise.getRandomNorm()
    assert abs(gamma - np.sum(xi * xi)) / np.sum(xi * xi) < 1e-5
    noise.generatePerpendicularNormalDistributions()
    xi = noise.getRandomNumbers().astype(np.float64)
    nu = noise.getPerpendicularRandomNumbers().astype(np.float64)
    pre = noise.getReductionBuffer()
    rel = np.sum(xi * xi)
    noise.findDoubleNormAndDot()
    post = noise.getReductionBuffer()
    assert abs(post[0, 0] / rel - np.sum(xi * xi) / rel) < 1e-6
    assert abs(post[0, 1] / rel - np.sum(nu * nu) / rel) < 1e-5
    assert abs(post[0, 0] / rel - pre[0, 0] / rel) < 1e-6
    assert abs(post[0, 1] / rel - pre[0, 1] / rel) < 1e-5
    assert abs(post[0, 2] / rel) < 1e-3 and abs(np.sum(xi * nu) / rel) < 1e-3
    for a in (xi, nu):
        assert abs(a.mean()) < 0.01 and abs(a.var() - 1.0) < 0.01


def test_seed_api():
    noise, _ = _noise_obj(NX, NY, True, True, 1, use_lcg=True)
    s0 = noise.getSeed()
    assert np.array_equal(s0, noise.host_seed)
    noise.generateNormalDistribution()
    s1 = noise.getSeed()
    assert np.array_equal(s1, O.lcg_next(O.lcg_next(s0)))
    noise.resetSeed()
    assert np.array_equal(noise.getSeed(), noise.host_seed)
    ph, _ = _noise_obj(NX, NY, True, True, 1, use_lcg=False)
    assert ph.seed is None and ph.host_seed is None
    with pytest.raises(AssertionError):
        ph.getSeed()


# ---------------------------------------------------------------------------
# the simulator's per-step model error and the batched form
# ---------------------------------------------------------------------------
def _sim_case(nx, ny, bc, f=1e-4, **kw):
    from SWESimulators import CDKLM16, Common
    rs = np.random.RandomState(7)
    j, i = np.mgrid[0:ny + 5, 0:nx + 5]
    Hi = (40 + 5 * np.sin(i / 5.0) + 3 * np.cos(j / 7.0)).astype(np.float32)
    eta = (0.1 * np.exp(-((i[:-1, :-1] - nx / 2) ** 2 + (j[:-1, :-1] - ny / 2) ** 2) / 40.0)).astype(np.float32)
    z = np.zeros((ny + 4, nx + 4), dtype=np.float32)
    ctx = Common.HIPContext(0)
    return CDKLM16.CDKLM16(ctx, eta, z, z, Hi, nx, ny, 500.0, 500.0, 5.0, 9.81, f, 0.0,
                           boundary_conditions=Common.BoundaryConditions(*bc), **kw)


@pytest.mark.parametrize("factor", [1, 3])
def test_simulator_step_applies_model_error(factor):
    """CDKLM16.step with small_scale_perturbation: state after one step =
    deterministic step + balanced perturbation of that state (CDKLM16.py:580-586)."""
    _say("simulator step with model error, factor", factor)
    nx, ny = 48, 36
    bc = (2, 2, 2, 2)
    a = _sim_case(nx, ny, bc, small_scale_perturbation=True, small_scale_perturbation_amplitude=1e-3,
                  small_scale_perturbation_interpolation_factor=factor, use_lcg=True)
    b = _sim_case(nx, ny, bc)
    seed0 = a.small_scale_model_error.getSeed()
    a.step(5.0)
    b.step(5.0)
    eb, ub, vb = (np.asarray(x, np.float64) for x in b.download())
    ea, ua, va = a.download()
    noise = a.small_scale_model_error
    s, xi = O.lcg_generate(seed0, int(noise.rand_nx))
    assert np.array_equal(noise.getSeed(), s)
    xi = noise.getRandomNumbers()
    Hi = a.downloadBathymetry()[0]
    cnx, cny = nx // factor, ny // factor
    cd = float(np.float32(500.0 * factor))
    coarse = O.soar_coarse(xi, cnx, cny, cd, cd, 1e-3, float(noise.soar_L), 1.0, True, True)
    if factor == 1:
        we, wu, wv = O.geostrophic_balance(coarse, eb, ub, vb, Hi, nx, ny, 500.0, 500.0, 9.81, 1e-4)
    else:
        we, wu, wv = O.bicubic_balance(coarse, eb, ub, vb, Hi, nx, ny, 500.0, 500.0, 9.81, 1e-4, factor=factor)
    I = (slice(2, -2), slice(2, -2))
    for name, g, w, base in (("eta", ea, we, eb), ("hu", ua, wu, ub), ("hv", va, wv, vb)):
        e = _relerr(g[I] - base[I], w[I] - base[I])
        _say("  ", name, "rel err of the perturbation", e)
        assert e <= REL
    # periodic ghosts refreshed after the perturbation
    assert np.array_equal(ea[0:2, 2:-2], ea[ny:ny + 2, 2:-2])


def test_batched_equals_single_domain():
    """One batched launch per kernel over members of uneven sizes and BCs
    gives, member by member, the single-domain result (same LCG seeds)."""
    from SWESimulators import CDKLM16Ensemble, Common, OceanStateNoise
    torch = _torch()
    import test_gpu_ensemble as TE
    ms = TE.members(seed=3)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.7)
    ens.enable_model_error(soar_q0=2e-3, use_lcg=True, seed=5)
    singles = []
    for k, m in enumerate(ms):
        pns, pew = m["bc"][0] == 2, m["bc"][1] == 2
        bc = Common.BoundaryConditions(*m["bc"])
        nz = OceanStateNoise.OceanStateNoise(ctx, ens.gpu_stream, m["nx"], m["ny"], m["dx"], m["dy"], bc,
                                             False, soar_q0=2e-3, use_lcg=True)
        so = int(ens.noise_offsets[k, 0])
        n_seed = int(nz.seed_nx) * int(nz.seed_ny)
        nz.setSeed(ens.noise_seed[so:so + n_seed].cpu().numpy().astype(np.uint64).reshape(nz.seed.shape))
        singles.append(nz)
    before = [ens.download(k) for k in range(ens.n)]
    ens.perturb()
    torch.cuda.synchronize()
    for k, (m, nz) in enumerate(zip(ms, singles)):
        e, u, v = (Common.DeviceArray2D(ens.gpu_stream, m["nx"], m["ny"], 2, 2, a) for a in before[k])
        Hi = ens.Hi[ens.hi_off[k]:ens.hi_off[k + 1]].view(m["ny"] + 5, m["nx"] + 5)
        nz.perturbOceanState(e, u, v, Hi, np.float32(m["f"]), 0.0, m["g"])
        torch.cuda.synchronize()
        got = ens.download(k)
        for a, b in zip(got, (e, u, v)):
            bb = b.download(ens.gpu_stream)
            # interior identical; the batch also refreshed periodic ghosts
            assert np.array_equal(a[2:-2, 2:-2], bb[2:-2, 2:-2]), k
        r, c = ens.noise_views(k)
        assert np.array_equal(r.cpu().numpy(), nz.getRandomNumbers())


def test_graph_replays_with_model_error_match_eager():
    """Captured steps with model error: every replay equals the eager run
    from the same state and generator counter (Philox state on the device)."""
    from SWESimulators import CDKLM16Ensemble, Common
    torch = _torch()
    import test_gpu_ensemble as TE
    ms = TE.members(seed=4)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.7)
    ens.enable_model_error(soar_q0=1e-3, use_lcg=False, seed=6)
    ens.capture(3, model_error=True)
    torch.cuda.synchronize()
    for rep in range(3):
        with torch.cuda.stream(ens.gpu_stream):
            s0 = [t.clone() for t in ens.state]
            c0 = ens.noise_counter.clone()
        ens.replay()
        torch.cuda.synchronize()
        with torch.cuda.stream(ens.gpu_stream):
            got = [t.clone() for t in ens.state]
        torch.cuda.synchronize()
        assert int(ens.noise_counter.item()) == int(c0.item()) + 3
        # eager from the same start
        with torch.cuda.stream(ens.gpu_stream):
            for t, s in zip(ens.state, s0):
                t.copy_(s)
            ens.noise_counter.copy_(c0)
        ens.step_n(3, model_error=True)
        torch.cuda.synchronize()
        for a, b in zip(got, ens.state):
            assert torch.equal(a, b), rep
        assert not torch.equal(got[0], s0[0])


@pytest.mark.parametrize("nx,ny", [(45, 38), (130, 9), (131, 40)])
@pytest.mark.parametrize("pns,pew", BCS)
@pytest.mark.parametrize("coriolis", ["fields", "constant"])
def test_fused_soar_balance_equals_two_launches(nx, ny, pns, pew, coriolis):
    """gocean_noise_soar_balance == gocean_noise_soar then
    gocean_noise_balance, bit for bit: the state and every coarse-buffer cell
    (interior and both ghost rings), with land, and with a beta plane and
    per-cell fields or one constant Coriolis parameter (whose g / f the fused
    kernel takes from the host)."""
    G, lib = _lib()
    torch = _torch()
    _say("fused soar+balance", nx, ny, pns, pew, coriolis)
    rs = np.random.RandomState(9)
    H, eta, hu, hv = _state(rs, nx, ny, [(5, 6), (ny, nx // 2)])
    cnx, cny, rnx, rny, *_ = O.noise_sizes(nx, ny, 1, pns, pew)
    xi = _dev(rs.randn(rny, rnx).astype(np.float32))
    fields = G.Fields()
    if coriolis == "fields":
        d = _desc(nx, ny, 1, pns, pew, beta=2e-6)
        ang = 0.2 * rs.rand(ny + 4, nx + 4)
        keep = [_dev(a.astype(np.float32)) for a in (F + 1e-3 * rs.rand(ny + 4, nx + 4), np.sin(ang), np.cos(ang))]
        fields.f, fields.north_x, fields.north_y = (t.data_ptr() for t in keep)
    else:
        d = _desc(nx, ny, 1, pns, pew)
    q0 = float(np.float32(d.soar_q0 * 0.9))
    outs = []
    for fused in (False, True):
        st = [_dev(a) for a in (eta, hu, hv, H)]
        coarse = torch.full((cny + 4, cnx + 4), -3.0, dtype=torch.float32, device="cuda")
        if fused:
            G.check(lib.gocean_noise_soar_balance(d, q0, 0.8, xi.data_ptr(), coarse.data_ptr(), st[0].data_ptr(),
                                                  st[1].data_ptr(), st[2].data_ptr(), st[3].data_ptr(), fields, None))
        else:
            G.check(lib.gocean_noise_soar(d, q0, 0.8, xi.data_ptr(), coarse.data_ptr(), 0, None))
            G.check(lib.gocean_noise_balance(d, coarse.data_ptr(), st[0].data_ptr(), st[1].data_ptr(),
                                             st[2].data_ptr(), st[3].data_ptr(), fields, 0, 0, None))
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in st[:3]] + [coarse.cpu().numpy()])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert not np.any(outs[1][3] == -3.0), "every coarse cell written"
