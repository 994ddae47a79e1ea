"""Pins of the model-error oracle (oracle/ocean_noise_np.py) on the CPU.

Each pin is something a plausible mistake would break, checked by a route
that does not retype the oracle's formula: closed forms (LCG jump, impulse
response, polynomial reproduction, geostrophic slope), published
known-answer vectors (Philox), invariants (Box-Muller radius and angle,
perpendicularity), the reference's own test expectations
(tests/stochastic/OceanStateNoise_test.py, RandomNumbers_test.py), and the
agreement of two independent restatements: the kernel-semantics functions
(written from ocean_noise.cu) against the CPU-twin loops (written from
OceanStateNoise.py:572-937), as the reference tests its kernels."""

import numpy as np
import pytest

from oracle import ocean_noise_np as O

# the reference's test setup (tests/stochastic/OceanStateNoise_parent.py:39-79)
NX, NY, DX, DY, F, G_, CUT = 30, 40, 7.0, 7.0, 0.02, 9.81, 2
BCS = [(True, True), (False, False), (True, False), (False, True)]   # (NS, EW)


def _seeds(rs, shape):
    return (rs.rand(*shape) * O.LCG_DENOM).astype(np.uint64)


# ---------------------------------------------------------------------------
# sizes: OceanStateNoise_test.py:40-111
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("pns,pew", BCS)
def test_sizes_as_reference_tests(pns, pew):
    cnx, cny, rnx, rny, snx, sny, gx, gy = O.noise_sizes(NX, NY, 1, pns, pew)
    assert rnx == (NX if pew else NX + 2 * (2 + CUT))
    assert rny == (NY if pns else NY + 2 * (2 + CUT))
    assert snx == rnx / 2 and sny == rny
    # global sizes of the reference's 16x16 launches (:46-47): SOAR grid
    # covers coarse+4, random grid covers the seeds
    assert int(np.ceil((cnx + 4) / 16)) == 3 and int(np.ceil((cny + 4) / 16)) == (3 if pns else 3)
    assert int(np.ceil(snx / 16)) == (1 if pew else 2)


def test_sizes_reject_even_factor():
    with pytest.raises(AssertionError):
        O.noise_sizes(30, 30, 2)


# ---------------------------------------------------------------------------
# LCG
# ---------------------------------------------------------------------------
def test_lcg_matches_closed_form_jump():
    rs = np.random.RandomState(3)
    seeds = _seeds(rs, (50,))
    for k in (1, 2, 7, 100):
        s = seeds.copy()
        for _ in range(k):
            s = O.lcg_next(s)
        want = [O.lcg_jump(int(x), k) for x in seeds]
        assert [int(x) for x in s] == want


def test_lcg_known_values():
    # (1 * 1103515245 + 12345) mod (2^31 - 1); 2^31 - 2 maps to 12345 - a mod m
    assert int(O.lcg_next(np.uint64(1))) == 1103527590
    assert int(O.lcg_next(np.uint64(O.LCG_M - 1))) == (12345 - 1103515245) % O.LCG_M
    assert int(O.lcg_next(np.uint64(0))) == 12345


def test_box_muller_invariants():
    rs = np.random.RandomState(4)
    seed = _seeds(rs, (20, 16))
    s2, out = O.lcg_generate(seed, 32, normal=True, uniform_float32=False)
    s1 = O.lcg_next(seed)
    assert np.array_equal(s2, O.lcg_next(s1))
    u1 = s1.astype(np.float64) / O.LCG_DENOM
    u2 = s2.astype(np.float64) / O.LCG_DENOM
    n1, n2 = out[:, 0::2], out[:, 1::2]
    np.testing.assert_allclose(n1 ** 2 + n2 ** 2, -2 * np.log(u1), rtol=1e-12)
    ang = np.mod(np.arctan2(n2, n1), 2 * np.pi)
    np.testing.assert_allclose(ang, np.mod(2 * np.pi * u2, 2 * np.pi), atol=1e-9)


def test_lcg_odd_column_rule():
    """Odd rand_nx: the last column is never written (ocean_noise.cu:186-192)."""
    seed = _seeds(np.random.RandomState(5), (6, 4))
    _, out = O.lcg_generate(seed, 7, normal=False)
    assert np.all(out[:, 6] == 0.0)
    assert np.all(out[:, :6] > 0.0)


@pytest.mark.slow
def test_lcg_statistics_as_reference():
    """RandomNumbers_test.py:56-99 on a 400x400 periodic field."""
    cnx, cny, rnx, rny, snx, sny, _, _ = O.noise_sizes(400, 400)
    rs = np.random.RandomState(6)
    _, U = O.lcg_generate(_seeds(rs, (sny, snx)), rnx, normal=False)
    assert abs(U.mean() - 0.5) < 0.005 and abs(U.var() - 1 / 12) < 0.001
    _, N = O.cpu_twin_lcg_normal(_seeds(rs, (sny, snx)), rnx)
    assert abs(N.mean()) < 0.01 and abs(N.var() - 1.0) < 0.01


# ---------------------------------------------------------------------------
# Philox4x32-10: known-answer vectors of the Random123 distribution
# (kat_vectors, philox4x32 R=10)
# ---------------------------------------------------------------------------
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    out = O.philox4x32_10(np.array(ctr, dtype=np.uint64).reshape(4, 1), key)
    assert tuple(int(v) for v in out[:, 0]) == want


def test_philox_field_statistics_and_counter():
    a = O.philox_generate(401, 400, key=12345, counter=0)
    b = O.philox_generate(401, 400, key=12345, counter=1)
    assert abs(a.mean()) < 0.01 and abs(a.var() - 1.0) < 0.01
    assert abs(np.corrcoef(a.ravel(), b.ravel())[0, 1]) < 0.01
    u = O.philox_generate(401, 400, key=12345, counter=0, normal=False)
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.001


# ---------------------------------------------------------------------------
# SOAR
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("pns,pew", BCS)
def test_soar_impulse_response(pns, pew):
    """A unit random value at one cell spreads as q0 (1 + d/L) exp(-d/L) to
    every coarse cell within the cutoff; coarse cell a is centred on random
    cell a+2 (ghosted random field) or (a-2) mod n (periodic)."""
    cnx, cny, rnx, rny, *_ = O.noise_sizes(NX, NY, 1, pns, pew)
    q0, L, cd = 1e-3, 0.75 * DX, DX
    for ry, rx in ((9, 11), (0, 1), (rny - 1, rnx - 2)):
        xi = np.zeros((rny, rnx))
        xi[ry, rx] = 1.0
        c = O.soar_coarse(xi, cnx, cny, cd, cd, q0, L, 1.0, pns, pew)

        def off(r, a, n, periodic):
            o = (r - (a - 2) + 2) % n - 2 if periodic else r - (a + 2)
            return o if -2 <= o <= 2 else None

        want = np.zeros_like(c)
        for ay in range(cny + 4):
            oy = off(ry, ay, cny, pns)
            for ax in range(cnx + 4):
                ox = off(rx, ax, cnx, pew)
                if oy is not None and ox is not None:
                    d = np.hypot(ox * cd, oy * cd)
                    want[ay, ax] = q0 * (1 + d / L) * np.exp(-d / L)
        np.testing.assert_allclose(c, want, rtol=1e-12, atol=1e-18)
        assert np.count_nonzero(want) >= 1


def test_soar_periodic_translation():
    """Periodic in both directions: shifting the random field by (1, 2)
    shifts the coarse interior by the same amount."""
    rs = np.random.RandomState(7)
    cnx, cny, rnx, rny, *_ = O.noise_sizes(NX, NY)
    xi = rs.randn(rny, rnx)
    a = O.soar_coarse(xi, cnx, cny, DX, DY, 1e-3, 5.0)
    b = O.soar_coarse(np.roll(xi, (1, 2), axis=(0, 1)), cnx, cny, DX, DY, 1e-3, 5.0)
    np.testing.assert_allclose(np.roll(a[2:-2, 2:-2], (1, 2), axis=(0, 1)), b[2:-2, 2:-2], rtol=1e-12)
    # the two ghost layers repeat the periodic image
    np.testing.assert_allclose(a[:2, 2:-2], a[cny:cny + 2, 2:-2], rtol=1e-12)


@pytest.mark.parametrize("pns,pew", BCS)
def test_soar_vectorised_equals_cpu_twin_loops(pns, pew):
    rs = np.random.RandomState(8)
    cnx, cny, rnx, rny, *_ = O.noise_sizes(18, 14, 1, pns, pew)
    xi = rs.randn(rny, rnx)
    a = O.soar_coarse(xi, cnx, cny, DX, DY, DX / 1e5, 0.75 * DX, 1.0, pns, pew)
    b = O.cpu_twin_applyQ(xi, cnx, cny, DX, DY, DX / 1e5, 0.75 * DX, pns, pew)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-20)


# ---------------------------------------------------------------------------
# interpolation
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("factor", [3, 5, 7])
def test_bicubic_reproduces_quadratics(factor):
    """Central-difference Hermite bicubic is exact for polynomials of degree
    <= 2 in each variable: a quadratic on the coarse grid comes out as the
    same quadratic at the fine cell centres."""
    nx, ny = 6 * factor, 5 * factor
    cnx, cny = nx // factor, ny // factor
    cd = float(np.float32(DX * factor))
    jc, ic = np.meshgrid(np.arange(cny + 4), np.arange(cnx + 4), indexing="ij")
    X = (ic - 2 + 0.5) * cd
    Y = (jc - 2 + 0.5) * cd

    def q(x, y):
        return 0.3 + 1e-3 * x - 2e-3 * y + 1e-5 * x * x + 3e-6 * x * y - 2e-5 * y * y

    d = O.bicubic_eta(q(X, Y), nx, ny, DX, DY, cnx, cny, cd, cd)
    jf, if_ = np.meshgrid(np.arange(1, ny + 3), np.arange(1, nx + 3), indexing="ij")
    want = q((if_ - 2 + 0.5) * DX, (jf - 2 + 0.5) * DY)
    np.testing.assert_allclose(d, want, rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("factor", [3, 5, 7])
def test_bicubic_hits_coarse_values_at_aligned_cells(factor):
    """OceanStateNoise_test.py:256-268: fine cells whose centre coincides with
    a coarse centre take the coarse value."""
    rs = np.random.RandomState(9)
    nx, ny = 4 * factor, 3 * factor
    cnx, cny = nx // factor, ny // factor
    C = rs.randn(cny + 4, cnx + 4)
    cd = float(np.float32(DX * factor))
    d = O.bicubic_eta(C, nx, ny, DX, DY, cnx, cny, cd, cd)
    fc = (factor - 1) // 2
    inner = d[1:-1, 1:-1]
    np.testing.assert_allclose(inner[fc::factor, fc::factor], C[2:-2, 2:-2], rtol=1e-6, atol=1e-6)


def test_alignment_offsets_shift_the_field():
    """OceanStateNoise_test.py:283-355: aligning the coarse grid with fine
    cell i shifts the interpolated field by the offset (factor 5, cells
    6, 7, 8 -> offsets 1, 0, -1)."""
    factor = 5
    assert [O.coarse_grid_offset(factor, a, a) for a in (6, 7, 8)] == [(1, 1), (0, 0), (-1, -1)]
    rs = np.random.RandomState(10)
    nx, ny = 6 * factor, 6 * factor
    cnx, cny = nx // factor, ny // factor
    C = rs.randn(cny + 4, cnx + 4)
    cd = float(np.float32(DX * factor))
    base = O.bicubic_eta(C, nx, ny, DX, DY, cnx, cny, cd, cd)
    for off in (1, -1, 2):
        sh = O.bicubic_eta(C, nx, ny, DX, DY, cnx, cny, cd, cd, off, off)
        # value at fine cell t with offset equals the unshifted value at t+off
        a = sh[4:-4, 4:-4]
        b = base[4 + off:base.shape[0] - 4 + off, 4 + off:base.shape[1] - 4 + off]
        np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6)


# ---------------------------------------------------------------------------
# balance
# ---------------------------------------------------------------------------
def test_geostrophic_balance_closed_form():
    """d_eta = s*y: d_hu = -(g/f) h s, d_hv = 0 where h = H + eta + d_eta."""
    nx, ny = 12, 10
    j = np.arange(ny + 4)[:, None] * np.ones((1, nx + 4))
    s = 1e-4
    d = s * (j * DY)
    eta = 0.01 * np.ones((ny + 4, nx + 4))
    H = 5.0 * np.ones((ny + 5, nx + 5))
    z = np.zeros_like(eta)
    e, u, v = O.geostrophic_balance(d, eta, z, z, H, nx, ny, DX, DY, G_, F)
    I = (slice(2, -2), slice(2, -2))
    np.testing.assert_allclose(e[I], eta[I] + d[I], rtol=1e-14)
    np.testing.assert_allclose(u[I], -(G_ / F) * (5.0 + eta[I] + d[I]) * s, rtol=1e-12)
    np.testing.assert_allclose(v[I], 0.0, atol=1e-14)
    assert np.all(e[:2] == eta[:2]) and np.all(u[:, :2] == 0)


def test_beta_plane_and_land():
    nx, ny = 12, 10
    rs = np.random.RandomState(11)
    d = rs.randn(ny + 4, nx + 4) * 1e-3
    eta = rs.randn(ny + 4, nx + 4) * 1e-2
    H = 5.0 * np.ones((ny + 5, nx + 5))
    H[6, 7] = 1e20
    z = np.zeros_like(eta)
    beta = 2e-6
    e, u, v = O.geostrophic_balance(d, eta, z + 1, z + 2, H, nx, ny, DX, DY, G_, F, beta=beta)
    # cells with corner (6, 7): ghosted (5..6, 6..7) -> all three zeroed
    for (jj, ii) in ((5, 6), (5, 7), (6, 6), (6, 7)):
        assert e[jj, ii] == 0 and u[jj, ii] == 0 and v[jj, ii] == 0
    # a wet cell: coriolis f + beta (j + 0.5) dy at ghosted j (ocean_noise.cu:473)
    jj, ii = 9, 3
    cor = F + beta * (jj + 0.5) * DY
    ey = (d[jj + 1, ii] - d[jj - 1, ii]) / (2 * DY)
    np.testing.assert_allclose(u[jj, ii], 1 - (G_ / cor) * (d[jj, ii] + 5.0 + eta[jj, ii]) * ey, rtol=1e-12)
    # bicubic path: a land corner leaves the state unchanged (ocean_noise.cu:758-763)
    nx3 = ny3 = 12
    C = rs.randn(4 + 4, 4 + 4) * 1e-3
    H3 = 5.0 * np.ones((ny3 + 5, nx3 + 5))
    H3[6, 7] = 1e20
    z3 = np.zeros((ny3 + 4, nx3 + 4))
    e3, u3, v3 = O.bicubic_balance(C, z3 + 0.5, z3 + 1, z3 + 2, H3, nx3, ny3, DX, DY, G_, F, factor=3)
    for (jj, ii) in ((5, 6), (5, 7), (6, 6), (6, 7)):
        assert (e3[jj, ii], u3[jj, ii], v3[jj, ii]) == (0.5, 1.0, 2.0)
    assert e3[9, 3] != 0.5


@pytest.mark.parametrize("factor", [1, 3, 5, 7])
@pytest.mark.parametrize("pns,pew", BCS)
def test_kernel_semantics_equal_cpu_twins(factor, pns, pew):
    """The reference tests its kernels against its CPU twins with eta = 0,
    beta = 0, H = 5 (OceanStateNoise_test.py:154-281): the two restatements
    agree the same way, to 1e-7 of the largest value."""
    nx = NX * factor if factor > 1 else NX
    ny = NY * factor if factor > 1 else NY
    if factor > 1:
        nx, ny = 8 * factor, 6 * factor
    cnx, cny, rnx, rny, snx, sny, *_ = O.noise_sizes(nx, ny, factor, pns, pew)
    rs = np.random.RandomState(12 + factor)
    _, xi = O.lcg_generate(_seeds(rs, (sny, snx)), rnx)
    xi = xi.astype(np.float32)
    cd = float(np.float32(DX * factor))
    q0, L = DX / 1e5, 0.75 * cd
    H = 5.0 * np.ones((ny + 5, nx + 5))
    z = np.zeros((ny + 4, nx + 4))
    coarse = O.soar_coarse(xi, cnx, cny, cd, cd, q0, L, 1.0, pns, pew)
    if factor == 1:
        e, u, v = O.geostrophic_balance(coarse, z, z, z, H, nx, ny, DX, DY, G_, F)
    else:
        e, u, v = O.bicubic_balance(coarse, z, z, z, H, nx, ny, DX, DY, G_, F, factor=factor)
    de, dhu, dhv = O.cpu_twin_perturbation(xi, H, nx, ny, DX, DY, factor, q0, L, F, 0.0, G_, pns, pew)
    I = (slice(2, -2), slice(2, -2))
    for a, b in ((e[I], de[I]), (u[I], dhu), (v[I], dhv)):
        assert np.max(np.abs(a - b)) <= 1e-7 * np.max(np.abs(b))


def test_cpu_twin_catches_a_planted_sign_error():
    """The cross-check is sharp: flipping the sign of d_hv in one restatement
    is caught."""
    nx, ny = 12, 10
    cnx, cny, rnx, rny, snx, sny, *_ = O.noise_sizes(nx, ny)
    rs = np.random.RandomState(13)
    xi = rs.randn(rny, rnx)
    H = 5.0 * np.ones((ny + 5, nx + 5))
    z = np.zeros((ny + 4, nx + 4))
    coarse = O.soar_coarse(xi, cnx, cny, DX, DY, DX / 1e5, 0.75 * DX)
    _, _, v = O.geostrophic_balance(coarse, z, z, z, H, nx, ny, DX, DY, G_, F)
    _, _, dhv = O.cpu_twin_perturbation(xi, H, nx, ny, DX, DY, 1, DX / 1e5, 0.75 * DX, F, 0.0, G_, True, True)
    assert np.max(np.abs(-v[2:-2, 2:-2] - dhv)) > 0.5 * np.max(np.abs(dhv))


# ---------------------------------------------------------------------------
# perpendicular pair: RandomNumbers_test.py:116-157
# ---------------------------------------------------------------------------
def test_make_perpendicular():
    rs = np.random.RandomState(14)
    xi = rs.randn(40, 30)
    nu = rs.randn(40, 30)
    red = O.norms(xi, nu)
    p = O.make_perpendicular(xi, nu, red)
    assert abs(np.sum(xi * p)) < 1e-9 * np.sum(xi * xi)
    np.testing.assert_allclose(np.sum(p * p), np.sum(nu * nu), rtol=1e-12)
    np.testing.assert_allclose(O.norms(xi)[0], np.linalg.norm(xi) ** 2, rtol=1e-12)
