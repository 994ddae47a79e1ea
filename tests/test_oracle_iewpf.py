"""Pins of the IEWPF oracle (oracle/iewpf_np.py, the reference's
IEWPFOcean.py) against the reference's own golden data
(gpu_ocean/tests/iewpfRefData, committed as tests/golden/iewpf_*.txt) for the
set-up of its unit test (IEWPFOcean_test.py:64-140), and by what the
mathematics fixes:

  * S = (H Q H^T + R)^-1 equals S.dat;
  * the local SVD block's full matrix (the square root is not unique) equals
    localSVD.dat's;
  * the block applied at coarse cell (30, 30) maps preLocalSVDtoGlobal.dat
    to postLocalSVDtoGlobal.dat (with the golden block);
  * the Kalman-gain stencil equals the matrix form Q U_GB^T H^T e whose
    matrices the golden SVD pins;
  * alpha solves (alpha - 1) gamma - N_x log alpha = c*, the equation whose
    closed form the Lambert W gives."""

import os

import numpy as np
import pytest

from oracle import iewpf_np as I

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ref_params(factor=1):
    """IEWPFOcean_test.py:64-140: 40 x 40, dx = 4, f = 0.05, H = 10,
    q0 = 0.5 dt f / (g H) with dt = 0.05, soar_L = 0.75 coarse dx,
    observation variance 0.01^2."""
    dt, f, g, H = 0.05, 0.05, 9.81, 10.0
    q0 = 0.5 * dt * f / (g * H)
    return I.Params(nx=40, ny=40, dx=4.0, dy=4.0, factor=factor, q0=q0,
                    L=0.75 * np.float32(4.0 * factor), f=f, g=g, H=H, R=np.eye(2) * 0.01 ** 2)


def test_S_equals_reference_data():
    p = ref_params()
    S = I.create_S(p)
    want = np.loadtxt(os.path.join(GOLD, "iewpf_S.txt"))
    np.testing.assert_allclose(S, want, rtol=1e-6, atol=1e-9 * np.abs(want).max())


def test_local_svd_full_matrix_equals_reference_data():
    p = ref_params()
    svd = I.local_svd(p, I.create_S(p)).astype(np.float64)
    want = np.loadtxt(os.path.join(GOLD, "iewpf_localSVD.txt"))
    assert svd.shape == want.shape == (49, 49)
    np.testing.assert_allclose(svd @ svd.T, want @ want.T, rtol=0, atol=1e-5)


def test_local_svd_application_equals_reference_data():
    svd = np.loadtxt(os.path.join(GOLD, "iewpf_localSVD.txt"))
    pre = np.loadtxt(os.path.join(GOLD, "iewpf_preLocalSVDtoGlobal.txt"))
    post = np.loadtxt(os.path.join(GOLD, "iewpf_postLocalSVDtoGlobal.txt"))
    got = I.apply_local_svd(pre.copy(), svd, 30, 30)
    assert np.linalg.norm(got - post) / np.max(post) < 1e-9
    # outside the 7 x 7 window nothing moves
    mask = np.ones_like(pre, dtype=bool)
    mask[27:34, 27:34] = False
    np.testing.assert_array_equal(got[mask], pre[mask])


@pytest.mark.parametrize("e", [(1.0, 0.0), (0.0, 1.0), (0.37, -1.3)])
def test_kalman_stencil_equals_matrix_form(e):
    p = ref_params()
    pat = I.half_kalman_pattern(p, np.array(e))
    H = I.h_matrix(7, 3, 3)
    want = (I.soar_matrix(p) @ I.ugb_matrix(p).T @ H.T @ np.array(e)).reshape(7, 7)
    np.testing.assert_allclose(pat, want, rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("rx,ry", [(0.0, 1 / 3), (1 / 3, 0.0), (1.0, 2 / 3), (0.2, 0.7)])
def test_bicubic_reproduces_quadratics(rx, ry):
    """Cubic Hermite interpolation with central-difference slopes is exact
    for quadratics (a wrong coefficient, sign or corner breaks this)."""
    j, i = np.mgrid[0:7, 0:7].astype(np.float64)
    q = lambda x, y: 0.3 + 1.1 * x - 0.7 * y + 0.25 * x * x - 0.4 * x * y + 0.15 * y * y
    c = q(i, j)
    for (ci, cj) in ((3, 3), (2, 2), (2, 3)):
        assert abs(I.bicubic_inner(c, ci, cj, rx, ry) - q(ci + rx, cj + ry)) < 1e-12
    # and not for a cubic in x (so the test can tell)
    c3 = c + 0.05 * i ** 3
    assert abs(I.bicubic_inner(c3, 3, 3, 0.25, 0.5) - (q(3.25, 3.5) + 0.05 * 3.25 ** 3)) > 1e-4


def test_S_with_interpolation_reads_the_balance_between_coarse_cells():
    """With interpolation factor 3 the geostrophic response is read one fine
    cell (a third of a coarse cell) from the centre on the bicubic surface:
    S stays symmetric positive definite and its response is smaller than the
    coarse-cell differences of factor 1 on the same coarse grid."""
    p3 = ref_params(3)
    p1 = I.Params(nx=40, ny=40, dx=12.0, dy=12.0, factor=1, q0=p3.q0, L=p3.L, f=p3.f, g=p3.g, H=p3.H, R=p3.R)
    HQ3 = np.linalg.inv(I.create_S(p3).astype(np.float64)) - p3.R
    HQ1 = np.linalg.inv(I.create_S(p1).astype(np.float64)) - p1.R
    assert np.all(np.linalg.eigvalsh(0.5 * (HQ3 + HQ3.T)) > 0)
    assert abs(HQ3[0, 1]) < 1e-3 * abs(HQ3[0, 0])
    assert abs(HQ3[0, 0] - HQ3[1, 1]) < 1e-6 * abs(HQ3[0, 0])


@pytest.mark.parametrize("gamma,c_star", [(4800.0, 3.0), (5200.0, -2.0), (4700.0, 40.0)])
def test_alpha_solves_the_implicit_equation(gamma, c_star):
    Nx = 4800
    a = I.solve_alpha(gamma, Nx, c_star)
    assert a > 0
    assert abs((a - 1) * gamma - Nx * np.log(a) - c_star) < 1e-8 * max(1.0, abs(c_star))
    # principal branch: alpha gamma / Nx <= 1
    assert a * gamma / Nx <= 1.0 + 1e-12


def test_two_stage_target_and_beta():
    """w_target is the mean of c, and beta is the largest value for which
    every particle's c* = w_target - c - (beta - 1) nu stays >= 0 (the
    particle with the smallest slack gets c* = 0)."""
    rs = np.random.RandomState(0)
    c = rs.uniform(2, 6, 8)
    nu = rs.uniform(4000, 5000, 8)
    w, b = I.target_two_stage(c, nu)
    assert abs(w - c.mean()) < 1e-12
    c_star = w - c - (b - 1) * nu
    assert np.all(c_star >= -1e-9) and abs(c_star.min()) < 1e-9


def test_one_stage_particle_with_largest_c_keeps_alpha_one():
    """In the one-stage filter c* = max(c) - c: the particle with the largest
    c has c* = 0, whose principal-branch root is alpha = 1 while gamma <= N_x."""
    p = ref_params()
    S = I.create_S(p)
    svd = I.local_svd(p, S)
    rs = np.random.RandomState(5)
    N, n = 4, 40
    states = 0.01 * rs.randn(N, 3, n + 4, n + 4)
    pos = np.array([[60.0, 80.0], [120.0, 20.0]])
    inn = 0.02 * rs.randn(N, 2, 2)
    xi = [0.9 * rs.randn(n, n) for _ in range(N)]          # gamma below N_x
    _, d = I.iewpf_1stage(p, S, svd, states, np.full((n + 5, n + 5), 10.0), pos, inn, xi)
    k = int(np.argmax(d["c"]))
    assert d["gamma"][k] <= p.Nx
    assert abs(d["alpha"][k] - 1.0) < 1e-9
    assert np.all(d["alpha"] > 0)
