"""Host side of the device IEWPF (SWESimulators/IEWPFOcean.py): S, the local
SVD block and the Kalman-gain patches equal the oracle's restatement of the
reference (IEWPFOcean.py:859-1196), for the reference's unit-test set-up and
with interpolated model error (factors 3 and 5, the double-jet
experiments' factor)."""

import os

import numpy as np
import pytest

from oracle import iewpf_np as I
from SWESimulators import IEWPFOcean as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pair(factor, nx=40, dx=4.0, f=0.05, H=10.0):
    dt, g = 0.05, 9.81
    q0 = 0.5 * dt * f / (g * H)
    L = 0.75 * np.float32(dx * factor)
    R = np.eye(2) * 0.01 ** 2
    o = I.Params(nx=nx, ny=nx, dx=dx, dy=dx, factor=factor, q0=q0, L=L, f=f, g=g, H=H, R=R)
    p = object.__new__(P.IEWPFOcean)          # the host methods only, no ensemble
    p.nx = p.ny = nx
    p.dx = p.dy = np.float32(dx)
    p.interpolation_factor = factor
    p.coarse_nx = p.coarse_ny = nx // factor
    p.coarse_dx = p.coarse_dy = np.float32(dx * factor)
    p.soar_q0, p.soar_L = np.float32(q0), np.float32(L)
    p.f, p.g, p.const_H = np.float32(f), np.float32(g), np.float32(H)
    p.geoBalanceConst = np.float32(p.g * p.const_H / (2.0 * p.f))
    p.R = R
    return p, o


@pytest.mark.parametrize("factor,nx", [(1, 40), (3, 45), (5, 50)])
def test_S_equals_oracle(factor, nx):
    p, o = _pair(factor, nx=nx)
    want = I.create_S(o)
    np.testing.assert_allclose(p._createS(), want, rtol=2e-6, atol=1e-9 * np.abs(want).max())


def test_S_equals_reference_data():
    p, _ = _pair(1)
    want = np.loadtxt(os.path.join(GOLD, "iewpf_S.txt"))
    np.testing.assert_allclose(p._createS(), want, rtol=2e-6, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("factor,nx", [(1, 40), (5, 50)])
def test_local_svd_equals_oracle(factor, nx):
    p, o = _pair(factor, nx=nx)
    p.S_host = p._createS()
    got = p._generateLocalSVDforP().astype(np.float64)
    want = I.local_svd(o, p.S_host).astype(np.float64)
    np.testing.assert_allclose(got @ got.T, want @ want.T, rtol=0, atol=1e-6)


def test_kalman_patches_equal_oracle():
    p, o = _pair(1)
    pat = p._kalman_patch()
    np.testing.assert_allclose(pat[0], I.half_kalman_pattern(o, np.array([1.0, 0.0])), rtol=1e-12, atol=1e-15 * np.abs(pat).max())
    np.testing.assert_allclose(pat[1], I.half_kalman_pattern(o, np.array([0.0, 1.0])), rtol=1e-12, atol=1e-15 * np.abs(pat).max())


def test_catmull_rom_equals_oracle_bicubic():
    rs = np.random.RandomState(4)
    c = rs.randn(7, 7)
    for (i, j, rx, ry) in ((3, 3, 0.0, 0.2), (2, 2, 1.0, 0.8), (3, 2, 0.4, 0.6)):
        assert abs(P.IEWPFOcean._catmull_rom(c, i, j, rx, ry) - I.bicubic_inner(c, i, j, rx, ry)) < 1e-12


def test_circular_cover_is_the_shortest_periodic_interval():
    """Against brute force over every start and length on small masks,
    including wrapped, full and empty ones."""
    rs = np.random.RandomState(4)
    masks = [np.zeros(7, bool), np.ones(7, bool), np.array([1, 0, 0, 0, 0, 0, 1], bool)]
    masks += [rs.rand(n) < p for n in (1, 5, 9, 16) for p in (0.1, 0.3, 0.7) for _ in range(6)]
    for m in masks:
        n = m.size
        start, length = P.circular_cover(m)
        if not m.any():
            assert (start, length) == (0, 0)
            continue
        cover = np.zeros(n, bool)
        cover[(start + np.arange(length)) % n] = True
        assert np.all(cover[m]), (m, start, length)
        best = min(L for L in range(1, n + 1) for s in range(n)
                   if np.all(np.isin(np.flatnonzero(m), (s + np.arange(L)) % n)))
        assert length == best


def test_kalman_window_tiles_list_exactly_the_drifters_meeting_each_segment():
    """Every 256-cell row segment a window meets is listed with that drifter,
    nothing else is, and each list is ascending."""
    nx, ny = 150, 37
    rs = np.random.RandomState(8)
    origins = np.column_stack([rs.randint(0, ny, 9), rs.randint(0, nx, 9)])
    sizes = np.column_stack([rs.randint(1, 20, 9), rs.randint(1, 90, 9)])
    sizes[3] = (0, 0)                                    # a drifter without a term
    sizes[5] = (ny, nx)                                  # one covering everything
    tiles, off, lst = P.kalman_window_tiles(origins, sizes, nx, ny)
    nseg = -(-nx // P.SEGMENT)
    want = {}
    for d in range(9):
        for i in range(sizes[d, 0]):
            for j in range(sizes[d, 1]):
                r, c = (origins[d, 0] + i) % ny, (origins[d, 1] + j) % nx
                want.setdefault((r, c // P.SEGMENT), set()).add(d)
    assert len(tiles) == len(want) == ny * nseg
    assert off[0] == 0 and off[-1] == len(lst) and np.all(np.diff(off) > 0)
    for t, (r, sg) in enumerate(tiles):
        got = lst[off[t]:off[t + 1]]
        assert np.all(np.diff(got) > 0)
        assert set(got.tolist()) == want[(r, sg)]
    none = P.kalman_window_tiles(origins[:1], np.zeros((1, 2), int), nx, ny)
    assert none[0].shape == (0, 2) and none[1].tolist() == [0] and none[2].size == 0


def test_kalman_window_coefficients_layout():
    c = np.random.RandomState(2).randn(37, 4, 5)
    out = P.kalman_window_coefficients(c)
    assert out.shape == (4, 5, 48) and out.dtype == np.float32
    np.testing.assert_array_equal(out[:, :, :37], c.transpose(1, 2, 0).astype(np.float32))
    assert not out[:, :, 37:].any()
