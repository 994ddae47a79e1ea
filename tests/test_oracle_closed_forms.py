"""Closed forms that pin the oracle's source terms and time integrators,
independently of the reference's golden runs (which have no wind, no friction
and only SSP-RK2):

* uniform wind on a flat lake at rest (periodic, so the state stays uniform):
  every face flux is the same, so after one substep hu = dt X and hv = dt Y
  exactly, eta stays 0, and after a full RK2 step the Coriolis rotation of the
  intermediate state adds the known second-order terms
  (CDKLM16_kernel.cu:817-864: X, Y and the Coriolis term enter the hu / hv
  sources with these signs);
* the wind time interpolation: weight t between two samples;
* quadratic bottom friction on uniform flow over a flat bottom: the implicit
  factor 1 / (1 + C), C = dt r |u| / h, on each RK2 stage (:884-928);
* SSP-RK1/2/3 composed by hand from Euler substeps (rk_step 0) against the
  oracle's own RK steps (:902-982);
* gridded wind: the texture coordinates of common.cu:616-650 (ghost-shifted
  cell index, texel centres, clamping) on fields whose interpolant is known."""

import numpy as np
import pytest

from oracle import cdklm16_np as O
import helpers as H


def _uniform(nx, ny, H0, hu, hv, **kw):
    p = O.Params(nx=nx, ny=ny, dx=300.0, dy=200.0, bc=(2, 2, 2, 2), **kw)
    Hi = np.full((ny + 5, nx + 5), H0, np.float32)
    Hi = O.bathymetry_halo(Hi, p)
    Hm = O.init_bm(Hi)
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    return p, Hi, Hm, eta, np.full_like(eta, hu), np.full_like(eta, hv)


def _interior(a):
    return np.asarray(a)[2:-2, 2:-2]


def test_uniform_wind_from_rest_one_substep_exact():
    X, Y, f, dt = np.float32(9.8e-5), np.float32(-5.1e-5), np.float32(1.2e-4), np.float32(7.5)
    p, Hi, Hm, eta, hu, hv = _uniform(24, 18, 50.0, 0.0, 0.0, f=f, wind_x=(X, X), wind_y=(Y, Y))
    e, u, v = O.cdklm_substep(dt, 0, eta, hu, hv, eta, hu, hv, Hi, Hm, p)
    assert np.all(_interior(e) == 0.0)
    assert np.all(_interior(u) == np.float32(dt * X))
    assert np.all(_interior(v) == np.float32(dt * Y))


def test_uniform_wind_from_rest_rk2_step():
    X, Y, f, dt = 9.8e-5, -5.1e-5, 1.2e-4, 7.5
    p, Hi, Hm, eta, hu, hv = _uniform(24, 18, 50.0, 0.0, 0.0, f=np.float32(f),
                                      wind_x=(X, X), wind_y=(Y, Y))
    sim = O.OracleSim(eta, hu, hv, Hi, p, dt=dt)
    sim.step_n(1)
    e, u, v = (_interior(a).astype(np.float64) for a in sim.download())
    # hu* = dt X, hv* = dt Y; Q^{n+1} = (Q^n + Q* + dt L(Q*)) / 2 with
    # L(Q*) = (X + f hv*, Y - f hu*)
    want_u = dt * X + 0.5 * f * dt * dt * Y
    want_v = dt * Y - 0.5 * f * dt * dt * X
    assert np.all(e == 0.0)
    assert np.max(np.abs(u - want_u)) <= 1e-6 * abs(want_u)
    assert np.max(np.abs(v - want_v)) <= 1e-6 * abs(want_v)
    # a sign error in the Coriolis term would move hu by f dt^2 |Y|
    assert abs(f * dt * dt * Y) > 100 * 1e-6 * abs(want_u)


@pytest.mark.parametrize("t", [0.0, 0.25, 0.5, 1.0])
def test_wind_time_interpolation(t):
    x0, x1, y0, y1, dt = 2e-4, -1e-4, 5e-5, 3e-4, 4.0
    p, Hi, Hm, eta, hu, hv = _uniform(12, 10, 40.0, 0.0, 0.0, f=np.float32(0.0),
                                      wind_x=(x0, x1), wind_y=(y0, y1))
    e, u, v = O.cdklm_substep(np.float32(dt), 0, eta, hu, hv, eta, hu, hv, Hi, Hm, p, wind_t=t)
    wx = (1 - t) * x0 + t * x1
    wy = (1 - t) * y0 + t * y1
    assert np.allclose(_interior(u), dt * wx, rtol=3e-7, atol=0)
    assert np.allclose(_interior(v), dt * wy, rtol=3e-7, atol=0)


@pytest.mark.parametrize("uv", [(0.5, -0.3), (-1.2, 0.0), (0.05, 0.8)])
def test_quadratic_friction_uniform_flow(uv):
    h, r, dt = 50.0, 0.003, 6.0
    u0, v0 = uv
    p, Hi, Hm, eta, hu, hv = _uniform(16, 14, h, u0 * h, v0 * h, f=np.float32(0.0), r=r)
    sim = O.OracleSim(eta, hu, hv, Hi, p, dt=dt)
    sim.step_n(1)
    e, u, v = (_interior(a).astype(np.float64) for a in sim.download())
    C0 = dt * r * np.hypot(u0, v0) / h
    us, vs = u0 / (1 + C0), v0 / (1 + C0)
    C1 = dt * r * np.hypot(us, vs) / h
    want_u = 0.5 * (u0 + us) * h / (1 + 0.5 * C1)
    want_v = 0.5 * (v0 + vs) * h / (1 + 0.5 * C1)
    assert np.all(e == 0.0)
    for got, want in ((u, want_u), (v, want_v)):
        assert np.max(np.abs(got - want)) <= 1e-6 * max(abs(want), 1e-30) + 1e-12
    # the friction changes the momentum by far more than the tolerance
    big = max((want_u, u0 * h), (want_v, v0 * h), key=lambda t: abs(t[1]))
    assert abs(big[0] - big[1]) > 100 * 1e-6 * abs(big[1])


def _euler(p, Hi, Hm, q, dt):
    """One forward-Euler substep (rk_step 0): Q + dt L(Q), dry clamp included,
    then the periodic ghost cells."""
    out = [np.array(a) for a in O.cdklm_substep(dt, 0, *q, *q, Hi, Hm, p)]
    O.periodic_bc(*out, p)
    return out


@pytest.mark.parametrize("order", [1, 2, 3])
def test_ssp_rk_steps_compose_from_euler_substeps(order):
    """SSP-RK by its definition: RK1 Q1 = E(Q); RK2 Q1 = (Q + E(E(Q)))/2;
    RK3 Q2 = 3/4 Q + 1/4 E(E(Q)), Q3 = 1/3 Q + 2/3 E(Q2), each E an Euler
    substep of the oracle on deep water (no clamp acts).  Periodic sides: at
    a wall face the upwinding of the tangential flux is decided by rounding
    (u- + u+ vanishes there), so an ulp of difference between two ways of
    forming a stage could flip it."""
    p, Hi, Hm, eta, hu, hv = H.random_case(31, 23, (2, 2, 2, 2), 8 + order)
    p.rk_order = order
    dt = np.float32(2.5)
    sim = O.OracleSim(eta, hu, hv, Hi, p, dt=float(dt))
    q0 = [a.astype(np.float64) for a in sim.download()]
    sim.step_n(1, dt)
    got = sim.download()
    q = [np.asarray(a, np.float32) for a in (np.maximum(eta, -Hm), hu, hv)]
    e1 = _euler(p, Hi, Hm, q, dt)
    if order == 1:
        want = [a.astype(np.float64) for a in e1]
    elif order == 2:
        e2 = _euler(p, Hi, Hm, e1, dt)
        want = [0.5 * (a + b.astype(np.float64)) for a, b in zip(q0, e2)]
    else:
        e2 = _euler(p, Hi, Hm, e1, dt)
        q2 = [(0.75 * a + 0.25 * b.astype(np.float64)).astype(np.float32) for a, b in zip(q0, e2)]
        O.periodic_bc(*q2, p)
        e3 = _euler(p, Hi, Hm, q2, dt)
        want = [a / 3.0 + 2.0 / 3.0 * b.astype(np.float64) for a, b in zip(q0, e3)]
    # float32 rounding of the stage combinations (one ulp in Q2 is carried
    # through the last Euler substep) stays within 3e-6 of the field maximum
    for g, w, name in zip(got, want, ("eta", "hu", "hv")):
        err = H.max_rel(_interior(g), _interior(w))
        assert err <= 3e-6, (order, name, err)
    # the wrong weights (RK2's for RK3) would be far off
    if order == 3:
        e2 = _euler(p, Hi, Hm, e1, dt)
        rk2 = [0.5 * (a + b.astype(np.float64)) for a, b in zip(q0, e2)]
        assert H.max_rel(_interior(got[1]), _interior(rk2[1])) > 1e-4


# ---------------------------------------------------------------------------
# gridded wind: the texture lookup of common.cu:616-650
# ---------------------------------------------------------------------------
def test_wind_texture_same_resolution_reads_the_shifted_texel():
    """With a field of nx x ny texels, cell (j, i) reads texel (j+2, i+2):
    s = (i + 2.5)/nx puts it on that texel's centre (the cell index is
    ghost-inclusive), so both weights vanish up to the rounding of s; cells
    past the field's end read the clamped last texel."""
    nx, ny = 13, 9
    rng = np.random.default_rng(1)
    F = rng.standard_normal((ny, nx)).astype(np.float32)
    got = O.wind_texture(F, nx, ny).astype(np.float64)
    j = np.minimum(np.arange(ny) + 2, ny - 1)
    i = np.minimum(np.arange(nx) + 2, nx - 1)
    want = F[np.ix_(j, i)].astype(np.float64)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(F))
    # one texel off (no ghost shift) is far away
    assert np.max(np.abs(got - F[np.ix_(np.arange(ny), np.arange(nx))])) > 0.5


def test_wind_texture_bilinear_reproduces_a_linear_field():
    """A field linear in the texel index, read on a coarser grid: the
    interpolated value is the linear function at texel coordinate
    x(i) = (i + 2.5) w / nx - 0.5 (clamped to [0, w-1]), in both directions."""
    nx, ny, wx, wy = 40, 30, 9, 7
    c0, cx, cy = 0.3, 0.05, -0.02
    kx = np.arange(wx)[None, :]
    ky = np.arange(wy)[:, None]
    F = (c0 + cx * kx + cy * ky).astype(np.float32)
    got = O.wind_texture(F, nx, ny, t=np.float64).astype(np.float64)
    x = np.clip((np.arange(nx) + 2.5) * wx / nx - 0.5, 0, wx - 1)
    y = np.clip((np.arange(ny) + 2.5) * wy / ny - 0.5, 0, wy - 1)
    want = c0 + cx * x[None, :] + cy * y[:, None]
    assert np.max(np.abs(got - want)) < 1e-6
    # the same coordinates without the two ghost cells would be off by a
    # quarter texel or more
    x_wrong = np.clip((np.arange(nx) + 0.5) * wx / nx - 0.5, 0, wx - 1)
    assert np.max(np.abs(cx * (x - x_wrong))) > 1e-2


def test_gridded_wind_from_rest_one_substep():
    """From rest on a flat periodic lake, one substep gives hu = dt X(cell)
    and hv = dt Y(cell) for the interpolated fields, time weight included."""
    nx, ny, dt, wt = 20, 16, np.float32(5.0), 0.25
    rng = np.random.default_rng(2)
    fields = tuple((1e-4 * rng.standard_normal((5, 6))).astype(np.float32) for _ in range(4))
    p, Hi, Hm, eta, hu, hv = _uniform(nx, ny, 45.0, 0.0, 0.0, f=np.float32(0.0))
    p.wind_grid = fields
    e, u, v = O.cdklm_substep(dt, 0, eta, hu, hv, eta, hu, hv, Hi, Hm, p, wind_t=wt)
    X = [O.wind_texture(a, nx, ny).astype(np.float64) for a in fields]
    wx = (1 - wt) * X[0] + wt * X[1]
    wy = (1 - wt) * X[2] + wt * X[3]
    for got, want in ((u, dt * wx), (v, dt * wy)):
        assert np.max(np.abs(_interior(got) - want)) <= 1e-6 * np.max(np.abs(want))
    assert np.all(_interior(e) == 0.0)
