"""Differentiable CPU restatement of the CDKLM16 step (TEST INFRASTRUCTURE ONLY).

The same arithmetic as ``oracle/cdklm16_np.py`` (gpu_ocean/SWESimulators/
gpu_kernels/CDKLM16_kernel.cu:80-1006) written with torch operations, so that
torch autograd in float64 gives the reference gradient the hand-written HIP
adjoint is checked against.  Branches (minmod selection, min/max of wave
speeds, upwinding, positivity adjustment, dry cells) follow the forward values,
which is what the adjoint does too.  Only tests import this module.

Branch pinning: inside ``pinned(pins)`` every data-dependent decision of the
step (minmod signs and minima, wave-speed extrema, upwinding, positivity and
dry-cell masks) goes through ``_cond``.  A first evaluation in float32 on the
float32 trajectory records the decisions; a second, float64 evaluation replays
them, so torch autograd differentiates the float64 arithmetic along exactly the
branches the float32 computation took.  Without it, states within float32
rounding of a tie (mirrored wall ghosts, minmod ties, u = 0 upwinding) pick
the other branch in float64 and the two gradients differ by the jump.

Scope: constant Coriolis parameter and north vector, wall/periodic
boundaries, RK order 2, no friction -- the adjoint's scope.
"""

from __future__ import annotations

import contextlib

import numpy as np
import torch

from .cdklm16_np import DRY_EPS, DRY_FLAG, PERIODIC, WALL, Params
from .cdklm16_np import _C as _np_consts

_f64 = torch.float64


class BranchPins:
    """Decisions of one substep evaluation, in evaluation order."""

    def __init__(self):
        self.conds = []
        self.replay = False
        self.pos = 0


_pins = None


@contextlib.contextmanager
def pinned(pins: BranchPins):
    """Record (``pins.replay`` False) or replay (True) the decisions of the
    substep evaluations run inside the block."""
    global _pins
    old, _pins = _pins, pins
    pins.pos = 0
    try:
        yield pins
    finally:
        _pins = old
    if pins.replay and pins.pos != len(pins.conds):
        raise RuntimeError(f"replayed {pins.pos} of {len(pins.conds)} recorded decisions")


def _cond(c):
    P = _pins
    if P is None:
        return c
    if not P.replay:
        P.conds.append(c.detach().clone())
        return c
    if P.pos >= len(P.conds):
        raise RuntimeError("more decisions than were recorded")
    r = P.conds[P.pos]
    P.pos += 1
    if r.shape != c.shape:
        raise RuntimeError(f"decision {P.pos - 1}: recorded shape {tuple(r.shape)}, replay {tuple(c.shape)}")
    return r


def _min(a, b):
    if _pins is None:
        return torch.minimum(a, b)
    return torch.where(_cond(a <= b), a, b)


def _max(a, b):
    if _pins is None:
        return torch.maximum(a, b)
    return torch.where(_cond(a >= b), a, b)


def _abs(x):
    if _pins is None:
        return x.abs()
    return torch.where(_cond(x >= 0), x, -x)


def _sign1(x, one):
    """copysign(1, x): the sign bit decides."""
    if _pins is None:
        return torch.copysign(one, x)
    return torch.where(_cond(torch.signbit(x)), -one, one)


def _pos(x):
    """max(x, 0)"""
    if _pins is None:
        return torch.clamp(x, min=0.0)
    return torch.where(_cond(x > 0), x, torch.zeros_like(x))


def _c(x, dtype):
    return torch.tensor(x, dtype=dtype)


def minmod_raw(b, c, f):
    one = torch.ones((), dtype=b.dtype)
    sb, sc, sf = _sign1(b, one), _sign1(c, one), _sign1(f, one)
    m = _min(_min(_abs(b), _abs(c)), _abs(f))
    return ((0.25 * sb) * (sb + sc)) * (sc + sf) * m


def minmod_slope(l, c, r, theta):
    return minmod_raw((c - l) * theta, (r - l) * 0.5, (r - c) * theta)


def desingularize(h, hu, eps):
    a = _div(h * h, 2.0 * eps)
    if a.dtype == torch.float32:
        # 0.5*eps is a double literal in the reference: that addition is in double
        a = (a.to(_f64) + 0.5 * eps).to(torch.float32)
    else:
        a = a + 0.5 * eps
    den = _max(_min(a, torch.full_like(a, eps)), _abs(h))
    return hu / den


def _div(x, c):
    """x / c for a Python constant c, as a true division: torch's device path
    turns division by a host scalar into multiplication by its reciprocal,
    which is not correctly rounded (its CPU path divides)."""
    return x / torch.full((), c, dtype=x.dtype, device=x.device)


def _sqrt(x):
    """Correctly rounded square root: torch's float32 sqrt is not (its CPU
    path is off by an ulp now and then); the float64 root rounded to float32
    is, since 53 >= 2*24 + 2 bits make that double rounding innocuous."""
    if x.dtype == torch.float32:
        return torch.sqrt(x.to(_f64)).to(torch.float32)
    return torch.sqrt(x)


def flux_cu(hm, um, vm, hp, up, vp, g, half_g, cutoff, flux_eps):
    z = torch.zeros_like(hp)
    wet_p = _cond(hp > cutoff)
    wet_m = _cond(hm > cutoff)
    Fp_x = torch.where(wet_p, hp * up, z)
    Fp_y = torch.where(wet_p, (hp * up) * up + (half_g * hp) * hp, z)
    Fp_z = torch.where(wet_p, (hp * up) * vp, z)
    u_p = torch.where(wet_p, up, z)
    c_p = torch.where(wet_p, _sqrt(g * torch.where(wet_p, hp, torch.ones_like(hp))), z)
    Fm_x = torch.where(wet_m, hm * um, z)
    Fm_y = torch.where(wet_m, (hm * um) * um + (half_g * hm) * hm, z)
    Fm_z = torch.where(wet_m, (hm * um) * vm, z)
    u_m = torch.where(wet_m, um, z)
    c_m = torch.where(wet_m, _sqrt(g * torch.where(wet_m, hm, torch.ones_like(hm))), z)
    am = _min(_min(u_m - c_m, u_p - c_p), z)
    ap = _max(_max(u_m + c_m, u_p + c_p), z)
    den = ap - am
    zero = _cond((ap - am).abs() < flux_eps)
    den_s = torch.where(zero, torch.ones_like(den), den)
    apam = ap * am
    Fx = ((ap * Fm_x - am * Fp_x) + apam * (hp - hm)) / den_s
    Fy = ((ap * Fm_y - am * Fp_y) + apam * (Fp_x - Fm_x)) / den_s
    Fz = torch.where(_cond(um + up > 0), Fm_z, Fp_z)
    return (torch.where(zero, z, Fx), torch.where(zero, z, Fy), torch.where(zero, z, Fz))


def bc_view(eta, hu, hv, p: Params):
    """Padded arrays whose ghost cells are the mirrored (walls) or wrapped
    (periodic) images of interior cells, as the forward kernel reads them;
    differentiable (built by indexing, not in-place writes)."""
    nx, ny = p.nx, p.ny
    dev = eta.device
    rows = torch.arange(ny + 4, device=dev)
    cols = torch.arange(nx + 4, device=dev)
    sr = rows.clone()
    sc = cols.clone()
    sgn_v = torch.ones(ny + 4, dtype=eta.dtype, device=dev)
    sgn_u = torch.ones(nx + 4, dtype=eta.dtype, device=dev)
    if p.bc_south == WALL:
        sr[0], sr[1] = 3, 2
        sgn_v[0] = sgn_v[1] = -1
    elif p.bc_south == PERIODIC:
        sr[0], sr[1] = ny, ny + 1
    if p.bc_north == WALL:
        sr[ny + 2], sr[ny + 3] = ny + 1, ny
        sgn_v[ny + 2] = sgn_v[ny + 3] = -1
    elif p.bc_north == PERIODIC:
        sr[ny + 2], sr[ny + 3] = 2, 3
    if p.bc_west == WALL:
        sc[0], sc[1] = 3, 2
        sgn_u[0] = sgn_u[1] = -1
    elif p.bc_west == PERIODIC:
        sc[0], sc[1] = nx, nx + 1
    if p.bc_east == WALL:
        sc[nx + 2], sc[nx + 3] = nx + 1, nx
        sgn_u[nx + 2] = sgn_u[nx + 3] = -1
    elif p.bc_east == PERIODIC:
        sc[nx + 2], sc[nx + 3] = 2, 3
    E = eta[sr][:, sc]
    U = hu[sr][:, sc] * sgn_u[None, :]
    V = hv[sr][:, sc] * sgn_v[:, None]
    return E, U, V


def substep(dt, rk_step, eta_in, hu_in, hv_in, eta_qn, hu_qn, hv_qn, Hi, Hm, p: Params):
    """One cdklm_swe_2D call on padded torch tensors; returns interior-updated
    copies of the *_qn arrays (their ghost cells are left as they are)."""
    dtp = eta_in.dtype
    nx, ny = p.nx, p.ny
    # constants rounded to the working precision, and their products formed
    # in it, as the kernel's float constants are (exact as Python floats)
    t = np.float32 if dtp == torch.float32 else np.float64
    C = _np_consts(p, t)
    g, theta = float(C.GRAV), float(C.THETA)
    two_g, half_g, theta_g = float(C.TWO_G), float(C.HALF_G), float(C.THETA_G)
    cutoff, desing, feps = float(C.CUTOFF), float(C.DESING), float(C.FLUX_EPS)
    dx, dy = float(C.DX), float(C.DY)
    f = float(t(p.f))
    dxf, dyf = float(C.DX * t(p.f)), float(C.DY * t(p.f))
    n_x, n_y = float(t(p.north_x)), float(t(p.north_y))
    dt = float(t(dt))
    e_x, e_y = n_y, -n_x
    land_v = float(p.land_value)
    R0, R1, R2 = bc_view(eta_in, hu_in, hv_in, p)
    Hif = torch.where((Hi - land_v).abs() < DRY_EPS, torch.full_like(Hi, DRY_FLAG), Hi)
    h = R0 + Hm
    land = (Hm - land_v).abs() <= DRY_EPS
    shallow = (~land) & _cond(h < desing)
    dry = shallow & _cond(h <= cutoff)
    des = shallow & ~dry
    wet = (~land) & ~shallow
    hs = torch.where(wet | des, h, torch.ones_like(h))
    U = torch.where(wet, R1 / hs, torch.where(des, desingularize(hs, R1, desing), torch.zeros_like(h)))
    V = torch.where(wet, R2 / hs, torch.where(des, desingularize(hs, R2, desing), torch.zeros_like(h)))
    E = torch.where(land, torch.full_like(h, DRY_FLAG), torch.where(dry, -Hm + cutoff, R0))
    I = slice(2, nx + 2)
    J = slice(2, ny + 2)
    Hm_reg = 0.25 * (((Hif[2:ny + 2, 2:nx + 2] + Hif[3:ny + 3, 2:nx + 2]) + Hif[2:ny + 2, 3:nx + 3])
                     + Hif[3:ny + 3, 3:nx + 3])
    Ec = E[J, I]
    hreg = Ec + Hm_reg
    hreg_wet = _cond(hreg > cutoff)
    hu_c = torch.where(hreg_wet, U[J, I] * hreg, torch.zeros_like(hreg))
    hv_c = torch.where(hreg_wet, V[J, I] * hreg, torch.zeros_like(hreg))

    # ---- x ----
    Xc, Xl, Xr = slice(1, nx + 3), slice(0, nx + 2), slice(2, nx + 4)
    ux = minmod_slope(U[J, Xl], U[J, Xc], U[J, Xr], theta)
    vx = minmod_slope(V[J, Xl], V[J, Xc], V[J, Xr], theta)
    flipc = torch.ones(nx + 4, dtype=dtp, device=eta_in.device)
    if p.bc_west == WALL:
        flipc[0:2] = -1
    if p.bc_east == WALL:
        flipc[nx + 2:nx + 4] = -1
    VK = V * flipc[None, :]
    fv_l = (n_x * U[J, Xl] + n_y * VK[J, Xl]) * f
    fv_c = (n_x * U[J, Xc] + n_y * VK[J, Xc]) * f
    fv_r = (n_x * U[J, Xr] + n_y * VK[J, Xr]) * f
    Vd = float(C.DX / C.TWO_G)
    el, ec, er = E[J, Xl], E[J, Xc], E[J, Xr]
    Kx = minmod_raw(theta_g * ((ec - el) - Vd * (fv_c + fv_l)),
                    half_g * ((er - el) - Vd * ((fv_r + 2.0 * fv_c) + fv_l)),
                    theta_g * ((er - ec) - Vd * (fv_c + fv_r)))
    vadj = V[J, Xc]
    if p.bc_west == WALL:
        vadj = torch.cat([-vadj[:, 0:1], vadj[:, 1:]], 1)
    dxfv = dxf * vadj
    H_west = 0.5 * (Hif[2:ny + 2, 1:nx + 3] + Hif[3:ny + 3, 1:nx + 3])
    H_east = 0.5 * (Hif[2:ny + 2, 2:nx + 4] + Hif[3:ny + 3, 2:nx + 4])
    tt = _div(Kx + dxfv, two_g)
    hw = (ec + H_west) - tt
    he = (ec + H_east) + tt
    Kx = torch.where(_cond(hw > 0), Kx, -dxfv + two_g * (ec + H_west))
    Kx = torch.where(_cond(he > 0), Kx, -dxfv - two_g * (ec + H_east))

    def fface(cm, ncol):
        cp = cm + 1
        eta_p, eta_m = E[J, cp], E[J, cm]
        up, um, vp, vm = U[J, cp], U[J, cm], V[J, cp], V[J, cm]
        Rpx = up - 0.5 * ux[:, cp - 1]
        Rpy = vp - 0.5 * vx[:, cp - 1]
        Rmx = um + 0.5 * ux[:, cm - 1]
        Rmy = vm + 0.5 * vx[:, cm - 1]
        Hf = 0.5 * (Hif[2:ny + 2, cp] + Hif[3:ny + 3, cp])
        Kp, Km = Kx[:, cp - 1], Kx[:, cm - 1]
        vmf = vm * torch.where((cp == 2) & (p.bc_west == WALL), -1.0, 1.0).to(dtp)
        vpf = vp * torch.where((cp == nx + 2) & (p.bc_east == WALL), -1.0, 1.0).to(dtp)
        vpn = up * n_x + vpf * n_y
        vmn = um * n_x + vmf * n_y
        hp = _pos((eta_p + Hf) - _div(Kp + dxf * vpn, two_g))
        hm = _pos((eta_m + Hf) + _div(Km + dxf * vmn, two_g))
        F = flux_cu(hm, Rmx, Rmy, hp, Rpx, Rpy, g, half_g, cutoff, feps)
        dryf = _cond((eta_p == DRY_FLAG) | (eta_m == DRY_FLAG))
        return tuple(torch.where(dryf, torch.zeros_like(x), x) for x in F)

    cols = torch.arange(2, nx + 2, device=eta_in.device)
    Fe = fface(cols, cols)
    Fw = fface(cols - 1, cols)
    fdx = [_div(Fe[k] - Fw[k], dx) for k in range(3)]

    # ---- y ----
    Yc, Yl, Yu = slice(1, ny + 3), slice(0, ny + 2), slice(2, ny + 4)
    uy = minmod_slope(U[Yl, I], U[Yc, I], U[Yu, I], theta)
    vy = minmod_slope(V[Yl, I], V[Yc, I], V[Yu, I], theta)
    flipr = torch.ones(ny + 4, dtype=dtp, device=eta_in.device)
    if p.bc_south == WALL:
        flipr[0:2] = -1
    if p.bc_north == WALL:
        flipr[ny + 2:ny + 4] = -1
    UK = U * flipr[:, None]
    fu_lo = (e_x * UK[Yl, I] + e_y * V[Yl, I]) * f
    fu_c = (e_x * UK[Yc, I] + e_y * V[Yc, I]) * f
    fu_up = (e_x * UK[Yu, I] + e_y * V[Yu, I]) * f
    Ud = float(C.DY / C.TWO_G)
    elo, ecy, eup = E[Yl, I], E[Yc, I], E[Yu, I]
    Ly = minmod_raw(theta_g * ((ecy - elo) + Ud * (fu_c + fu_lo)),
                    half_g * ((eup - elo) + Ud * ((fu_up + 2.0 * fu_c) + fu_lo)),
                    theta_g * ((eup - ecy) + Ud * (fu_c + fu_up)))
    uadj = U[Yc, I]
    if p.bc_south == WALL:
        uadj = torch.cat([-uadj[0:1, :], uadj[1:, :]], 0)
    dyfu = dyf * uadj
    H_south = 0.5 * (Hif[1:ny + 3, 2:nx + 2] + Hif[1:ny + 3, 3:nx + 3])
    H_north = 0.5 * (Hif[2:ny + 4, 2:nx + 2] + Hif[2:ny + 4, 3:nx + 3])
    tt = _div(Ly - dyfu, two_g)
    hs_ = (ecy + H_south) - tt
    hn_ = (ecy + H_north) + tt
    Ly = torch.where(_cond(hs_ > 0), Ly, dyfu + two_g * (ecy + H_south))
    Ly = torch.where(_cond(hn_ > 0), Ly, dyfu - two_g * (ecy + H_north))

    def gface(rm, nrow):
        rp = rm + 1
        eta_p, eta_m = E[rp, I], E[rm, I]
        up, um, vp, vm = U[rp, I], U[rm, I], V[rp, I], V[rm, I]
        Rpx = up - 0.5 * uy[rp - 1, :]
        Rpy = vp - 0.5 * vy[rp - 1, :]
        Rmx = um + 0.5 * uy[rm - 1, :]
        Rmy = vm + 0.5 * vy[rm - 1, :]
        Hf = 0.5 * (Hif[rp, 2:nx + 2] + Hif[rp, 3:nx + 3])
        Lp, Lm = Ly[rp - 1, :], Ly[rm - 1, :]
        umf = um * torch.where((rp == 2) & (p.bc_south == WALL), -1.0, 1.0).to(dtp)[:, None]
        upf = up * torch.where((rp == ny + 2) & (p.bc_north == WALL), -1.0, 1.0).to(dtp)[:, None]
        upe = upf * e_x + vp * e_y
        ume = umf * e_x + vm * e_y
        hp = _pos((eta_p + Hf) - _div(Lp - dyf * upe, two_g))
        hm = _pos((eta_m + Hf) + _div(Lm - dyf * ume, two_g))
        a, b, c = flux_cu(hm, Rmy, Rmx, hp, Rpy, Rpx, g, half_g, cutoff, feps)
        dryf = _cond((eta_p == DRY_FLAG) | (eta_m == DRY_FLAG))
        z = torch.zeros_like(a)
        return torch.where(dryf, z, a), torch.where(dryf, z, c), torch.where(dryf, z, b)

    rows = torch.arange(2, ny + 2, device=eta_in.device)
    Gn = gface(rows, rows)
    Gs = gface(rows - 1, rows)
    fd = [fdx[k] + _div(Gn[k] - Gs[k], dy) for k in range(3)]

    tx = _div(Kx[:, 1:nx + 1] + dxf * V[J, I], two_g)
    ty = _div(Ly[1:ny + 1, :] - dyf * U[J, I], two_g)
    eta_we = 0.5 * ((Ec - tx) + (Ec + tx))
    eta_sn = 0.5 * ((Ec - ty) + (Ec + ty))
    RHxp = 0.5 * (Hif[2:ny + 2, 3:nx + 3] + Hif[3:ny + 3, 3:nx + 3])
    RHxm = 0.5 * (Hif[2:ny + 2, 2:nx + 2] + Hif[3:ny + 3, 2:nx + 2])
    RHyp = 0.5 * (Hif[3:ny + 3, 2:nx + 2] + Hif[3:ny + 3, 3:nx + 3])
    RHym = 0.5 * (Hif[2:ny + 2, 2:nx + 2] + Hif[2:ny + 2, 3:nx + 3])
    b1 = (g * (eta_we + Hm_reg)) * (RHxp - RHxm)
    b2 = (g * (eta_sn + Hm_reg)) * (RHyp - RHym)
    hu_east = hu_c * e_x + hv_c * e_y
    hv_north = hu_c * n_x + hv_c * n_y
    hu_ec = f * hv_north
    hv_nc = -f * hu_east
    up_x, up_y = -n_x, n_y
    hu_cor = up_y * hu_ec + (-up_x) * hv_nc
    hv_cor = up_x * hu_ec + up_y * hv_nc
    X, Y = float(t(p.wind_x[0])), float(t(p.wind_y[0]))
    active = _cond((hreg >= cutoff) & (Ec != DRY_FLAG))
    st1 = torch.where(active, (X + hu_cor) + _div(b1, dx), torch.zeros_like(hreg))
    st2 = torch.where(active, (Y + hv_cor) + _div(b2, dy), torch.zeros_like(hreg))
    L1 = -fd[0]
    L2 = -fd[1] + st1
    L3 = -fd[2] + st2
    if rk_step == 0:
        ue = Ec + dt * L1
        uu = hu_c + dt * L2
        uv = hv_c + dt * L3
    else:
        ue = 0.5 * (eta_qn[J, I] + (Ec + dt * L1))
        uu = 0.5 * (hu_qn[J, I] + (hu_c + dt * L2))
        uv = 0.5 * (hv_qn[J, I] + (hv_c + dt * L3))
    # dry clamp of the update (CDKLM16_kernel.cu:986-1001)
    floor = -Hm_reg + cutoff
    dry_after = _cond((ue + Hm_reg) <= cutoff)
    zero = torch.zeros_like(ue)
    ue = _max(floor, torch.where(dry_after, floor, ue))
    uu = torch.where(dry_after, zero, uu)
    uv = torch.where(dry_after, zero, uv)
    out = []
    for base, val in zip((eta_qn, hu_qn, hv_qn), (ue, uu, uv)):
        top = base[:2, :]
        bot = base[ny + 2:, :]
        mid = torch.cat([base[J, :2], val, base[J, nx + 2:]], 1)
        out.append(torch.cat([top, mid, bot], 0))
    return tuple(out)


def rk2_steps(eta, hu, hv, Hi, Hm, p: Params, dt, n):
    """n fixed-dt RK2 steps from (eta, hu, hv) (ghost cells are not used: the
    boundaries are applied by index remapping)."""
    q = (eta, hu, hv)
    for _ in range(n):
        qs = substep(dt, 0, *q, *q, Hi, Hm, p)
        q = substep(dt, 1, *qs, *q, Hi, Hm, p)
    return q
