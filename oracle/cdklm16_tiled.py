"""Literal thread-block emulation of ``cdklm_swe_2D`` (TEST INFRASTRUCTURE ONLY).

A second, independent CPU restatement of gpu_ocean/SWESimulators/gpu_kernels/
CDKLM16_kernel.cu:429-1006: it walks the launch grid block by block and, inside
a block, thread by thread between the ``__syncthreads()`` of the reference,
with the shared-memory arrays R, Qx and Hi at their reference shapes and the
same index arithmetic.  Scalar ``numpy.float32`` arithmetic is IEEE correctly
rounded, and ``--fmad=false`` means the reference rounds every product too.

It is slow (pure Python) and meant for grids of a few hundred cells.  The
tests use it to pin ``oracle/cdklm16_np.py`` (whole-grid vectorised form), and
to document the one place where the two differ: the reference reads an
unmirrored global ghost row when (ny-1) is a multiple of the block height (or
(nx-1) of the block width) and that side is a wall.
"""

from __future__ import annotations

import math

import numpy as np

from .cdklm16_np import DRY_FLAG, DRY_EPS, WALL, Params

f32 = np.float32


def _copysign1(x):
    return f32(math.copysign(1.0, float(x)))


def _minmod_raw(b, c, f):
    """common.cu:544-551"""
    sb, sc, sf = _copysign1(b), _copysign1(c), _copysign1(f)
    m = min(min(abs(b), abs(c)), abs(f))
    return ((f32(0.25) * sb) * (sb + sc)) * (sc + sf) * m


def _minmod_slope(l, c, r, theta):
    """common.cu:532-542"""
    return _minmod_raw((c - l) * theta, (r - l) * f32(0.5), (r - c) * theta)


def _desing(h, hu, eps):
    """common.cu:822-824 (``0.5*eps`` is a double literal)."""
    a = (h * h) / (f32(2.0) * eps)
    a = f32(float(a) + 0.5 * float(eps))
    return hu / max(min(a, eps), abs(h))


def _clampi(v, lo, hi):
    return max(lo, min(v, hi))


def tiled_reference_substep(dt, rk_step, eta0, hu0, hv0, eta1, hu1, hv1,
                            Hi, Hm, p: Params, block=(4, 3), wind_t=0.0):
    """Emulate one launch of cdklm_swe_2D with block (bw, bh).  Arrays are
    padded (ny+4, nx+4) float32; (eta1,hu1,hv1) hold Q^n for rk_step 1.
    Returns new output arrays.  RK order 1 and 2 only."""
    bw, bh = block
    nx, ny = p.nx, p.ny
    NX, NY = nx, ny
    DX, DY, GRAV = f32(p.dx), f32(p.dy), f32(p.g)
    THETA = f32(p.theta)
    DESING, FLUX_EPS, CUTOFF = f32(p.desing_eps), f32(p.flux_slope_eps), f32(p.depth_cutoff)
    FRIC = f32(p.r)
    land_value = f32(p.land_value)
    dt = f32(dt)
    two_g = f32(2.0) * GRAV
    fld = lambda v: np.full((ny + 4, nx + 4), v, np.float32) if np.ndim(v) == 0 else np.asarray(v, np.float32)
    cor = fld(p.f)
    nox = fld(p.north_x)
    noy = fld(p.north_y)

    def coriolisF(i, j):
        # texture lookup at a texel centre == the per-cell value (clamped)
        return cor[_clampi(j, 0, ny + 3), _clampi(i, 0, nx + 3)]

    def getNorth(i, j):
        jj, ii = _clampi(j, 0, ny + 3), _clampi(i, 0, nx + 3)
        return (nox[jj, ii], noy[jj, ii])

    X = f32(wind_t) * f32(p.wind_x[1]) + (f32(1.0) - f32(wind_t)) * f32(p.wind_x[0])
    Y = f32(wind_t) * f32(p.wind_y[1]) + (f32(1.0) - f32(wind_t)) * f32(p.wind_y[0])

    bc_north, bc_east, bc_south, bc_west = p.bc
    packed = (bc_north << 24) | (bc_south << 16) | (bc_east << 8) | bc_west

    out = [np.array(a, dtype=np.float32, copy=True) for a in (eta1, hu1, hv1)]
    ins = [np.asarray(a, dtype=np.float32) for a in (eta0, hu0, hv0)]

    def F_func(Q):
        h, u, v = Q
        return (h * u, h * u * u + f32(0.5) * GRAV * h * h, h * u * v)

    def flux(Qm, Qp):
        """:101-140"""
        Fp = (f32(0), f32(0), f32(0)); up = f32(0); cp = f32(0)
        if Qp[0] > CUTOFF:
            Fp = F_func(Qp); up = Qp[1]; cp = f32(math.sqrt(float(GRAV * Qp[0])))
        Fm = (f32(0), f32(0), f32(0)); um = f32(0); cm = f32(0)
        if Qm[0] > CUTOFF:
            Fm = F_func(Qm); um = Qm[1]; cm = f32(math.sqrt(float(GRAV * Qm[0])))
        am = min(min(um - cm, up - cp), f32(0))
        ap = max(max(um + cm, up + cp), f32(0))
        if abs(ap - am) < FLUX_EPS:
            return (f32(0), f32(0), f32(0))
        Fx = ((ap * Fm[0] - am * Fp[0]) + ap * am * (Qp[0] - Qm[0])) / (ap - am)
        Fy = ((ap * Fm[1] - am * Fp[1]) + ap * am * (Fp[0] - Fm[0])) / (ap - am)
        Fz = Fm[2] if (Qm[1] + Qp[1] > 0) else Fp[2]
        return (Fx, Fy, Fz)

    gx = -(-nx // bw)
    gy = -(-ny // bh)
    for byi in range(gy):
        for bxi in range(gx):
            bx = bw * bxi
            by = bh * byi
            R = np.zeros((3, bh + 4, bw + 4), np.float32)
            Qx = np.zeros((3, bh + 2, bw + 2), np.float32)
            Hs = np.zeros((bh + 3, bw + 3), np.float32)
            threads = [(tx, ty) for ty in range(bh) for tx in range(bw)]
            # read state (:490-506)
            for j in range(bh + 4):
                l = _clampi(by + j, 0, NY + 3)
                for i in range(bw + 4):
                    k = _clampi(bx + i, 0, NX + 3)
                    for c in range(3):
                        R[c, j, i] = ins[c][l, k]
            # read Hi (:511-524)
            for j in range(bh + 3):
                l = _clampi(by + j + 1, 1, NY + 4)
                for i in range(bw + 3):
                    k = _clampi(bx + i + 1, 1, NX + 4)
                    v = f32(Hi[l, k])
                    Hs[j, i] = f32(DRY_FLAG) if abs(v - land_value) < f32(DRY_EPS) else v
            # wall BCs (:535-542 -> handleWallBC :373-424)
            if packed > 0:
                for tx, ty in threads:
                    ti = bx + tx + 2; tj = by + ty + 2
                    i = tx + 2; j = ty + 2
                    if bc_north == WALL and tj == NY + 1:
                        R[0, j + 1, i] = R[0, j, i]; R[1, j + 1, i] = R[1, j, i]; R[2, j + 1, i] = -R[2, j, i]
                        R[0, j + 2, i] = R[0, j - 1, i]; R[1, j + 2, i] = R[1, j - 1, i]; R[2, j + 2, i] = -R[2, j - 1, i]
                    if bc_south == WALL and tj == 2:
                        R[0, j - 1, i] = R[0, j, i]; R[1, j - 1, i] = R[1, j, i]; R[2, j - 1, i] = -R[2, j, i]
                        R[0, j - 2, i] = R[0, j + 1, i]; R[1, j - 2, i] = R[1, j + 1, i]; R[2, j - 2, i] = -R[2, j + 1, i]
                    if bc_east == WALL and ti == NX + 1:
                        R[0, j, i + 1] = R[0, j, i]; R[1, j, i + 1] = -R[1, j, i]; R[2, j, i + 1] = R[2, j, i]
                        R[0, j, i + 2] = R[0, j, i - 1]; R[1, j, i + 2] = -R[1, j, i - 1]; R[2, j, i + 2] = R[2, j, i - 1]
                    if bc_west == WALL and ti == 2:
                        R[0, j, i - 1] = R[0, j, i]; R[1, j, i - 1] = -R[1, j, i]; R[2, j, i - 1] = R[2, j, i]
                        R[0, j, i - 2] = R[0, j, i + 1]; R[1, j, i - 2] = -R[1, j, i + 1]; R[2, j, i - 2] = R[2, j, i + 1]
            # register Hm (:547)
            Hm_reg = {}
            for tx, ty in threads:
                Hm_reg[(tx, ty)] = f32(0.25) * (((Hs[ty + 1, tx + 1] + Hs[ty + 2, tx + 1]) + Hs[ty + 1, tx + 2]) + Hs[ty + 2, tx + 2])
            # conversion (:551-592)
            for j in range(bh + 4):
                l = _clampi(by + j, 0, NY + 3)
                for i in range(bw + 4):
                    k = _clampi(bx + i, 0, NX + 3)
                    lHm = f32(Hm[l, k])
                    h = R[0, j, i] + lHm
                    if abs(lHm - land_value) <= f32(DRY_EPS):
                        R[0, j, i] = f32(DRY_FLAG); R[1, j, i] = 0; R[2, j, i] = 0
                    elif h < DESING:
                        if h <= CUTOFF:
                            R[0, j, i] = -lHm + CUTOFF; R[1, j, i] = 0; R[2, j, i] = 0
                        else:
                            R[1, j, i] = _desing(h, R[1, j, i], DESING)
                            R[2, j, i] = _desing(h, R[2, j, i], DESING)
                    else:
                        R[1, j, i] = R[1, j, i] / h
                        R[2, j, i] = R[2, j, i] / h
            # hu, hv (:597-602)
            huv = {}
            for tx, ty in threads:
                Hmr = Hm_reg[(tx, ty)]
                e = R[0, ty + 2, tx + 2]
                if (e + Hmr) > CUTOFF:
                    huv[(tx, ty)] = (R[1, ty + 2, tx + 2] * (e + Hmr), R[2, ty + 2, tx + 2] * (e + Hmr))
                else:
                    huv[(tx, ty)] = (f32(0), f32(0))
            # x slopes (:610-665)
            for j in range(bh):
                l = j + 2
                for i in range(bw + 2):
                    k = i + 1
                    le, ce, re = R[0, l, k - 1], R[0, l, k], R[0, l, k + 1]
                    lu, cu, ru = R[1, l, k - 1], R[1, l, k], R[1, l, k + 1]
                    lv, cv, rv = R[2, l, k - 1], R[2, l, k], R[2, l, k + 1]
                    Qx[0, j, i] = _minmod_slope(lu, cu, ru, THETA)
                    Qx[1, j, i] = _minmod_slope(lv, cv, rv, THETA)
                    gid = bx + i + 1
                    if bc_west == 1:
                        if gid < 3: lv = -lv
                        if gid < 2: cv = -cv
                    if bc_east == 1:
                        if gid > NX: rv = -rv
                        if gid > NX + 1: cv = -cv
                    nth = getNorth(bx + k, by + l)
                    lf, cf, rf = coriolisF(bx + k - 1, by + l), coriolisF(bx + k, by + l), coriolisF(bx + k + 1, by + l)
                    lfv = (nth[0] * lu + nth[1] * lv) * lf
                    cfv = (nth[0] * cu + nth[1] * cv) * cf
                    rfv = (nth[0] * ru + nth[1] * rv) * rf
                    Vc = DX / (f32(2.0) * GRAV)
                    b_ = THETA * GRAV * (ce - le - Vc * (cfv + lfv))
                    c_ = f32(0.5) * GRAV * (re - le - Vc * (rfv + f32(2.0) * cfv + lfv))
                    f_ = THETA * GRAV * (re - ce - Vc * (cfv + rfv))
                    Qx[2, j, i] = _minmod_raw(b_, c_, f_)
            # adjustSlopes_x (:148-191)
            for j in range(bh):
                l = j + 2; H_j = j + 1
                for i in range(bw + 2):
                    k = i + 1; H_i = i
                    eta = R[0, l, k]; v = R[2, l, k]
                    if bc_west == 1 and bx + k < 2: v = -v
                    if bc_east == 1 and bx + k > NX + 2: v = -v
                    cfv = DX * coriolisF(bx + k, by + l) * v
                    Hw = f32(0.5) * (Hs[H_j, H_i] + Hs[H_j + 1, H_i])
                    He = f32(0.5) * (Hs[H_j, H_i + 1] + Hs[H_j + 1, H_i + 1])
                    hw = eta + Hw - (Qx[2, j, i] + cfv) / (f32(2.0) * GRAV)
                    he = eta + He + (Qx[2, j, i] + cfv) / (f32(2.0) * GRAV)
                    Qx[2, j, i] = Qx[2, j, i] if hw > 0 else -cfv + f32(2.0) * GRAV * (eta + Hw)
                    Qx[2, j, i] = Qx[2, j, i] if he > 0 else -cfv - f32(2.0) * GRAV * (eta + He)

            def Fface(i, j, cfm, cfp, north):
                """computeFFaceFlux :246-304"""
                l = j + 2; k = i + 1; H_i = i + 1; H_j = j + 1
                ebp, ebm = R[0, l, k + 1], R[0, l, k]
                up, um = R[1, l, k + 1], R[1, l, k]
                vp, vm = R[2, l, k + 1], R[2, l, k]
                if ebp == f32(DRY_FLAG) or ebm == f32(DRY_FLAG):
                    return (f32(0), f32(0), f32(0))
                Rp = (up - f32(0.5) * Qx[0, j, i + 1], vp - f32(0.5) * Qx[1, j, i + 1])
                Rm = (um + f32(0.5) * Qx[0, j, i], vm + f32(0.5) * Qx[1, j, i])
                Hf = f32(0.5) * (Hs[H_j, H_i] + Hs[H_j + 1, H_i])
                Kp, Km = Qx[2, j, i + 1], Qx[2, j, i]
                if bc_west == 1 and bx + i + 2 == 2: vm = -vm
                if bc_east == 1 and bx + i + 2 == NX + 2: vp = -vp
                vpn = up * north[0] + vp * north[1]
                vmn = um * north[0] + vm * north[1]
                hp = max(f32(0), ebp + Hf - (Kp + DX * cfp * vpn) / (f32(2.0) * GRAV))
                hm = max(f32(0), ebm + Hf + (Km + DX * cfm * vmn) / (f32(2.0) * GRAV))
                return flux((hm, Rm[0], Rm[1]), (hp, Rp[0], Rp[1]))

            fd = {}
            etaWE = {}
            for tx, ty in threads:
                ti = bx + tx + 2; tj = by + ty + 2
                cfc = coriolisF(ti, tj)
                north = getNorth(ti, tj)
                Fe = Fface(tx + 1, ty, cfc, coriolisF(ti + 1, tj), north)
                Fw = Fface(tx, ty, coriolisF(ti - 1, tj), cfc, north)
                fd[(tx, ty)] = tuple((Fe[c] - Fw[c]) / DX for c in range(3))
                tmp = (Qx[2, ty, tx + 1] + DX * cfc * R[2, ty + 2, tx + 2]) / (f32(2.0) * GRAV)
                etaWE[(tx, ty)] = (R[0, ty + 2, tx + 2] - tmp, R[0, ty + 2, tx + 2] + tmp)
            # y slopes (:715-769)
            for j in range(bh + 2):
                l = j + 1
                for i in range(bw):
                    k = i + 2
                    le, ce, ue = R[0, l - 1, k], R[0, l, k], R[0, l + 1, k]
                    lu, cu, uu = R[1, l - 1, k], R[1, l, k], R[1, l + 1, k]
                    lv, cv, uv = R[2, l - 1, k], R[2, l, k], R[2, l + 1, k]
                    Qx[0, j, i] = _minmod_slope(lu, cu, uu, THETA)
                    Qx[1, j, i] = _minmod_slope(lv, cv, uv, THETA)
                    gid = by + j + 1
                    if bc_south == 1:
                        if gid < 3: lu = -lu
                        if gid < 2: cu = -cu
                    if bc_north == 1:
                        if gid > NY: uu = -uu
                        if gid > NY + 1: cu = -cu
                    nth = getNorth(bx + k, by + l)
                    east = (nth[1], -nth[0])
                    lf, cf, uf = coriolisF(bx + k, by + l - 1), coriolisF(bx + k, by + l), coriolisF(bx + k, by + l + 1)
                    lfu = (east[0] * lu + east[1] * lv) * lf
                    cfu = (east[0] * cu + east[1] * cv) * cf
                    ufu = (east[0] * uu + east[1] * uv) * uf
                    Uc = DY / (f32(2.0) * GRAV)
                    b_ = THETA * GRAV * (ce - le + Uc * (cfu + lfu))
                    c_ = f32(0.5) * GRAV * (ue - le + Uc * (ufu + f32(2.0) * cfu + lfu))
                    f_ = THETA * GRAV * (ue - ce + Uc * (cfu + ufu))
                    Qx[2, j, i] = _minmod_raw(b_, c_, f_)
            # adjustSlopes_y (:198-241)
            for j in range(bh + 2):
                l = j + 1; H_j = j
                for i in range(bw):
                    k = i + 2; H_i = i + 1
                    eta = R[0, l, k]; u = R[1, l, k]
                    if bc_south == 1 and by + l < 2: u = -u
                    if bc_north == 1 and by + l > NY + 2: u = -u
                    dyfu = DY * coriolisF(bx + k, by + l) * u
                    Hso = f32(0.5) * (Hs[H_j, H_i] + Hs[H_j, H_i + 1])
                    Hno = f32(0.5) * (Hs[H_j + 1, H_i] + Hs[H_j + 1, H_i + 1])
                    hs = eta + Hso - (Qx[2, j, i] - dyfu) / (f32(2.0) * GRAV)
                    hn = eta + Hno + (Qx[2, j, i] - dyfu) / (f32(2.0) * GRAV)
                    Qx[2, j, i] = Qx[2, j, i] if hs > 0 else dyfu + f32(2.0) * GRAV * (eta + Hso)
                    Qx[2, j, i] = Qx[2, j, i] if hn > 0 else dyfu - f32(2.0) * GRAV * (eta + Hno)

            def Gface(i, j, cfm, cfp, east):
                """computeGFaceFlux :309-370"""
                l = j + 1; k = i + 2; H_i = i + 1; H_j = j + 1
                ebp, ebm = R[0, l + 1, k], R[0, l, k]
                up, um = R[1, l + 1, k], R[1, l, k]
                vp, vm = R[2, l + 1, k], R[2, l, k]
                if ebp == f32(DRY_FLAG) or ebm == f32(DRY_FLAG):
                    return (f32(0), f32(0), f32(0))
                Rp = (up - f32(0.5) * Qx[0, j + 1, i], vp - f32(0.5) * Qx[1, j + 1, i])
                Rm = (um + f32(0.5) * Qx[0, j, i], vm + f32(0.5) * Qx[1, j, i])
                Hf = f32(0.5) * (Hs[H_j, H_i] + Hs[H_j, H_i + 1])
                Lp, Lm = Qx[2, j + 1, i], Qx[2, j, i]
                if bc_south == 1 and by + j + 2 == 2: um = -um
                if bc_north == 1 and by + j + 2 == NY + 2: up = -up
                upe = up * east[0] + vp * east[1]
                ume = um * east[0] + vm * east[1]
                hp = max(f32(0), ebp + Hf - (Lp - DY * cfp * upe) / (f32(2.0) * GRAV))
                hm = max(f32(0), ebm + Hf + (Lm - DY * cfm * ume) / (f32(2.0) * GRAV))
                F = flux((hm, Rm[1], Rm[0]), (hp, Rp[1], Rp[0]))
                return (F[0], F[2], F[1])

            for tx, ty in threads:
                ti = bx + tx + 2; tj = by + ty + 2
                cfc = coriolisF(ti, tj)
                north = getNorth(ti, tj)
                east = (north[1], -north[0])
                Gn = Gface(tx, ty + 1, cfc, coriolisF(ti, tj + 1), east)
                Gs = Gface(tx, ty, coriolisF(ti, tj - 1), cfc, east)
                fdx = fd[(tx, ty)]
                fdt = tuple(fdx[c] + (Gn[c] - Gs[c]) / DY for c in range(3))
                tmp = (Qx[2, ty + 1, tx] - DY * cfc * R[1, ty + 2, tx + 2]) / (f32(2.0) * GRAV)
                eta_s = R[0, ty + 2, tx + 2] - tmp
                eta_n = R[0, ty + 2, tx + 2] + tmp
                if not (ti > 1 and ti < NX + 2 and tj > 1 and tj < NY + 2):
                    continue
                i = tx + 2; j = ty + 2; H_i = tx + 1; H_j = ty + 1
                Hmr = Hm_reg[(tx, ty)]
                hu, hv = huv[(tx, ty)]
                st1 = f32(0); st2 = f32(0)
                h = R[0, j, i] + Hmr
                if h >= CUTOFF and R[0, j, i] != f32(DRY_FLAG):
                    RHxp = f32(0.5) * (Hs[H_j, H_i + 1] + Hs[H_j + 1, H_i + 1])
                    RHxm = f32(0.5) * (Hs[H_j, H_i] + Hs[H_j + 1, H_i])
                    RHyp = f32(0.5) * (Hs[H_j + 1, H_i] + Hs[H_j + 1, H_i + 1])
                    RHym = f32(0.5) * (Hs[H_j, H_i] + Hs[H_j, H_i + 1])
                    H_x = RHxp - RHxm; H_y = RHyp - RHym
                    eta_sn = f32(0.5) * (eta_n + eta_s)
                    eta_we = f32(0.5) * (etaWE[(tx, ty)][0] + etaWE[(tx, ty)][1])
                    b1 = GRAV * (eta_we + Hmr) * H_x
                    b2 = GRAV * (eta_sn + Hmr) * H_y
                    hu_e = hu * east[0] + hv * east[1]
                    hv_n = hu * north[0] + hv * north[1]
                    hu_ec = cfc * hv_n
                    hv_nc = -cfc * hu_e
                    upv = (-north[0], north[1])
                    right = (upv[1], -upv[0])
                    hu_cor = right[0] * hu_ec + right[1] * hv_nc
                    hv_cor = upv[0] * hu_ec + upv[1] * hv_nc
                    st1 = X + hu_cor + b1 / DX
                    st2 = Y + hv_cor + b2 / DY
                L1 = -fdt[0]; L2 = -fdt[1] + st1; L3 = -fdt[2] + st2
                Cf = f32(0)
                if FRIC > 0:
                    if h < DESING:
                        u = _desing(h, hu, DESING); v = _desing(h, hv, DESING)
                        Cf = dt * FRIC * _desing(h, f32(math.sqrt(float(u * u + v * v))), DESING)
                    else:
                        u = hu / h; v = hv / h
                        Cf = dt * FRIC * f32(math.sqrt(float(u * u + v * v))) / h
                if rk_step == 0:
                    ue = R[0, j, i] + dt * L1
                    uu = (hu + dt * L2) / (f32(1.0) + Cf)
                    uv = (hv + dt * L3) / (f32(1.0) + Cf)
                else:
                    ea, ha, va = out[0][tj, ti], out[1][tj, ti], out[2][tj, ti]
                    ue = f32(0.5) * (ea + (R[0, j, i] + dt * L1))
                    uu = f32(0.5) * (ha + (hu + dt * L2)) / (f32(1.0) + f32(0.5) * Cf)
                    uv = f32(0.5) * (va + (hv + dt * L3)) / (f32(1.0) + f32(0.5) * Cf)
                if ue + Hmr <= CUTOFF:
                    ue = -Hmr + CUTOFF; uu = f32(0); uv = f32(0)
                out[0][tj, ti] = max(-Hmr + CUTOFF, ue)
                out[1][tj, ti] = uu
                out[2][tj, ti] = uv
    return tuple(out)
