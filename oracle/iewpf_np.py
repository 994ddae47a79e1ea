"""NumPy restatement of gpu-ocean's implicit equal-weights particle filter,
two-stage and one-stage (IEWPFOcean.iewpf_2stage and iewpf,
gpu_ocean/SWESimulators/IEWPFOcean.py:195-394, with the kernels of
gpu_kernels/iewpf_kernels.cu), float64.

TEST INFRASTRUCTURE ONLY: only ``tests/`` import this module; the product's
filter (SWESimulators/IEWPFOcean.py) never calls it.

The model error (SOAR + geostrophic balance) comes from the model-error
oracle (ocean_noise_np).  The filter's assumptions are the reference's:
doubly periodic domain, constant depth H, constant f, dx == dy.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.special import lambertw

from oracle import ocean_noise_np as ON


@dataclass
class Params:
    nx: int
    ny: int
    dx: float
    dy: float
    factor: int
    q0: float
    L: float
    f: float
    g: float
    H: float
    R: np.ndarray            # 2x2 observation covariance

    def __post_init__(self):
        f32 = np.float32
        self.cnx = self.nx // self.factor
        self.cny = self.ny // self.factor
        self.cdx = float(f32(f32(self.dx) * self.factor))
        self.cdy = float(f32(f32(self.dy) * self.factor))
        self.q0 = float(f32(self.q0))
        self.L = float(f32(self.L))
        # geoBalanceConst = float32(g H / (2 f)) (IEWPFOcean.py:103)
        self.geo = float(f32(f32(self.g) * f32(self.H) / (2.0 * f32(self.f))))
        self.Nx = 3 * self.nx * self.ny
        self.ratio = self.Nx / (self.cnx * self.cny)


def soar_q(p, ax, ay, bx, by):
    """SOAR covariance between coarse cells, periodic distances
    (IEWPFOcean._SOAR_Q_CPU, :1025-1036)."""
    ddx = min((ax - bx) ** 2, (ax - (bx + p.cnx)) ** 2, (ax - (bx - p.cnx)) ** 2)
    ddy = min((ay - by) ** 2, (ay - (by + p.cny)) ** 2, (ay - (by - p.cny)) ** 2)
    d = np.sqrt(p.cdx * p.cdx * ddx + p.cdy * p.cdy * ddy)
    return p.q0 * (1.0 + d / p.L) * np.exp(-d / p.L)


def bicubic_inner(c, i, j, rx, ry):
    """Bicubic value at (i + rx, j + ry) of the coarse array c from the
    values and central-difference derivatives at the four corners
    (OceanStateNoise._bicubic_interpolation_inner, OceanStateNoise.py:883-936)."""
    B = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [-3, 3, -2, -1], [2, -2, 1, 1]], dtype=np.float64)
    f = lambda jj, ii: c[jj, ii]
    dxf = lambda jj, ii: (c[jj, ii + 1] - c[jj, ii - 1]) / 2
    dyf = lambda jj, ii: (c[jj + 1, ii] - c[jj - 1, ii]) / 2
    F = np.array([[f(j, i), f(j + 1, i), dyf(j, i), dyf(j + 1, i)],
                  [f(j, i + 1), f(j + 1, i + 1), dyf(j, i + 1), dyf(j + 1, i + 1)],
                  [dxf(j, i), dxf(j + 1, i), (dyf(j, i + 1) - dyf(j, i - 1)) / 2,
                   (dyf(j + 1, i + 1) - dyf(j + 1, i - 1)) / 2],
                  [dxf(j, i + 1), dxf(j + 1, i + 1), (dyf(j, i + 2) - dyf(j, i)) / 2,
                   (dyf(j + 1, i + 2) - dyf(j + 1, i)) / 2]])
    A = B @ F @ B.T
    xv = np.array([1.0, rx, rx * rx, rx ** 3])
    yv = np.array([1.0, ry, ry * ry, ry ** 3])
    return xv @ A @ yv


def create_S(p):
    """S = (H Q H^T + R)^-1 for one drifter (IEWPFOcean._createS, :859-980):
    the observed hu, hv response to a unit innovation, spread by U_GB^T,
    twice by SOAR, read back through the (interpolated) geostrophic balance."""
    m = 3
    xc = np.zeros((7, 7))
    yc = np.zeros((7, 7))
    xc[m + 1, m] = -p.geo / p.cdy
    xc[m - 1, m] = p.geo / p.cdy
    yc[m, m + 1] = p.geo / p.cdx
    yc[m, m - 1] = -p.geo / p.cdx
    tx = np.zeros((7, 7))
    ty = np.zeros((7, 7))
    for (j, i) in ((m + 1, m), (m - 1, m)):
        for b in range(j - 2, j + 3):
            for a in range(i - 2, i + 3):
                tx[b, a] += xc[j, i] * soar_q(p, a, b, i, j)
    for (j, i) in ((m, m + 1), (m, m - 1)):
        for b in range(j - 2, j + 3):
            for a in range(i - 2, i + 3):
                ty[b, a] += yc[j, i] * soar_q(p, a, b, i, j)
    for j in range(m - 2, m + 3):
        for i in range(m - 2, m + 3):
            xc[j, i] = 0.0
            yc[j, i] = 0.0
            for b in range(max(j - 2, 0), min(j + 3, 7)):
                for a in range(max(i - 2, 0), min(i + 3, 7)):
                    q = soar_q(p, a, b, i, j)
                    xc[j, i] += tx[b, a] * q
                    yc[j, i] += ty[b, a] * q
    if p.factor > 1:
        r = 1.0 / p.factor
        north = lambda c: bicubic_inner(c, m, m, 0.0, r)
        east = lambda c: bicubic_inner(c, m, m, r, 0.0)
        south = lambda c: bicubic_inner(c, m - 1, m - 1, 1.0, 1.0 - r)
        west = lambda c: bicubic_inner(c, m - 1, m - 1, 1.0 - r, 1.0)
    else:
        north = lambda c: c[m + 1, m]
        south = lambda c: c[m - 1, m]
        west = lambda c: c[m, m - 1]
        east = lambda c: c[m, m + 1]
    x_hu = -p.geo * (north(xc) - south(xc)) / p.dy
    x_hv = p.geo * (east(xc) - west(xc)) / p.dx
    y_hu = -p.geo * (north(yc) - south(yc)) / p.dy
    y_hv = p.geo * (east(yc) - west(yc)) / p.dx
    HQHT = np.array([[x_hu, y_hu], [x_hv, y_hv]])
    return np.linalg.inv(HQHT + np.asarray(p.R, dtype=np.float64)).astype(np.float32)


def soar_matrix(p, n=7, cutoff=2):
    """Q_SOAR on an n x n coarse patch, cut off at 2 cells, periodic
    (IEWPFOcean._createCutoffSOARMatrixQ, :1039-1075)."""
    Q = np.zeros((n * n, n * n))
    for ay in range(n):
        for ax in range(n):
            for by in range(ay - cutoff, ay + cutoff + 1):
                for bx in range(ax - cutoff, ax + cutoff + 1):
                    Q[ay * n + ax, (by % n) * n + bx % n] = soar_q(p, ax, ay, bx % n, by % n)
    return Q


def ugb_matrix(p, n=7):
    """[I; A_hu; A_hv]: eta and its geostrophic transports by central
    differences on the periodic n x n coarse patch (IEWPFOcean._createUGBmatrix,
    :1078-1124)."""
    I = np.eye(n * n)
    Ahu = np.zeros((n * n, n * n))
    Ahv = np.zeros((n * n, n * n))
    for ay in range(n):
        for ax in range(n):
            k = ay * n + ax
            Ahu[k, ((ay + 1) % n) * n + ax] = 1.0
            Ahu[k, ((ay - 1) % n) * n + ax] = -1.0
            Ahv[k, ay * n + (ax + 1) % n] = 1.0
            Ahv[k, ay * n + (ax - 1) % n] = -1.0
    return np.vstack([I, Ahu * (-p.geo / p.cdy), Ahv * (p.geo / p.cdx)])


def h_matrix(n, px, py):
    """Observation of hu, hv at patch cell (px, py) (:1126-1138)."""
    H = np.zeros((2, 3 * n * n))
    H[0, n * n + py * n + px] = 1.0
    H[1, 2 * n * n + py * n + px] = 1.0
    return H


def local_svd(p, S):
    """U sqrt(Sigma) of I - Q U_GB^T H^T S H U_GB Q on the 7 x 7 patch
    (IEWPFOcean._generateLocaleSVDforP, :1140-1196)."""
    H = h_matrix(7, 3, 3)
    Q = soar_matrix(p)
    U = ugb_matrix(p)
    M = np.eye(49) - Q @ U.T @ H.T @ np.asarray(S, dtype=np.float64) @ H @ U @ Q
    u, s, _ = np.linalg.svd(M, full_matrices=True)
    return (u @ np.diag(np.sqrt(s))).astype(np.float32)


def apply_local_svd(xi, svd, cx, cy):
    """In place on the coarse field xi: the 7 x 7 window centred on coarse
    cell (cx, cy) (periodic) is replaced by svd @ window
    (localSVDOnGlobalXi, iewpf_kernels.cu:213-258)."""
    ny, nx = xi.shape
    ys = (cy - 3 + np.arange(7)) % ny
    xs = (cx - 3 + np.arange(7)) % nx
    w = np.asarray(xi[np.ix_(ys, xs)], dtype=np.float64).ravel()
    xi[np.ix_(ys, xs)] = (np.asarray(svd, dtype=np.float64) @ w).reshape(7, 7)
    return xi


def half_kalman_pattern(p, e):
    """7 x 7 coarse patch Q U_GB^T H^T e centred on the drifter's coarse cell
    (halfTheKalmanGain, iewpf_kernels.cu:115-176): the innovation spreads to
    the four geostrophic neighbours and each is spread by SOAR."""
    g = p.geo
    pts = ((3, 4, -e[0] * g / p.cdy), (3, 2, e[0] * g / p.cdy),
           (4, 3, e[1] * g / p.cdx), (2, 3, -e[1] * g / p.cdx))
    K = np.zeros((7, 7))
    for ty in range(7):
        for tx in range(7):
            for (bx, by, w) in pts:
                if abs(tx - bx) <= 2 and abs(ty - by) <= 2:
                    K[ty, tx] += w * soar_q(p, tx, ty, bx, by)
    return K


def perturb(p, random, eta, hu, hv, Hi, fine_i=None, fine_j=None):
    """eta, hu, hv plus the SOAR + geostrophic perturbation of the coarse
    random field (perturbSim, OceanStateNoise.py:376-496), aligned with fine
    cell (fine_i, fine_j) when given."""
    coarse = ON.soar_coarse(random, p.cnx, p.cny, p.cdx, p.cdy, p.q0, p.L)
    if p.factor == 1:
        return ON.geostrophic_balance(coarse, eta, hu, hv, Hi, p.nx, p.ny, p.dx, p.dy, p.g, p.f)
    oi, oj = ON.coarse_grid_offset(p.factor, fine_i, fine_j)
    return ON.bicubic_balance(coarse, eta, hu, hv, Hi, p.nx, p.ny, p.dx, p.dy, p.g, p.f,
                              factor=p.factor, offset_i=oi, offset_j=oj)


def kalman_gain(p, S, state, Hi, positions, innovation):
    """One particle's state plus the Kalman-gain term of every drifter, and
    phi = sum_d d^T S d (IEWPFOcean.addKalmanGain, :548-595)."""
    eta, hu, hv = (np.asarray(a, dtype=np.float64) for a in state)
    phi = 0.0
    for d in range(positions.shape[0]):
        dd = np.asarray(innovation[d], dtype=np.float64)
        e = np.asarray(S, dtype=np.float64) @ dd
        e32 = e.astype(np.float32).astype(np.float64)
        cx = int(np.floor(positions[d, 0] / p.cdx))
        cy = int(np.floor(positions[d, 1] / p.cdy))
        xi = np.zeros((p.cny, p.cnx))
        ys = (cy - 3 + np.arange(7)) % p.cny
        xs = (cx - 3 + np.arange(7)) % p.cnx
        xi[np.ix_(ys, xs)] += half_kalman_pattern(p, e32)
        phi += dd @ e
        fi = int(np.floor(positions[d, 0] / p.dx))
        fj = int(np.floor(positions[d, 1] / p.dy))
        eta, hu, hv = perturb(p, xi, eta, hu, hv, Hi, fi, fj)
    return (eta, hu, hv), phi


def target_two_stage(c, nu_norm):
    """w_target = mean(c), beta = min((w_target - c) / nu_norm) + 1
    (IEWPFOcean.obtainTargetWeightTwoStage, :764-777)."""
    w = np.mean(c)
    return w, np.min((w - c) / nu_norm) + 1.0


def solve_alpha(gamma, Nx, c_star):
    """alpha from the principal branch of the Lambert W
    (IEWPFOcean.solveImplicitEquation, :804-857)."""
    arg = -(gamma / Nx) * np.exp(-gamma / Nx) * np.exp(-c_star / Nx)
    return -(Nx / gamma) * np.real(lambertw(arg, k=0))


def iewpf_2stage(p, S, svd, states, Hi, positions, innovations, xi, nu, beta=None):
    """Two-stage IEWPF update of every particle (IEWPFOcean.iewpf_2stage,
    :195-329) given the perpendicular N(0, I) coarse fields xi, nu of each
    particle (generatePerpendicularNormalDistributions).  states: (N, 3,
    ny+4, nx+4); returns the updated states (interior perturbed, ghost cells
    as the perturbation leaves them) and the diagnostics."""
    N = len(states)
    w_rest = -np.log(1.0 / N) * np.ones(N)
    out, phi = [], np.zeros(N)
    for k in range(N):
        s, phi[k] = kalman_gain(p, S, states[k], Hi, positions, innovations[k])
        out.append(s)
    gamma = np.array([np.sum(np.asarray(x, np.float64) ** 2) for x in xi]) * p.ratio
    nu_norm = np.array([np.sum(np.asarray(x, np.float64) ** 2) for x in nu]) * p.ratio
    c = phi + w_rest
    w_target, b = target_two_stage(c, nu_norm)
    if b < 0:
        b = 0.0
    if beta is not None:
        b = beta
    alphas = np.zeros(N)
    res = []
    for k in range(N):
        c_star = w_target - c[k] - (b - 1) * nu_norm[k]
        alphas[k] = solve_alpha(gamma[k], p.Nx, c_star)
        z = np.sqrt(alphas[k]) * np.asarray(xi[k], np.float64) + np.sqrt(b) * np.asarray(nu[k], np.float64)
        for d in range(positions.shape[0]):
            apply_local_svd(z, svd, int(np.floor(positions[d, 0] / p.cdx)), int(np.floor(positions[d, 1] / p.cdy)))
        res.append(perturb(p, z, *out[k], Hi))
    diag = dict(phi=phi, gamma=gamma, nu_norm=nu_norm, c=c, w_target=w_target, beta=b, alpha=alphas)
    return np.array(res), diag


def iewpf_1stage(p, S, svd, states, Hi, positions, innovations, xi):
    """One-stage IEWPF update of every particle (IEWPFOcean.iewpf, :331-394)
    given each particle's N(0, I) coarse field xi: Kalman gain, target weight
    max(c) (obtainTargetWeight, :742-762), gamma = |xi|^2 scaled, alpha from
    the Lambert W, the local SVD blocks at the drifters (sampleFromP,
    :597-627) and the perturbation scaled by sqrt(alpha)."""
    N = len(states)
    w_rest = -np.log(1.0 / N) * np.ones(N)
    out, phi = [], np.zeros(N)
    for k in range(N):
        s, phi[k] = kalman_gain(p, S, states[k], Hi, positions, innovations[k])
        out.append(s)
    c = phi + w_rest
    w_target = np.max(c)
    gamma = np.array([np.sum(np.asarray(x, np.float64) ** 2) for x in xi]) * p.ratio
    alphas = np.zeros(N)
    res = []
    for k in range(N):
        alphas[k] = solve_alpha(gamma[k], p.Nx, w_target - c[k])
        z = np.array(xi[k], dtype=np.float64)
        for d in range(positions.shape[0]):
            apply_local_svd(z, svd, int(np.floor(positions[d, 0] / p.cdx)), int(np.floor(positions[d, 1] / p.cdy)))
        res.append(perturb(p, np.sqrt(alphas[k]) * z, *out[k], Hi))
    diag = dict(phi=phi, gamma=gamma, c=c, w_target=w_target, alpha=alphas)
    return np.array(res), diag
