"""NumPy restatement of gpu-ocean's global ETKF analysis
(gpu_ocean/SWESimulators/ETKFOcean.py:74-301), in float64.

TEST INFRASTRUCTURE ONLY: only ``tests/`` import this module; the product's
analysis (SWESimulators/ETKFOcean.py) never calls it.

Layout as the reference: X_f is (n_state, N_e) with columns
[eta.ravel(), hu.ravel(), hv.ravel()] of each member (ghost cells included),
HX_f is (2 N_d, N_e) with rows [hu_1..hu_D, hv_1..hv_D], y is (2 N_d,).
"""

from __future__ import annotations

import numpy as np


def construct_rinv(R2, n_drifters):
    """R for 2 N_d observations from the 2x2 per-drifter covariance, inverted
    (ETKFOcean.py:127-141)."""
    Nd = n_drifters
    R = np.zeros((2 * Nd, 2 * Nd))
    for l in range(Nd):
        R[l, l] = R2[0, 0]
        R[Nd + l, Nd + l] = R2[1, 1]
        R[l, Nd + l] = R2[0, 1]
        R[Nd + l, l] = R2[1, 0]
    return np.linalg.inv(R)


def analysis(X_f, HX_f, y, R2):
    """X_a (n_state, N_e) of the global ETKF (ETKFOcean.py:74-284)."""
    X_f = np.asarray(X_f, dtype=np.float64)
    HX_f = np.asarray(HX_f, dtype=np.float64)
    N = X_f.shape[1]
    Nd = HX_f.shape[0] // 2
    X_mean = X_f.mean(axis=1)
    X_pert = X_f - X_mean[:, None]
    HX_mean = HX_f.mean(axis=1)
    HX_pert = HX_f - HX_mean[:, None]
    Rinv = construct_rinv(np.asarray(R2, dtype=np.float64), Nd)
    D = np.asarray(y, dtype=np.float64) - HX_mean
    P = np.linalg.inv((N - 1) * np.eye(N) + HX_pert.T @ Rinv @ HX_pert)
    K = X_pert @ (P @ (HX_pert.T @ Rinv))
    Xa_mean = X_mean + K @ D
    sigma, V = np.linalg.eigh((N - 1) * P)
    Xa_pert = X_pert @ (V @ np.diag(np.sqrt(np.real(sigma))) @ V.T)
    return Xa_pert + Xa_mean[:, None]


# ---------------------------------------------------------------------------
# Localised ETKF (LETKF) of ETKFOcean.py:299-744, with static-buoy
# observations (observe_particles_from_X_f, ETKFOcean.py:777-793).
#
# The state here is the interior only: X_f has shape (N, 3, ny, nx) in float32
# like the reference's downloaded X_f (ETKFOcean.py:594, 749-757).
# ---------------------------------------------------------------------------

def gaspari_cohn(dist, c):
    """Gaspari & Cohn (1999, eq. 4.10) correlation at distance ``dist`` for
    the half-width ``c`` (ETKFOcean.distGC, ETKFOcean.py:356-392): 1 at 0,
    zero from 2c on."""
    z = np.atleast_1d(np.asarray(dist, dtype=np.float64)) / c
    out = np.zeros_like(z)
    for i, q in enumerate(z):
        if q < 1:
            out[i] = 1 - 5 / 3 * q ** 2 + 5 / 8 * q ** 3 + 1 / 2 * q ** 4 - 1 / 4 * q ** 5
        elif q < 2:
            out[i] = 4 - 5 * q + 5 / 3 * q ** 2 + 5 / 8 * q ** 3 - 1 / 2 * q ** 4 + 1 / 12 * q ** 5 - 2 / (3 * q)
    return out


def periodic_distance(a, b, lx, ly):
    """The scalar branch of distGC (ETKFOcean.py:370-374): the smallest of the
    norms of |a-b|, |a-b| - (lx, 0), |a-b| - (0, ly) and |a-b| - (lx, ly)."""
    d = np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64))
    return min(np.linalg.norm(d - s) for s in
               (np.array([0.0, 0.0]), np.array([lx, 0.0]), np.array([0.0, ly]), np.array([lx, ly])))


def local_weight_shape(r, dx, dy, nx, ny, scale_w=1.0):
    """(2R+1)^2 weights, R = ceil(1.5 r), centred on the middle cell:
    min(1, GC) inside 1.5 r dx of the centre, 0 outside, times scale_w
    (ETKFOcean.getLocalWeightShape, ETKFOcean.py:396-418)."""
    n = int(np.ceil(r * 1.5) * 2 + 1)
    centre = np.array([n * dx / 2, n * dy / 2])
    w = np.zeros((n, n))
    for j in range(n):
        for i in range(n):
            p = np.array([(i + 0.5) * dx, (j + 0.5) * dy])
            if np.linalg.norm(centre - p) <= 1.5 * r * dx:
                w[j, i] = min(1.0, gaspari_cohn(periodic_distance(centre, p, nx * dx, ny * dy), r * dx)[0])
    return scale_w * w


def local_indices(pos, r, dx, dy, nx, ny):
    """Mask of the (2R+1)^2 box around an observation and the roll that
    brings the masked cells (taken in row-major order) into box order
    (ETKFOcean.getLocalIndices, ETKFOcean.py:302-352).  The box starts R
    cells before round(pos/dx), numpy's round half to even, and wraps
    periodically on at most one side per axis."""
    R = dx * np.ceil(r * 1.5)
    lo_x = int(np.round(pos[0] / dx)) - int(np.round(R / dx))
    hi_x = int(np.round(pos[0] / dx)) + int(np.round((R + dx) / dx))
    lo_y = int(np.round(pos[1] / dy)) - int(np.round(R / dy))
    hi_y = int(np.round(pos[1] / dy)) + int(np.round((R + dy) / dy))

    def ranges(lo, hi, n):
        if lo < 0:
            return [(n + lo, n), (0, hi)], lo
        if hi > n:
            return [(0, hi - n), (lo, n)], hi - n
        return [(lo, hi)], 0

    xs, xroll = ranges(lo_x, hi_x, nx)
    ys, yroll = ranges(lo_y, hi_y, ny)
    mask = np.zeros((ny, nx), dtype=bool)
    for a, b in xs:
        for c, d in ys:
            mask[c:d, a:b] = True
    return mask, xroll, yroll


def combined_weights(positions, r, dx, dy, nx, ny, scale_w=1.0):
    """Sum over one group's observations of the local weights placed on the
    grid (ETKFOcean.getCombinedWeights, ETKFOcean.py:421-443)."""
    W = np.zeros((ny, nx))
    W_loc = local_weight_shape(r, dx, dy, nx, ny, scale_w)
    for p in positions:
        mask, xroll, yroll = local_indices(p, r, dx, dy, nx, ny)
        W[mask] += np.roll(np.roll(W_loc, yroll, axis=0), xroll, axis=1).ravel()
    return W


def observation_groups(positions, r, dx, dy, nx, ny):
    """Greedy split of the observations into groups whose members are at least
    2 * 1.5 r dx apart, periodic distances (ETKFOcean.initializeGroups,
    ETKFOcean.py:478-514): while a group holds a closer pair, the observation
    in the column of the first closest entry (row-major) moves to the next
    group."""
    P = np.asarray(positions, dtype=np.float64)
    lx, ly = dx * nx, dy * ny
    n = P.shape[0]
    dist = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            ddx = abs(P[i, 0] - P[j, 0])
            ddx = lx - ddx if ddx > lx / 2 else ddx
            ddy = abs(P[i, 1] - P[j, 1])
            ddy = ly - ddy if ddy > ly / 2 else ddy
            dist[i, j] = np.sqrt(ddx ** 2 + ddy ** 2)
    np.fill_diagonal(dist, np.sqrt(lx ** 2 + ly ** 2))
    thr = 2.0 * 1.5 * r * dx
    groups = [list(range(n))]
    g = 0
    while dist[np.ix_(groups[g], groups[g])].min() < thr:
        while dist[np.ix_(groups[g], groups[g])].min() < thr:
            sub = dist[np.ix_(groups[g], groups[g])]
            col = np.where(sub == sub.min())[1][0]
            moved = groups[g][col]
            groups[g] = [k for k in groups[g] if k != moved]
            if len(groups) < g + 2:
                groups.append([moved])
            else:
                groups[g + 1].append(moved)
        g += 1
    return groups


def observe_buoys(X_f, observations, dx, dy):
    """(N, N_d, 2) hu, hv of each member at the cell holding each observation
    (ETKFOcean.observe_particles_from_X_f, ETKFOcean.py:777-793)."""
    out = np.empty((X_f.shape[0], observations.shape[0], 2))
    for d in range(observations.shape[0]):
        ix = int(np.floor(observations[d, 0] / dx))
        iy = int(np.floor(observations[d, 1] / dy))
        out[:, d, 0] = X_f[:, 1, iy, ix]
        out[:, d, 1] = X_f[:, 2, iy, ix]
    return out


def _local_transform(Xp, xm, hp, D, Rinv, N, ff):
    """Local analysis mean + anomalies (ETKFOcean.py:681-704) for the
    forgetting factor ff (1: no inflation)."""
    A2 = hp.T @ Rinv @ hp
    A = (N - 1) * ff * np.eye(N) + A2
    K = Xp @ np.linalg.solve(A, hp.T @ Rinv)
    mean_a = xm + K @ D
    lam, V = np.linalg.eigh(A / (N - 1))
    return Xp @ V @ np.diag(np.sqrt(1 / np.real(lam))) @ V.T + mean_a[:, None]


def letkf(X_f, observations, positions, R2, r, dx, dy, scale_w=1.0, inflation=1.0):
    """Analysis (N, 3, ny, nx) float32 of ETKFOcean.LETKF (ETKFOcean.py:
    583-744).  observations: (N_d, 4) [x, y, hu, hv]; positions: (N_d, 2)
    buoy positions that place the local boxes and groups; R2: 2x2
    observation covariance; inflation: 1 none, 0 adaptive (Saetrom and Omre),
    else the factor applied to hu and hv."""
    X = np.array(X_f, dtype=np.float32)
    N, _, ny, nx = X.shape
    obs = np.asarray(observations, dtype=np.float64)
    pos = np.asarray(positions, dtype=np.float64)
    W_loc = local_weight_shape(r, dx, dy, nx, ny, scale_w)
    n_loc = W_loc.size
    side = W_loc.shape[0]
    boxes = [local_indices(p, r, dx, dy, nx, ny) for p in pos]
    Rinv = np.linalg.inv(np.asarray(R2, dtype=np.float64))
    for grp in observation_groups(pos, r, dx, dy, nx, ny):
        W = combined_weights(pos[grp], r, dx, dy, nx, ny, scale_w)
        W_forecast = 1.0 - W / np.maximum(W, 1.0)
        mean = np.mean(X, axis=0)                               # float32, :749-759
        pert = X - mean
        HX = observe_buoys(X, obs, dx, dy)                      # :763-775
        HX_mean = np.nansum(HX, axis=0) / N
        HX_pert = HX - HX_mean
        X_a = np.zeros_like(X)
        for d in grp:
            mask, xroll, yroll = boxes[d]
            loc = X[:, :, mask].astype(np.float64)
            locp = pert[:, :, mask].astype(np.float64)
            locm = mean[:, mask].astype(np.float64)
            if xroll or yroll:
                loc = np.roll(np.roll(loc.reshape(N, 3, side, side), -yroll, 2), -xroll, 3).reshape(N, 3, n_loc)
                locp = np.roll(np.roll(locp.reshape(N, 3, side, side), -yroll, 2), -xroll, 3).reshape(N, 3, n_loc)
                locm = np.roll(np.roll(locm.reshape(3, side, side), -yroll, 1), -xroll, 2).reshape(3, n_loc)
            xm = locm.reshape(-1)
            Xp = locp.reshape(N, 3 * n_loc).T
            hp = HX_pert[:, d, :].T
            D = obs[d, 2:4] - HX_mean[d]
            if inflation == 0.0:
                ff = 1.0 / (1.0 + np.trace(Rinv @ np.outer(D, D)) / (N - 2))
            else:
                ff = 1.0 / inflation ** 2
            Xa = _local_transform(Xp, xm, hp, D, Rinv, N, 1.0)
            if ff != 1.0:
                Xa[n_loc:] = _local_transform(Xp, xm, hp, D, Rinv, N, ff)[n_loc:]
            for i in range(3):
                wx = Xa[i * n_loc:(i + 1) * n_loc] * W_loc.reshape(-1)[:, None]
                if xroll or yroll:
                    wx = np.roll(np.roll(wx.reshape(side, side, N), yroll, 0), xroll, 1)
                X_a[:, i, mask] += wx.reshape(n_loc, N).T
        X = (W_forecast[None, None] * X + X_a).astype(np.float32)   # :733-740
    return X


# ---------------------------------------------------------------------------
# Stochastic EnKF (EnKFOcean.EnKF, gpu_ocean/SWESimulators/EnKFOcean.py:64-296)
# ---------------------------------------------------------------------------

def enkf(X_f, HX_f, innovations, Y_pert, R2, inflation=1.0):
    """X_a (n_state, N) of the reference's EnKF with perturbed observations:
    HX_f (2 N_d, N) rows [hu_1..hu_D, hv_1..hv_D] (ETKFOcean layout),
    innovations (2 N_d, N) y - H x of each member, Y_pert (2 N_d, N) the
    observation perturbations (multivariate normal N(0, R) in the
    reference, :186-221); HPHT = infl^2/(N-1) HX' HX'^T, F = HPHT + R,
    E = HX'^T F^-1 (D), X_a = X_f + infl^2/(N-1) X' E (:172-293)."""
    X_f = np.asarray(X_f, dtype=np.float64)
    HX_f = np.asarray(HX_f, dtype=np.float64)
    N = X_f.shape[1]
    Nd = HX_f.shape[0] // 2
    R = np.linalg.inv(construct_rinv(np.asarray(R2, dtype=np.float64), Nd))
    HXp = HX_f - np.nanmean(HX_f, axis=1)[:, None]
    F = inflation ** 2 / (N - 1) * HXp @ HXp.T + R
    D = np.asarray(innovations, dtype=np.float64) + np.asarray(Y_pert, dtype=np.float64)
    E = HXp.T @ (np.linalg.inv(F) @ D)
    Xp = X_f - X_f.mean(axis=1)[:, None]
    return X_f + inflation ** 2 / (N - 1) * Xp @ E
