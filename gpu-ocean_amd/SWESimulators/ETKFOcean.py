"""Global and localised ETKF analyses on the batched ensemble state
(gpu_ocean/SWESimulators/ETKFOcean.py:34-301 and 299-812).

The reference downloads every member, assembles X_f (n_state x N_e) in
float64 on the host, and uploads every member again
(ETKFOcean.py:143-191, 286-301).  Here the ensemble state stays where
CDKLM16Ensemble keeps it -- N members of three (ny+4) x (nx+4) planes back
to back, i.e. an [N, 3, ny+4, nx+4] float32 tensor in HBM:

  * H X_f: the batched observation kernel (observeUnderlyingFlow for every
    member at the true state's drifter positions, one launch);
  * the ensemble-space algebra (2 N_d observations, N x N matrices) in
    float64 on the host, as the reference does (ETKFOcean.py:232-284);
  * the state update X_a = mean + (X_f - mean) T with T = W + w 1^T, one
    library GEMM of [N x N]^T by [N x n_state] on the device, written back in
    place.  W is the symmetric square root of (N-1) P and w = P HX'^T R^-1 D,
    so that X' w = K D of the reference.

The LETKF (ETKFOcean.LETKF, ETKFOcean.py:583-744) keeps the reference's
geometry -- the Gaspari-Cohn weight stencil, the periodic local boxes, the
greedy groups of distant buoys -- computed once on the host per buoy set, and
runs each group as two device launches (``gocean_letkf_coefficients``,
``gocean_letkf_apply`` in ``lib/libgocean_hip.so``): the per-buoy local
transform in a rank-2 closed form (float64), then the weighted blend of the
local analyses into every member in place.  Nothing is downloaded.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import Common
from . import _gocean as G


class ETKFOcean(object):

    def __init__(self, ensemble, observation_cov, inflation_factor=1.0, scale_w=1.0):
        """ensemble: a CDKLM16Ensemble whose members share nx, ny (and whose
        drifters are enabled, for the observation kernel); observation_cov:
        the 2x2 covariance of one (hu, hv) observation."""
        self.ensemble = ensemble
        self.N_e = ensemble.n
        assert len(set(zip(ensemble.nx, ensemble.ny))) == 1, "the ETKF needs members of one size"
        self.n_i = ensemble.ny[0] + 4
        self.n_j = ensemble.nx[0] + 4
        self.ghost_cells_x = self.ghost_cells_y = 2
        self.plane = ensemble.planes[0]
        assert all(int(ensemble.state_off[k]) == 3 * self.plane * k for k in range(self.N_e))
        self.R = np.asarray(observation_cov, dtype=np.float64).reshape(2, 2)
        self.inflation_factor = inflation_factor
        self.scale_w = scale_w
        self.last = {}
        # localisation (ETKFOcean.py:64-70)
        self.r_factor = 15.0
        self.W_loc = None
        self.groups = None
        self._letkf = None

    def state_matrix(self):
        """[N_e, 3 (ny+4)(nx+4)] view of the ensemble's current state."""
        return self.ensemble.state[0][:self.N_e * 3 * self.plane].view(self.N_e, 3 * self.plane)

    def _constructRinv(self, n_drifters):
        Nd = n_drifters
        R = np.zeros((2 * Nd, 2 * Nd))
        i = np.arange(Nd)
        R[i, i] = self.R[0, 0]
        R[Nd + i, Nd + i] = self.R[1, 1]
        R[i, Nd + i] = self.R[0, 1]
        R[Nd + i, i] = self.R[1, 0]
        return np.linalg.inv(R)

    def ETKF(self, observations, obs_positions=None):
        """Analysis of the whole ensemble in place.

        observations: (N_d, 4) [x, y, hu, hv] of the true state
        (observeTrueState, BaseOceanStateEnsemble.py:179-190), on the host.
        obs_positions: device (N_d, 2) or (N_d+1, 2) positions to observe the
        members at; default: the x, y columns of ``observations``."""
        torch = Common._torch()
        ens = self.ensemble
        st = ens.gpu_stream
        obs = np.asarray(observations, dtype=np.float64)
        Nd = obs.shape[0]
        assert Nd == ens.num_drifters, "observations must match the ensemble's drifter count"
        N = self.N_e
        dev = ens.state[0].device
        with torch.cuda.stream(st):
            if obs_positions is None:
                obs_positions = torch.from_numpy(obs[:, :2].astype(np.float32)).to(dev)
            HX = ens.observe(obs_positions)                       # [N, Nd, 2] on the device
            HX_f = HX[:, :Nd, :].permute(2, 1, 0).reshape(2 * Nd, N).to("cpu").numpy().astype(np.float64)
            X = self.state_matrix()
            X_mean = X.mean(dim=0, keepdim=True)
            X_pert = X - X_mean
        # ensemble-space algebra (ETKFOcean.py:232-284), float64.  One
        # eigendecomposition of A = (N-1) I + HX'^T R^-1 HX' gives both
        # P = A^-1 and the symmetric square root of (N-1) P; R^-1 is built
        # from the inverse of its 2 x 2 (hu_l, hv_l) blocks.
        HX_mean = HX_f.mean(axis=1)
        HX_pert = HX_f - HX_mean[:, None]
        r = np.linalg.inv(self.R)
        U, Vv = HX_pert[:Nd], HX_pert[Nd:]
        y = np.concatenate([obs[:, 2], obs[:, 3]])
        D = y - HX_mean
        A = (N - 1) * np.eye(N) + r[0, 0] * (U.T @ U) + r[0, 1] * (U.T @ Vv) + r[1, 0] * (Vv.T @ U) \
            + r[1, 1] * (Vv.T @ Vv)
        lam, V = np.linalg.eigh(A)
        P = (V / lam) @ V.T
        RD = np.concatenate([r[0, 0] * D[:Nd] + r[0, 1] * D[Nd:], r[1, 0] * D[:Nd] + r[1, 1] * D[Nd:]])
        w = P @ (HX_pert.T @ RD)
        W = (V * np.sqrt((N - 1) / lam)) @ V.T
        T = W + w[:, None]
        self.last = {"P": P, "w": w, "W": W, "D": D}
        with torch.cuda.stream(st):
            Tt = torch.from_numpy(np.ascontiguousarray(T.T).astype(np.float32)).to(dev)
            X_a = torch.addmm(X_mean.expand(N, -1), Tt, X_pert)
            X.copy_(X_a)
        return self.last

    # ------------------------------------------------------------------
    # LETKF (ETKFOcean.py:299-744)
    # ------------------------------------------------------------------
    def _letkf_setup(self, positions):
        """Weights, boxes, groups and per-group tile lists for one buoy set
        (ETKFOcean._prepare_LETKF, ETKFOcean.py:543-580), cached."""
        torch = Common._torch()
        ens = self.ensemble
        nx, ny = ens.nx[0], ens.ny[0]
        dx, dy = ens.dxy[0]
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 2)
        key = (pos.tobytes(), float(self.r_factor), float(self.scale_w), nx, ny, dx, dy)
        if self._letkf is not None and self._letkf["key"] == key:
            return self._letkf
        assert np.all((pos >= 0) & (pos < np.array([nx * dx, ny * dy]))), "buoys outside the domain"
        r = float(self.r_factor)
        W_loc = local_weight_shape(r, dx, dy, nx, ny, self.scale_w)
        n = W_loc.shape[0]
        R = n // 2
        assert n <= min(nx, ny), f"local box {n} wider than the {nx} x {ny} domain"
        # the reference's box is R cells high only when dx*R/dy rounds to R
        # (ETKFOcean.py:308, 317-318)
        assert int(np.round(dx * R / dy)) == R and int(np.round((dx * R + dy) / dy)) == R + 1, \
            "dx, dy give a non-square local box"
        bx = (np.round(pos[:, 0] / dx).astype(np.int64) - R) % nx
        by = (np.round(pos[:, 1] / dy).astype(np.int64) - R) % ny
        groups = observation_groups(pos, r, dx, dy, nx, ny)
        tw, th, maxc, maxn = (ctypes.c_int32() for _ in range(4))
        G.check(ens.lib.gocean_letkf_tile_shape(ctypes.byref(tw), ctypes.byref(th), ctypes.byref(maxc),
                                                ctypes.byref(maxn)), "gocean_letkf_tile_shape")
        assert self.N_e <= maxn.value, f"the LETKF takes at most {maxn.value} members"
        jj, ii = np.nonzero(W_loc)
        dev = ens.state[0].device
        tables = []
        for g in groups:
            tiles, tile_obs, cover = letkf_tiles(g, bx, by, jj, ii, nx, ny, tw.value, th.value)
            if cover > maxc.value:
                raise ValueError(f"{cover} buoys of one group weigh on one cell; at most {maxc.value}")
            tables.append((g, tiles, tile_obs))
        N, Nd = self.N_e, pos.shape[0]
        # device tables belong to the ensemble's stream, which runs the kernels
        with torch.cuda.stream(ens.gpu_stream):
            per_group = [dict(group=torch.from_numpy(np.asarray(g, dtype=np.int32)).to(dev), n=len(g),
                              tiles=torch.from_numpy(tiles).to(dev), n_tiles=tiles.shape[0],
                              tile_obs=torch.from_numpy(tile_obs).to(dev), n_tile_obs=tile_obs.size)
                         for g, tiles, tile_obs in tables]
            S = dict(key=key, W_loc=W_loc, groups=groups, per_group=per_group, local_n=n,
                     w_loc=torch.from_numpy(W_loc.astype(np.float64)).to(dev),
                     boxes=torch.from_numpy(np.column_stack([bx, by]).astype(np.int32)).to(dev),
                     coef=torch.empty(Nd * N * 6, dtype=torch.float32, device=dev),
                     oc=torch.empty(Nd * G.LETKF_COEF_BYTES, dtype=torch.uint8, device=dev),
                     layout=G.LetkfLayout(N, nx, ny, self.n_j, self.plane, 3 * self.plane))
        self.W_loc, self.groups = W_loc, groups
        self._letkf = S
        return S

    def LETKF(self, observations, r_factor=0.0, positions=None):
        """Localised analysis of the whole ensemble in place
        (ETKFOcean.LETKF, ETKFOcean.py:583-744, static buoys).

        observations: (N_d, 4) [x, y, hu, hv] of the true state; the buoy
        cell floor(x/dx), floor(y/dy) is observed (ETKFOcean.py:777-793).
        r_factor: localisation scale in cells (> 0 replaces the current one,
        15 by default, ETKFOcean.py:65, 557-561).  positions: (N_d, 2) buoy
        positions that place the local boxes and groups; default the x, y
        columns of ``observations`` (observeTrueDrifters and observeTrueState
        agree for static buoys).  inflation_factor and scale_w are the
        constructor's (1: none; 0: adaptive, ETKFOcean.py:664-671)."""
        torch = Common._torch()
        ens = self.ensemble
        obs = np.asarray(observations, dtype=np.float64).reshape(-1, 4)
        Nd = obs.shape[0]
        if r_factor > 0.0:
            self.r_factor = float(r_factor)
        S = self._letkf_setup(obs[:, :2] if positions is None else positions)
        assert S["boxes"].shape[0] == Nd, "positions and observations differ in length"
        nx, ny = ens.nx[0], ens.ny[0]
        dx, dy = ens.dxy[0]
        cells = np.column_stack([np.floor(obs[:, 0] / dx), np.floor(obs[:, 1] / dy)]).astype(np.int64)
        assert np.all((cells >= 0) & (cells < np.array([nx, ny]))), "observations outside the domain"
        st = ens.gpu_stream
        dev = ens.state[0].device
        cov = (ctypes.c_double * 4)(*self.R.ravel())
        lay = ctypes.byref(S["layout"])
        state = ens.state[0]
        with torch.cuda.stream(st):
            d_cells = torch.from_numpy(cells.astype(np.int32)).to(dev)
            d_y = torch.from_numpy(np.ascontiguousarray(obs[:, 2:4])).to(dev)
            for P in S["per_group"]:
                G.check(ens.lib.gocean_letkf_coefficients(
                    lay, state.data_ptr(), P["group"].data_ptr(), P["n"], d_cells.data_ptr(), d_y.data_ptr(), Nd,
                    cov, float(self.inflation_factor), S["coef"].data_ptr(), S["oc"].data_ptr(),
                    G.stream_handle(st)), "gocean_letkf_coefficients")
                G.check(ens.lib.gocean_letkf_apply(
                    lay, state.data_ptr(), S["w_loc"].data_ptr(), S["local_n"], S["boxes"].data_ptr(), Nd,
                    S["coef"].data_ptr(), S["oc"].data_ptr(), P["tiles"].data_ptr(), P["n_tiles"],
                    P["tile_obs"].data_ptr(), P["n_tile_obs"], G.stream_handle(st)), "gocean_letkf_apply")
            # uploadAnalysis refreshes the ghost cells (ETKFOcean.py:796-812)
            G.check(ens.lib.gocean_ensemble_periodic_bc(ens.handle, state.data_ptr(), G.stream_handle(st)),
                    "ensemble periodic bc")
        self.last = {"groups": S["groups"]}
        return self.last


# ----------------------------------------------------------------------
# LETKF geometry (host, once per buoy set)
# ----------------------------------------------------------------------
def gaspari_cohn(z):
    """Gaspari & Cohn (1999) eq. 4.10 at z = distance / half-width
    (ETKFOcean.distGC, ETKFOcean.py:356-392); zero from z = 2 on."""
    z = np.asarray(z, dtype=np.float64)
    inner = 1 - 5 / 3 * z ** 2 + 5 / 8 * z ** 3 + 1 / 2 * z ** 4 - 1 / 4 * z ** 5
    with np.errstate(divide="ignore"):
        outer = 4 - 5 * z + 5 / 3 * z ** 2 + 5 / 8 * z ** 3 - 1 / 2 * z ** 4 + 1 / 12 * z ** 5 - 2 / (3 * z)
    return np.where(z < 1, inner, np.where(z < 2, outer, 0.0))


def local_weight_shape(r, dx, dy, nx, ny, scale_w=1.0):
    """(2R+1)^2 stencil, R = ceil(1.5 r): min(1, GC(d / (r dx))) within
    1.5 r dx of the centre cell, the distance taken periodically over the
    nx dx x ny dy domain, times scale_w (ETKFOcean.getLocalWeightShape,
    ETKFOcean.py:396-418)."""
    n = int(np.ceil(r * 1.5) * 2 + 1)
    j, i = np.mgrid[0:n, 0:n]
    ddx = np.abs(n * dx / 2 - (i + 0.5) * dx)
    ddy = np.abs(n * dy / 2 - (j + 0.5) * dy)
    lx, ly = nx * dx, ny * dy
    dist = np.minimum.reduce([np.sqrt(ddx ** 2 + ddy ** 2), np.sqrt((ddx - lx) ** 2 + ddy ** 2),
                              np.sqrt(ddx ** 2 + (ddy - ly) ** 2), np.sqrt((ddx - lx) ** 2 + (ddy - ly) ** 2)])
    inside = np.sqrt(ddx ** 2 + ddy ** 2) <= 1.5 * r * dx
    return scale_w * np.where(inside, np.minimum(1.0, gaspari_cohn(dist / (r * dx))), 0.0)


def observation_groups(pos, r, dx, dy, nx, ny):
    """Buoys split so that each group's members are at least 3 r dx apart
    (periodic distance), by the reference's greedy rule: while a group has a
    closer pair, the buoy in the column of the first closest entry (row-major)
    moves to the next group (ETKFOcean.initializeGroups, ETKFOcean.py:478-514)."""
    pos = np.asarray(pos, dtype=np.float64)
    lx, ly = dx * nx, dy * ny
    d = np.abs(pos[:, None, :] - pos[None, :, :])
    d[..., 0] = np.where(d[..., 0] > lx / 2, lx - d[..., 0], d[..., 0])
    d[..., 1] = np.where(d[..., 1] > ly / 2, ly - d[..., 1], d[..., 1])
    dist = np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2)
    np.fill_diagonal(dist, np.sqrt(lx ** 2 + ly ** 2))
    thr = 3.0 * r * dx
    assert np.sqrt(lx ** 2 + ly ** 2) >= thr, "localisation radius larger than the domain"
    groups = [list(range(pos.shape[0]))]
    g = 0
    while g < len(groups):
        while len(groups[g]) > 1:
            idx = np.asarray(groups[g])
            sub = dist[np.ix_(idx, idx)]
            flat = int(np.argmin(sub))          # first minimum, row-major
            if sub.flat[flat] >= thr:
                break
            moved = groups[g].pop(flat % len(idx))
            if len(groups) == g + 1:
                groups.append([])
            groups[g + 1].append(moved)
        g += 1
    return groups


def letkf_tiles(group, bx, by, jj, ii, nx, ny, tw, th):
    """Tiles of tw x th cells that the nonzero weights (rows jj, columns ii of
    the stencil) of the group's boxes reach, each with its buoys in group
    order; and the most buoys weighing on one cell."""
    # tiles start two cells left of a multiple of tw: with the two ghost
    # columns in front, each tile row is then tw-float aligned in memory
    sh = 2
    ntx = (nx + sh + tw - 1) // tw
    tids, oids = [], []
    cover = np.zeros(nx * ny, dtype=np.int32)
    for o in group:
        xs = (bx[o] + ii) % nx
        ys = (by[o] + jj) % ny
        cover[ys * nx + xs] += 1
        t = np.unique((ys // th) * ntx + (xs + sh) // tw)
        tids.append(t)
        oids.append(np.full(t.size, o))
    tid = np.concatenate(tids) if tids else np.zeros(0, np.int64)
    oid = np.concatenate(oids) if oids else np.zeros(0, np.int64)
    order = np.argsort(tid, kind="stable")
    tid, oid = tid[order], oid[order]
    uniq, begin, count = np.unique(tid, return_index=True, return_counts=True)
    tiles = np.column_stack([(uniq % ntx) * tw - sh, (uniq // ntx) * th, begin, count]).astype(np.int32)
    return np.ascontiguousarray(tiles), oid.astype(np.int32), int(cover.max()) if cover.size else 0
