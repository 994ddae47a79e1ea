"""Two-stage implicit equal-weights particle filter on the batched ensemble
state (IEWPFOcean.iewpf_2stage, gpu_ocean/SWESimulators/IEWPFOcean.py:195-329).

The reference loops over particles from Python: per particle and drifter a
zeroed random buffer, a 7 x 7 Kalman-gain patch and a full-domain model-error
perturbation; then perpendicular samples, a 49 x 49 block at each drifter and
one more perturbation.  Here every particle of a CDKLM16Ensemble (with its
batched model error) advances together:

  * Kalman gain.  The gain term of one drifter is the model-error perturbation
    (SOAR + geostrophic balance, aligned with the drifter's cell) of the 7 x 7
    patch Q U_GB^T H^T e, e = S d.  With interpolated model error the
    balance reads H + d_eta, so the term is a fixed quadratic function of e
    (d_eta is linear, the transports carry (H + d_eta) grad d_eta): five basis
    fields per drifter, computed once per drifter set by the model-error
    kernels and kept on the window of cells where they are non-zero; one
    gather launch adds every drifter's e_x, e_y, e_x^2, e_y^2, e_x e_y
    combination of them to every particle (gocean_iewpf_kalman_window).  Without
    interpolation the balance reads H + eta + d_eta, so the term depends on
    the state: the drifters go one after the other as in the reference, each
    as one patch launch and one batched perturbation over all particles.
  * Perpendicular samples, the per-particle alpha^1/2 xi + beta^1/2 nu and
    the local SVD blocks run as batched coarse-grid kernels
    (``gocean_iewpf_*``), and the final perturbation is the batched
    model-error launch with the random field kept.
  * The scalar work -- phi, the target weight, beta and each particle's
    alpha from the Lambert W function -- stays on the host in float64, as
    in the reference.

Assumptions are the reference's (IEWPFOcean.py:94-118): doubly periodic
members of one size, constant depth, constant f, dx == dy.
"""

from __future__ import annotations

import ctypes

import numpy as np
from scipy.special import lambertw

from . import Common
from . import OceanStateNoise
from . import _gocean as G


SEGMENT = 256    # interior cells per row tile of gocean_iewpf_kalman_window


def circular_cover(mask):
    """(start, length) of the shortest periodic interval of indices holding
    every True of ``mask``; (0, 0) when there is none."""
    mask = np.asarray(mask, dtype=bool)
    n = mask.size
    idx = np.flatnonzero(mask)
    if idx.size == 0:
        return 0, 0
    gaps = np.diff(np.concatenate([idx, [idx[0] + n]]))
    k = int(np.argmax(gaps))
    return int(idx[(k + 1) % idx.size]), int(n - (gaps[k] - 1))


def kalman_window_coefficients(coef):
    """(N, N_d, 5) per-member coefficients -> the (N_d, 5, N rounded up to
    16) float32 layout gocean_iewpf_kalman_window reads (zero padding)."""
    coef = np.asarray(coef)
    n = coef.shape[0]
    out = np.zeros((coef.shape[1], 5, -(-n // 16) * 16), dtype=np.float32)
    out[:, :, :n] = coef.transpose(1, 2, 0)
    return out


def kalman_window_tiles(origins, sizes, nx, ny):
    """The 256-cell interior row segments that drifters' periodic windows meet
    (window d: sizes[d] = (rows, columns) from origins[d] = (row, column)),
    each with the drifters meeting it in ascending order.  Returns tiles
    (T, 2) [row, segment], list_off (T + 1,) and the drifter list."""
    nseg = -(-nx // SEGMENT)
    tid, dd = [], []
    for d, ((r0, c0), (h, w)) in enumerate(zip(origins, sizes)):
        if h <= 0 or w <= 0:
            continue
        rows = (int(r0) + np.arange(h)) % ny
        segs = np.unique(((int(c0) + np.arange(w)) % nx) // SEGMENT)
        t = (rows[:, None] * nseg + segs[None, :]).ravel()
        tid.append(t)
        dd.append(np.full(t.size, d))
    if not tid:
        return np.zeros((0, 2), np.int32), np.zeros(1, np.int32), np.zeros(0, np.int32)
    tid, dd = np.concatenate(tid), np.concatenate(dd)
    order = np.lexsort((dd, tid))
    tid, dd = tid[order], dd[order]
    ut, first = np.unique(tid, return_index=True)
    tiles = np.column_stack([ut // nseg, ut % nseg]).astype(np.int32)
    list_off = np.append(first, tid.size).astype(np.int32)
    return tiles, list_off, dd.astype(np.int32)


class IEWPFOcean(object):

    def __init__(self, ensemble, observation_cov, beta=None):
        """ensemble: a CDKLM16Ensemble with enable_model_error(...) and
        enable_drifters(N_d) done; observation_cov: the 2 x 2 covariance R of
        one (hu, hv) observation; beta: fixed beta (None: the two-stage
        rule)."""
        torch = Common._torch()
        ens = self.ensemble = ensemble
        if ens.noise is None:
            raise RuntimeError("the IEWPF needs the ensemble's model error (enable_model_error)")
        assert len(set(zip(ens.nx, ens.ny))) == 1, "the IEWPF needs members of one size"
        assert all(b == (Common.PERIODIC,) * 4 for b in ens.bcs), \
            "IEWPF requires periodic boundary conditions in north-south and east-west"
        assert len(set(ens.f)) == 1, "IEWPF requires constant Coriolis forcing"
        self.nx, self.ny = ens.nx[0], ens.ny[0]
        self.dx, self.dy = (np.float32(v) for v in ens.dxy[0])
        assert self.dx == self.dy, "IEWPF requires square grid cells"
        nz = ens.noise_members[0].noise
        self.interpolation_factor = int(nz.interpolation_factor)
        self.coarse_nx = self.nx // self.interpolation_factor
        self.coarse_ny = self.ny // self.interpolation_factor
        self.coarse_dx = np.float32(self.dx * self.interpolation_factor)
        self.coarse_dy = np.float32(self.dy * self.interpolation_factor)
        self.soar_q0 = np.float32(nz.soar_q0)
        self.soar_L = np.float32(nz.soar_L)
        self.f = np.float32(ens.f[0])
        self.g = np.float32(ens.g[0])
        st = ens.gpu_stream
        st.synchronize()
        Hm = ens.Hm[int(ens.hm_off[0]):int(ens.hm_off[1])].to("cpu").numpy().reshape(self.ny + 4, self.nx + 4)
        H = Hm[2:-2, 2:-2]
        assert H.max() == H.min(), "IEWPF can not be used with a non-constant ocean depth"
        self.const_H = np.float32(H[0, 0])
        self.geoBalanceConst = np.float32(self.g * self.const_H / (2.0 * self.f))
        self.Nx = 3 * self.nx * self.ny
        self.random_numbers_ratio = self.Nx / (self.coarse_nx * self.coarse_ny)
        self.numParticles = ens.n
        self.R = np.asarray(observation_cov, dtype=np.float64).reshape(2, 2)
        self.beta = beta
        self.betas = np.zeros(0)
        self.S_host = self._createS()
        self.localSVD_host = self._generateLocalSVDforP()
        self.lib = ens.lib
        dev = ens.state[0].device
        self.n_coarse = self.coarse_nx * self.coarse_ny
        with torch.cuda.stream(st):
            self.offsets = torch.from_numpy(np.ascontiguousarray(ens.noise_offsets[:, 1])).to(dev)
            self.nu = torch.zeros_like(ens.noise_random)
            self.kalman_field = torch.zeros_like(ens.noise_random)   # Kalman-gain patches, factor 1
            self.dots = torch.zeros(3 * ens.n, dtype=torch.float64, device=dev)
            self.localSVD_device = torch.from_numpy(self.localSVD_host).to(dev)
        self._basis = None
        self._patch_device = None
        self.last = {}

    # ------------------------------------------------------------------
    # constant matrices (host, float64)
    # ------------------------------------------------------------------
    def _soar(self, ax, ay, bx, by):
        """SOAR covariance between coarse cells with periodic distances
        (IEWPFOcean._SOAR_Q_CPU, IEWPFOcean.py:1025-1036); arrays allowed."""
        n_x, n_y = self.coarse_nx, self.coarse_ny
        ddx = np.minimum.reduce([(ax - bx) ** 2, (ax - bx - n_x) ** 2, (ax - bx + n_x) ** 2])
        ddy = np.minimum.reduce([(ay - by) ** 2, (ay - by - n_y) ** 2, (ay - by + n_y) ** 2])
        d = np.sqrt(float(self.coarse_dx) ** 2 * ddx + float(self.coarse_dy) ** 2 * ddy)
        L = float(self.soar_L)
        return float(self.soar_q0) * (1.0 + d / L) * np.exp(-d / L)

    def _patch_matrices(self, n=7, periodic=True):
        """Q (cut-off SOAR), U_GB = [I; A_hu; A_hv] and H on the n x n coarse
        patch observed at its centre (IEWPFOcean.py:1039-1138).  Q wraps
        around the patch when periodic, else drops sources outside it."""
        cut = 2
        j, i = np.divmod(np.arange(n * n), n)
        Q = np.zeros((n * n, n * n))
        for oy in range(-cut, cut + 1):
            for ox in range(-cut, cut + 1):
                bj, bi = j + oy, i + ox
                inside = (bj >= 0) & (bj < n) & (bi >= 0) & (bi < n)
                if periodic:
                    inside[:] = True
                bj, bi = bj % n, bi % n
                rows = np.arange(n * n)[inside]
                Q[rows, (bj * n + bi)[inside]] = self._soar(i, j, bi, bj)[inside]
        Ahu = np.zeros((n * n, n * n))
        Ahv = np.zeros((n * n, n * n))
        r = np.arange(n * n)
        Ahu[r, ((j + 1) % n) * n + i] = 1.0
        Ahu[r, ((j - 1) % n) * n + i] = -1.0
        Ahv[r, j * n + (i + 1) % n] = 1.0
        Ahv[r, j * n + (i - 1) % n] = -1.0
        geo = float(self.geoBalanceConst)
        U = np.vstack([np.eye(n * n), Ahu * (-geo / float(self.coarse_dy)), Ahv * (geo / float(self.coarse_dx))])
        H = np.zeros((2, 3 * n * n))
        c = (n // 2) * n + n // 2
        H[0, n * n + c] = 1.0
        H[1, 2 * n * n + c] = 1.0
        return Q, U, H

    @staticmethod
    def _catmull_rom(c, i, j, rx, ry):
        """Bicubic value at (i + rx, j + ry) with central-difference slopes,
        as a tensor product of cubic Hermite curves
        (OceanStateNoise._bicubic_interpolation_inner, OceanStateNoise.py:883-936)."""
        def hermite(p0, p1, m0, m1, t):
            t2, t3 = t * t, t * t * t
            return (2 * t3 - 3 * t2 + 1) * p0 + (t3 - 2 * t2 + t) * m0 + (-2 * t3 + 3 * t2) * p1 + (t3 - t2) * m1

        def along_x(jj):
            row = c[jj]
            return hermite(row[i], row[i + 1], (row[i + 1] - row[i - 1]) / 2, (row[i + 2] - row[i]) / 2, rx)

        v = [along_x(jj) for jj in (j - 1, j, j + 1, j + 2)]
        return hermite(v[1], v[2], (v[2] - v[0]) / 2, (v[3] - v[1]) / 2, ry)

    def _createS(self):
        """S = (H Q H^T + R)^-1 of one drifter (IEWPFOcean._createS,
        IEWPFOcean.py:859-980): the 7 x 7 patch response Q Q U_GB^T H^T e_k,
        kept on the 5 x 5 centre, read back through the geostrophic balance
        at the drifter's fine cell (on the bicubic surface when the model
        error is interpolated)."""
        # both SOAR passes stay inside the 7 x 7 patch (IEWPFOcean.py:884-914)
        Q, U, H = self._patch_matrices(periodic=False)
        resp = Q @ Q @ U.T @ H.T                        # (49, 2)
        m = 3
        keep = np.zeros((7, 7), dtype=bool)
        keep[1:6, 1:6] = True
        geo = float(self.geoBalanceConst)
        HQHT = np.zeros((2, 2))
        for k in range(2):
            c = np.where(keep, resp[:, k].reshape(7, 7), 0.0)
            if self.interpolation_factor > 1:
                r = 1.0 / self.interpolation_factor
                north = self._catmull_rom(c, m, m, 0.0, r)
                east = self._catmull_rom(c, m, m, r, 0.0)
                south = self._catmull_rom(c, m - 1, m - 1, 1.0, 1.0 - r)
                west = self._catmull_rom(c, m - 1, m - 1, 1.0 - r, 1.0)
            else:
                north, south, east, west = c[m + 1, m], c[m - 1, m], c[m, m + 1], c[m, m - 1]
            HQHT[0, k] = -geo * (north - south) / float(self.dy)
            HQHT[1, k] = geo * (east - west) / float(self.dx)
        return np.linalg.inv(HQHT + self.R).astype(np.float32)

    def _generateLocalSVDforP(self):
        """U sqrt(Sigma) of I - Q U_GB^T H^T S H U_GB Q on the 7 x 7 patch
        (IEWPFOcean._generateLocaleSVDforP, IEWPFOcean.py:1140-1196)."""
        Q, U, H = self._patch_matrices()
        M = np.eye(49) - Q @ U.T @ H.T @ self.S_host.astype(np.float64) @ H @ U @ Q
        u, s, _ = np.linalg.svd(M, full_matrices=True)
        return np.ascontiguousarray((u @ np.diag(np.sqrt(s))).astype(np.float32))

    def _kalman_patch(self):
        """The 7 x 7 patches Q U_GB^T H^T e for e = (1, 0) and (0, 1)
        (halfTheKalmanGain, iewpf_kernels.cu:115-176)."""
        Q, U, H = self._patch_matrices()
        return np.ascontiguousarray((Q @ U.T @ H.T).T.reshape(2, 7, 7))

    # ------------------------------------------------------------------
    # Kalman-gain basis (device, once per drifter set)
    # ------------------------------------------------------------------
    def _kalman_basis(self, positions):
        """The Kalman-gain term of every drifter as a function of e = S d
        (addKalmanGain, IEWPFOcean.py:548-595): the model-error perturbation of the patch e_x P_x + e_y P_y,
        aligned with the drifter's fine cell, is
            e_x L_x + e_y L_y + e_x^2 Q_xx + e_y^2 Q_yy + e_x e_y Q_xy
        exactly -- d_eta is linear in e and the balanced transports
        -(g/f)(H + d_eta) grad d_eta are quadratic.  The five fields come
        from perturbations at e = (+-s, 0), (0, +-s), (s, s), s of the order
        of the innovations' S d, so the quadratic parts are well resolved.

        Each drifter's fields are kept on the smallest periodic window of
        interior cells outside which all five are zero (ghost cells are left
        out: the analysis ends with the perturbation, which refreshes them).
        Returns a dict: basis (N_d, 5, 3, wh, ww) device floats, origins
        (N_d, 2) rows/columns, the window shape and the 256-cell row tiles
        with their drifter lists (kalman_window_tiles)."""
        torch = Common._torch()
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 2)
        key = pos.tobytes()
        if self._basis is not None and self._basis[0] == key:
            return self._basis[1]
        ens = self.ensemble
        st = ens.gpu_stream
        dev = ens.state[0].device
        nx, ny = self.nx, self.ny
        noise = OceanStateNoise.OceanStateNoise(
            ens.gpu_ctx, st, nx, ny, float(self.dx), float(self.dy), Common.BoundaryConditions(2, 2, 2, 2), False,
            soar_q0=float(self.soar_q0), soar_L=float(self.soar_L), interpolation_factor=self.interpolation_factor)
        assert (int(noise.rand_nx), int(noise.rand_ny)) == (self.coarse_nx, self.coarse_ny)
        patch = self._kalman_patch()
        plane = (ny + 4) * (nx + 4)
        Hi = ens.Hi[int(ens.hi_off[0]):int(ens.hi_off[1])].view(ny + 5, nx + 5)
        sc = float(np.abs(self.S_host).max() * np.sqrt(np.abs(np.diag(self.R)).max()))
        sc = sc if np.isfinite(sc) and sc > 0 else 1.0
        with torch.cuda.stream(st):
            tmp = torch.zeros((5, 3 * plane), dtype=torch.float32, device=dev)
        windows, origins, sizes = [], np.zeros((len(pos), 2), dtype=np.int32), np.zeros((len(pos), 2), dtype=np.int64)
        for d, (px, py) in enumerate(pos):
            cx = int(np.floor(px / float(self.coarse_dx)))
            cy = int(np.floor(py / float(self.coarse_dy)))
            fi = int(np.floor(px / float(self.dx)))
            fj = int(np.floor(py / float(self.dy)))
            ys = (cy - 3 + np.arange(7)) % self.coarse_ny
            xs = (cx - 3 + np.arange(7)) % self.coarse_nx
            with torch.cuda.stream(st):
                tmp.zero_()
            for k, (ex, ey) in enumerate(((sc, 0.0), (-sc, 0.0), (0.0, sc), (0.0, -sc), (sc, sc))):
                xi = np.zeros((self.coarse_ny, self.coarse_nx), dtype=np.float32)
                xi[np.ix_(ys, xs)] = (ex * patch[0] + ey * patch[1]).astype(np.float32)
                noise.random_numbers.upload(st, xi)
                row = tmp[k].view(3, ny + 4, nx + 4)
                noise.perturbOceanState(row[0], row[1], row[2], Hi, float(self.f), g=float(self.g),
                                        update_random_field=False, align_with_cell_i=fi, align_with_cell_j=fj)
            with torch.cuda.stream(st):
                t = tmp.double()
                Lx = (t[0] - t[1]) / (2 * sc)
                Qxx = (t[0] + t[1]) / (2 * sc * sc)
                Ly = (t[2] - t[3]) / (2 * sc)
                Qyy = (t[2] + t[3]) / (2 * sc * sc)
                Qxy = (t[4] - sc * (Lx + Ly)) / (sc * sc) - Qxx - Qyy
                Bd = torch.stack([Lx, Ly, Qxx, Qyy, Qxy]).float().view(5, 3, ny + 4, nx + 4)[:, :, 2:-2, 2:-2]
                nz = Bd != 0
                rows = nz.any(dim=3).any(dim=1).any(dim=0).to("cpu").numpy()
                cols = nz.any(dim=2).any(dim=1).any(dim=0).to("cpu").numpy()
                (r0, wh), (c0, ww) = circular_cover(rows), circular_cover(cols)
                origins[d], sizes[d] = (r0, c0), (wh, ww)
                ri = torch.from_numpy((r0 + np.arange(wh)) % ny).to(dev)
                ci = torch.from_numpy((c0 + np.arange(ww)) % nx).to(dev)
                windows.append(Bd.index_select(2, ri).index_select(3, ci).contiguous())
        wh, ww = max(1, int(sizes[:, 0].max())), max(1, int(sizes[:, 1].max()))
        tiles, list_off, lst = kalman_window_tiles(origins, sizes, nx, ny)
        with torch.cuda.stream(st):
            W = torch.zeros((len(pos), 5, 3, wh, ww), dtype=torch.float32, device=dev)
            for d, w in enumerate(windows):
                W[d, :, :, :w.shape[2], :w.shape[3]] = w
            basis = dict(basis=W, wh=wh, ww=ww, n_tiles=len(tiles),
                         origins=torch.from_numpy(np.ascontiguousarray(origins)).to(dev),
                         tiles=torch.from_numpy(np.ascontiguousarray(tiles, dtype=np.int32)).to(dev),
                         list_off=torch.from_numpy(np.ascontiguousarray(list_off, dtype=np.int32)).to(dev),
                         list=torch.from_numpy(np.ascontiguousarray(lst, dtype=np.int32)).to(dev))
        st.synchronize()
        noise.cleanUp()
        self._basis = (key, basis)
        return basis

    # ------------------------------------------------------------------
    # the filter
    # ------------------------------------------------------------------
    def _coarse_cells(self, positions):
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 2)
        c = np.column_stack([np.floor(pos[:, 0] / float(self.coarse_dx)),
                             np.floor(pos[:, 1] / float(self.coarse_dy))]).astype(np.int64)
        assert np.all((c >= 0) & (c < np.array([self.coarse_nx, self.coarse_ny]))), "drifters outside the domain"
        return c.astype(np.int32)

    def samplePerpendicular(self):
        """xi, nu ~ N(0, I) on every particle's coarse grid with nu
        perpendicular to xi (generatePerpendicularNormalDistributions,
        OceanStateNoise.py:556-567); returns gamma = |xi|^2 and |nu|^2 scaled
        by N_x / (coarse cells) (IEWPFOcean.samplePerpendicular, :668-691)."""
        ens = self.ensemble
        s = G.stream_handle(ens.gpu_stream)
        for buf in (ens.noise_random, self.nu):
            G.check(self.lib.gocean_noise_batch_generate(ens.noise, G.ptr(ens.noise_seed), buf.data_ptr(),
                                                         ens.noise_key, ens.noise_counter.data_ptr(), s),
                    "gocean_noise_batch_generate")
        self._dots()
        G.check(self.lib.gocean_iewpf_make_perpendicular(ens.noise_random.data_ptr(), self.nu.data_ptr(),
                                                         self.offsets.data_ptr(), ens.n, self.n_coarse,
                                                         self.dots.data_ptr(), s), "gocean_iewpf_make_perpendicular")
        return self._norms()

    def _dots(self):
        ens = self.ensemble
        G.check(self.lib.gocean_iewpf_dots(ens.noise_random.data_ptr(), self.nu.data_ptr(), self.offsets.data_ptr(),
                                           ens.n, self.n_coarse, self.dots.data_ptr(),
                                           G.stream_handle(ens.gpu_stream)), "gocean_iewpf_dots")

    def _norms(self):
        """gamma, |nu|^2 (scaled) of the current xi, nu."""
        torch = Common._torch()
        with torch.cuda.stream(self.ensemble.gpu_stream):
            d = self.dots.view(-1, 3).to("cpu").numpy()
        return d[:, 0] * self.random_numbers_ratio, d[:, 1] * self.random_numbers_ratio

    def observeParticles(self, positions):
        """(N, N_d, 2) hu, hv of every particle at the true drifter positions
        (observeUnderlyingFlow, BaseOceanStateEnsemble.py:105-136)."""
        torch = Common._torch()
        ens = self.ensemble
        pos = np.asarray(positions, dtype=np.float32).reshape(-1, 2)
        with torch.cuda.stream(ens.gpu_stream):
            d_pos = torch.from_numpy(pos).to(ens.state[0].device)
            HX = ens.observe(d_pos)[:, :pos.shape[0], :].to("cpu").numpy().astype(np.float64)
        return HX

    def _innovations(self, obs, innovations):
        pos = obs[:, :2]
        N, Nd = self.ensemble.n, pos.shape[0]
        if innovations is None:
            innovations = obs[None, :, 2:4] - self.observeParticles(pos)
        return pos, np.asarray(innovations, dtype=np.float64).reshape(N, Nd, 2)

    def _add_kalman_gain(self, pos, d):
        """Every particle's state plus the Kalman-gain term of every drifter;
        returns phi = sum_d d^T S d per particle (addKalmanGain,
        IEWPFOcean.py:548-595)."""
        torch = Common._torch()
        ens = self.ensemble
        st = ens.gpu_stream
        s = G.stream_handle(st)
        N, Nd = d.shape[0], d.shape[1]
        S = self.S_host.astype(np.float64)
        e = np.einsum("ij,pdj->pdi", S, d)
        phi = np.einsum("pdi,pdi->p", d, e)
        e32 = e.astype(np.float32)                       # the reference passes e as float32
        dev = ens.state[0].device
        if self.interpolation_factor == 1:
            # the balance reads H + eta + d_eta (ocean_noise.cu:443-504): each
            # drifter's term depends on the state, so the drifters go one after
            # the other, every particle per launch
            cells = self._coarse_cells(pos)
            with torch.cuda.stream(st):
                d_e = torch.from_numpy(np.ascontiguousarray(e32)).to(dev)
                if self._patch_device is None:
                    self._patch_device = torch.from_numpy(self._kalman_patch().astype(np.float32)).to(dev)
            for dd in range(Nd):
                G.check(self.lib.gocean_iewpf_kalman_patch(
                    self.kalman_field.data_ptr(), self.offsets.data_ptr(), N, self.coarse_nx, self.coarse_ny,
                    self._patch_device.data_ptr(), d_e.data_ptr(), Nd, dd, int(cells[dd, 0]), int(cells[dd, 1]), s),
                    "gocean_iewpf_kalman_patch")
                G.check(self.lib.gocean_noise_batch_perturb(
                    ens.noise, ens.state[0].data_ptr(), ens.Hi.data_ptr(), None, 1.0, 1.0, 0,
                    G.ptr(ens.noise_seed), self.kalman_field.data_ptr(), ens.noise_coarse.data_ptr(),
                    ens.noise_key, ens.noise_counter.data_ptr(), s), "gocean_noise_batch_perturb")
        else:
            # the interpolated balance reads H + d_eta only: the term is a fixed
            # quadratic function of e per drifter, one gather for all particles
            W = self._kalman_basis(pos)
            plane = (self.ny + 4) * (self.nx + 4)
            assert all(int(ens.state_off[k]) == 3 * plane * k for k in range(N)), "members must be contiguous"
            ex, ey = e32[:, :, 0].astype(np.float64), e32[:, :, 1].astype(np.float64)
            coef = kalman_window_coefficients(np.stack([ex, ey, ex * ex, ey * ey, ex * ey], axis=2))
            with torch.cuda.stream(st):
                E = torch.from_numpy(coef).to(dev)
                G.check(self.lib.gocean_iewpf_kalman_window(
                    ens.state[0].data_ptr(), N, self.nx, self.ny, 3 * plane, W["basis"].data_ptr(), Nd, W["wh"],
                    W["ww"], W["origins"].data_ptr(), E.data_ptr(), coef.shape[2], W["tiles"].data_ptr(),
                    W["n_tiles"], W["list_off"].data_ptr(), W["list"].data_ptr(), s), "gocean_iewpf_kalman_window")
        return phi

    def iewpf_2stage(self, observations, innovations=None, sample=True):
        """Two-stage IEWPF analysis of every particle in place
        (IEWPFOcean.iewpf_2stage, IEWPFOcean.py:195-329; the final model
        step before it is the caller's).

        observations: (N_d, 4) [x, y, hu, hv] of the true drifters
        (observeTrueState).  innovations: (N, N_d, 2) y - H x of each
        particle; default: computed here from the particles' observations at
        the true drifter positions.  sample: draw new perpendicular xi, nu
        (False: use the ones samplePerpendicular left)."""
        torch = Common._torch()
        ens = self.ensemble
        st = ens.gpu_stream
        s = G.stream_handle(st)
        obs = np.asarray(observations, dtype=np.float64).reshape(-1, 4)
        pos, d = self._innovations(obs, innovations)
        N, Nd = d.shape[0], d.shape[1]
        dev = ens.state[0].device
        w_rest = -np.log(1.0 / N) * np.ones(N)

        # stage 1: Kalman gain
        phi = self._add_kalman_gain(pos, d)
        if sample:
            gamma, nu_norm = self.samplePerpendicular()
        else:
            self._dots()
            gamma, nu_norm = self._norms()
        if not (np.all(np.isfinite(phi)) and np.all(np.isfinite(gamma)) and np.all(np.isfinite(nu_norm))):
            raise RuntimeError("IEWPF: non-finite phi, gamma or nu norm")

        # target weight and beta (IEWPFOcean.py:263-289)
        c = phi + w_rest
        w_target = np.mean(c)
        beta = np.min((w_target - c) / nu_norm) + 1.0
        if beta < 0:
            beta = 0.0
        if self.beta is not None:
            beta = self.beta
        self.betas = np.append(self.betas, beta)

        # stage 2: alpha from the Lambert W (IEWPFOcean.solveImplicitEquation, :804-857)
        c_star = w_target - c - (beta - 1) * nu_norm
        arg = -(gamma / self.Nx) * np.exp(-gamma / self.Nx) * np.exp(-c_star / self.Nx)
        alpha = -(self.Nx / gamma) * np.real(lambertw(arg, k=0))
        cells = self._coarse_cells(pos)
        with torch.cuda.stream(st):
            a = torch.from_numpy(np.sqrt(alpha).astype(np.float32)).to(dev)
            d_cells = torch.from_numpy(cells).to(dev)
            G.check(self.lib.gocean_iewpf_axpby(ens.noise_random.data_ptr(), self.nu.data_ptr(),
                                                self.offsets.data_ptr(), N, self.n_coarse, a.data_ptr(),
                                                float(np.float32(np.sqrt(beta))), s), "gocean_iewpf_axpby")
            G.check(self.lib.gocean_iewpf_local_svd(ens.noise_random.data_ptr(), self.offsets.data_ptr(), N,
                                                    self.coarse_nx, self.coarse_ny,
                                                    self.localSVD_device.data_ptr(), d_cells.data_ptr(), Nd, s),
                    "gocean_iewpf_local_svd")
        # P^1/2 (alpha^1/2 xi + beta^1/2 nu) added to every particle, ghosts refreshed
        ens.perturb(update_random_field=False)
        self.last = dict(phi=phi, gamma=gamma, nu_norm=nu_norm, c=c, w_target=w_target, beta=beta, alpha=alpha)
        return self.last

    # ------------------------------------------------------------------
    # one-stage IEWPF (IEWPFOcean.iewpf, IEWPFOcean.py:331-394)
    # ------------------------------------------------------------------
    def sampleXi(self):
        """xi ~ N(0, I) on every particle's coarse grid (generateNormalDistribution)."""
        ens = self.ensemble
        G.check(self.lib.gocean_noise_batch_generate(ens.noise, G.ptr(ens.noise_seed), ens.noise_random.data_ptr(),
                                                     ens.noise_key, ens.noise_counter.data_ptr(),
                                                     G.stream_handle(ens.gpu_stream)), "gocean_noise_batch_generate")

    def iewpf(self, observations, innovations=None, sample=True):
        """One-stage IEWPF analysis of every particle in place
        (IEWPFOcean.iewpf, IEWPFOcean.py:331-394): the Kalman gain, the target
        weight max(c) (obtainTargetWeight, :742-762), xi ~ N(0, I) with gamma
        = |xi|^2 scaled (sampleFromP, :597-627), alpha from the Lambert W, and
        the perturbation P^1/2 xi scaled by sqrt(alpha) (the scale is applied
        to xi, the same by linearity).  sample=False uses the xi sampleXi
        left."""
        torch = Common._torch()
        ens = self.ensemble
        st = ens.gpu_stream
        s = G.stream_handle(st)
        obs = np.asarray(observations, dtype=np.float64).reshape(-1, 4)
        pos, d = self._innovations(obs, innovations)
        N, Nd = d.shape[0], d.shape[1]
        dev = ens.state[0].device
        w_rest = -np.log(1.0 / N) * np.ones(N)
        phi = self._add_kalman_gain(pos, d)
        c = phi + w_rest
        w_target = np.max(c)
        if sample:
            self.sampleXi()
        G.check(self.lib.gocean_iewpf_dots(ens.noise_random.data_ptr(), ens.noise_random.data_ptr(),
                                           self.offsets.data_ptr(), N, self.n_coarse, self.dots.data_ptr(), s),
                "gocean_iewpf_dots")
        gamma = self._norms()[0]
        c_star = w_target - c
        arg = -(gamma / self.Nx) * np.exp(-gamma / self.Nx) * np.exp(-c_star / self.Nx)
        alpha = -(self.Nx / gamma) * np.real(lambertw(arg, k=0))
        if not np.all(np.isfinite(alpha)):
            raise RuntimeError("IEWPF: non-finite alpha")
        cells = self._coarse_cells(pos)
        with torch.cuda.stream(st):
            a = torch.from_numpy(np.sqrt(alpha).astype(np.float32)).to(dev)
            d_cells = torch.from_numpy(cells).to(dev)
            G.check(self.lib.gocean_iewpf_local_svd(ens.noise_random.data_ptr(), self.offsets.data_ptr(), N,
                                                    self.coarse_nx, self.coarse_ny,
                                                    self.localSVD_device.data_ptr(), d_cells.data_ptr(), Nd, s),
                    "gocean_iewpf_local_svd")
            G.check(self.lib.gocean_iewpf_axpby(ens.noise_random.data_ptr(), ens.noise_random.data_ptr(),
                                                self.offsets.data_ptr(), N, self.n_coarse, a.data_ptr(), 0.0, s),
                    "gocean_iewpf_axpby")
        ens.perturb(update_random_field=False)
        self.last = dict(phi=phi, gamma=gamma, c=c, w_target=w_target, alpha=alpha)
        return self.last
