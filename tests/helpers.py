"""Shared test helpers: the reference test-suite's initial conditions, oracle
parameter construction, and a thin device harness around the C ABI."""

from __future__ import annotations

import numpy as np

from oracle import cdklm16_np as O

# ---------------------------------------------------------------------------
# the reference's CDKLM16 regression setup (gpu_ocean/tests/schemes/
# CDKLM16_test.py:39-67) and IC generators (gpu_ocean/tests/testUtils.py:99-167)
# ---------------------------------------------------------------------------
REF_NX, REF_NY, REF_DX, REF_DY, REF_DT, REF_T = 50, 70, 200.0, 200.0, 0.9, 50.0


def ref_bump(eta, nx, ny, dx, dy, xc, yc):
    """testUtils.addCentralBump/addCornerBump/addUpperCornerBump: adds
    exp(-(x^2/size + y^2/size)) with size = 500*min(dx,dy) where the distance
    is below ``size`` (always true at these sizes)."""
    i = np.arange(-2, nx + 2)
    j = np.arange(-2, ny + 2)
    X = dx * i[None, :] - xc
    Y = dy * j[:, None] - yc
    size = 500.0 * min(dx, dy)
    val = np.exp(-(X ** 2 / size + Y ** 2 / size))
    val = np.where(np.sqrt(X ** 2 + Y ** 2) < size, val, 0.0)
    eta += val.astype(eta.dtype) if eta.dtype == np.float64 else val
    return eta


def ref_bottom_topography(Hi, nx, ny):
    """testUtils.makeBottomTopography (intersections) :99-106."""
    j = np.arange(-2, ny + 3)[:, None].astype(np.float64)
    i = np.arange(-2, nx + 3)[None, :].astype(np.float64)
    Hi[:, :] = 6 + 2.0 * np.cos(0.3 * (i + i / (np.sin(0.5 * j) + 2.5))) + \
        2.0 * np.sin(2 * np.pi * (j + i) / (2.0 * ny))
    return Hi


REF_BCS = {"wall": (1, 1, 1, 1), "periodic": (2, 2, 2, 2),
           "periodicNS": (2, 1, 2, 1), "periodicEW": (1, 2, 1, 2)}


def ref_cases():
    """The 13 CDKLM16 regression cases: (name, bc, ic, golden stem, f, beta,
    bathymetry)."""
    cases = []
    for bcn, bc in REF_BCS.items():
        for ic in ("central", "corner", "upperCorner"):
            stem = ("wallBC" if (ic == "central" or bcn == "wall") else bcn) + "_" + ic + "_"
            cases.append((bcn + "_" + ic, bc, ic, stem, 0.0, 0.0, False))
    cases.append(("coriolis_central", (1, 1, 1, 1), "central", "coriolis_central_", 0.01, 0.0, False))
    cases.append(("betamodel_central", (1, 1, 1, 1), "central", "betamodel_central_",
                  0.01 - 2 * 1e-6 * REF_DY, 1e-6, False))
    cases.append(("bathymetry_central", (1, 1, 1, 1), "central", "wallBC_central_bathymetry_",
                  0.0, 0.0, True))
    return cases


def ref_initial(ic, bathymetry):
    nx, ny, dx, dy = REF_NX, REF_NY, REF_DX, REF_DY
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    centres = {"central": (dx * nx / 2.0, dy * ny / 2.0), "corner": (4 * dx, 4 * dy),
               "upperCorner": ((nx - 4) * dx, (ny - 4) * dy)}
    eta64 = ref_bump(np.zeros_like(eta, dtype=np.float64), nx, ny, dx, dy, *centres[ic])
    eta = eta64.astype(np.float32)
    Hi = np.ones((ny + 5, nx + 5), np.float32) * 60
    if bathymetry:
        Hi = ref_bottom_topography(np.zeros((ny + 5, nx + 5)), nx, ny).astype(np.float32)
    return eta, np.zeros_like(eta), np.zeros_like(eta), Hi


def legacy_beta_field(f, beta, nx=REF_NX, ny=REF_NY, dy=REF_DY):
    """Coriolis field f + beta*(j+1/2)*dy (ghost-inclusive row j): the
    convention the reference's beta-plane golden file was produced with."""
    j = np.arange(ny + 4)[:, None]
    return np.repeat(np.float32(f) + np.float32(beta) * ((j + 0.5) * dy), nx + 4, 1).astype(np.float32)


def current_beta_field(f, beta, nx=REF_NX, ny=REF_NY, dx=REF_DX, dy=REF_DY):
    """Coriolis field the current reference constructor builds
    (CDKLM16.py:146-148 compensation, :295-302 cell-centre coordinates)."""
    f_used = np.float32(f - beta * (2 * dx * 0.0 + 2 * dy * 1.0))
    y = np.linspace(-1.5 * np.float32(dy), (ny + 1.5) * np.float32(dy), ny + 4)
    return np.repeat((f_used + np.float32(beta) * y)[:, None], nx + 4, 1).astype(np.float32)


def load_goldens():
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "timestep50_cdklm16.npz")
    return np.load(p)


def golden_errors(res, gold, stem):
    """The reference's checkResults metrics (CDKLM16_test.py:114-145):
    max signed, max abs and L2 error over the interior, each divided by
    max|ref|."""
    out = {}
    for a, k in zip(res, ("eta", "u", "v")):
        r = gold[stem + k].astype(np.float64)
        d = np.asarray(a)[2:-2, 2:-2].astype(np.float64) - r
        m = np.max(np.abs(r))
        out[k] = (float(np.max(d) / m), float(np.max(np.abs(d)) / m), float(np.linalg.norm(d) / m))
    return out


def oracle_params(inp, **kw):
    return O.Params(nx=inp.nx, ny=inp.ny, dx=inp.dx, dy=inp.dy, g=inp.g, r=inp.r, bc=inp.bc,
                    f=inp.f, wind_x=(inp.wind_x, inp.wind_x), wind_y=(inp.wind_y, inp.wind_y), **kw)


TINY32 = np.float32(2.0 ** -126)


def normal_equal(a, b):
    """Bit equality of every value of normal float32 magnitude; values below
    2^-126 count as zero on both sides.  The reference's CUDA build flushes
    subnormals to zero (--ftz=true, gpu_ocean/SWESimulators/CDKLM16.py:197),
    and the HIP kernels' division corrections are exact for normal quotients
    only, so subnormal results may differ from the IEEE oracle in their last
    bits."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    fa = np.where(np.abs(a) < TINY32, np.float32(0.0), a)
    fb = np.where(np.abs(b) < TINY32, np.float32(0.0), b)
    return np.array_equal(fa, fb)


def diff_report(a, b, n=4):
    """First few cells where a and b differ: (index, a, b)."""
    a = np.asarray(a)
    b = np.asarray(b)
    idx = np.argwhere(a != b)[:n]
    return [(tuple(int(v) for v in i), float(a[tuple(i)]), float(b[tuple(i)])) for i in idx]


def max_rel(a, b):
    """max |a-b| / max |b| (the normalisation of the reference's tests)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (m if m > 0 else 1.0))


# ---------------------------------------------------------------------------
# device harness over the C ABI (GPU tests only)
# ---------------------------------------------------------------------------
class DeviceDomain:
    """Device copies of one domain's state and depth plus the C-ABI calls."""

    def __init__(self, inp, pitch_extra=0, fields=None, rk_order=2, r=None, theta=1.3,
                 north=(0.0, 1.0), f=None, wind=None, device="cuda"):
        import torch
        from SWESimulators import Common
        from SWESimulators import _gocean as G
        self.G = G
        self.lib = G.load()
        self.torch = torch
        self.inp = inp
        nx, ny = inp.nx, inp.ny
        self.pitch = nx + 4 + pitch_extra
        self.hi_pitch = nx + 5 + pitch_extra
        self.dev = torch.device(device)
        fval = inp.f if f is None else f
        self.dom = Common.make_domain(nx, ny, inp.dx, inp.dy, inp.g, inp.r if r is None else r,
                                      theta, rk_order, inp.bc, pitch=self.pitch,
                                      hi_pitch=self.hi_pitch, f=fval if fields is None else 0.0,
                                      north_x=north[0], north_y=north[1])
        self.wind = G.Wind()
        if wind is not None:
            self.wind.t, self.wind.x0, self.wind.x1, self.wind.y0, self.wind.y1 = wind
        else:
            self.wind.x0 = self.wind.x1 = inp.wind_x
            self.wind.y0 = self.wind.y1 = inp.wind_y
        self.Hi = self._pad(inp.Hi, ny + 5, self.hi_pitch)
        self.Hm = torch.zeros((ny + 4, self.pitch), dtype=torch.float32, device=self.dev)
        self.fields = G.Fields()
        self.field_t = None
        if fields is not None:
            self.field_t = [self._pad(a, ny + 4, self.pitch) for a in fields]
            self.fields.f, self.fields.north_x, self.fields.north_y = (t.data_ptr() for t in self.field_t)
        self.scratch = torch.zeros(1026, dtype=torch.float32, device=self.dev)

    def _pad(self, a, rows, pitch):
        torch = self.torch
        t = torch.zeros((rows, pitch), dtype=torch.float32, device=self.dev)
        t[:, :a.shape[1]] = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(self.dev)
        return t

    def state(self, eta, hu, hv):
        ny = self.inp.ny
        return [self._pad(x, ny + 4, self.pitch) for x in (eta, hu, hv)]

    def host(self, ts, cols=None):
        cols = self.inp.nx + 4 if cols is None else cols
        return [t[:, :cols].cpu().numpy() for t in ts]

    def bathymetry_init(self):
        self.G.check(self.lib.gocean_bathymetry_init(self.dom, self.Hi.data_ptr(), self.Hm.data_ptr(), None),
                     "bathymetry_init")

    def substep(self, dt, rk, src, dst, qn=None, tile=-1, dt_dev=None):
        P = self.G.ptr
        q = qn if qn is not None else (None, None, None)
        self.G.check(self.lib.gocean_cdklm_substep(
            self.dom, float(dt), P(dt_dev), rk, P(src[0]), P(src[1]), P(src[2]),
            P(q[0]), P(q[1]), P(q[2]), P(dst[0]), P(dst[1]), P(dst[2]),
            self.Hi.data_ptr(), self.fields, self.wind, tile, None), "substep")

    def rk_steps(self, n, dt, h0, h1, h2=None, tile=-1, dt_dev=None):
        P = self.G.ptr
        h2 = h2 if h2 is not None else (None, None, None)
        self.G.check(self.lib.gocean_rk_steps(
            self.dom, int(n), float(dt), P(dt_dev), P(h0[0]), P(h0[1]), P(h0[2]),
            P(h1[0]), P(h1[1]), P(h1[2]), P(h2[0]), P(h2[1]), P(h2[2]),
            self.Hi.data_ptr(), self.fields, self.wind, tile, None), "rk_steps")

    def periodic_bc(self, h):
        self.G.check(self.lib.gocean_periodic_bc(self.dom, h[0].data_ptr(), h[1].data_ptr(),
                                                 h[2].data_ptr(), None), "periodic_bc")

    def max_dt(self, h, courant=0.0):
        self.G.check(self.lib.gocean_max_dt(self.dom, h[0].data_ptr(), h[1].data_ptr(), h[2].data_ptr(),
                                            self.Hm.data_ptr(), float(courant), self.scratch.data_ptr(), None),
                     "max_dt")
        self.torch.cuda.synchronize()
        return self.scratch[:2].cpu().numpy()


def random_case(nx, ny, bc, seed, land=False, shallow=False, fvar=False, angvar=False, r=0.0,
          wind=(0.0, 0.0), dx=200.0, dy=150.0, dyadic=False):
    """Random smooth-bathymetry case with random state (no flat regions, so
    branch ties are improbable); optional shallow strip, land block, Coriolis /
    rotation fields, friction and wind.  Returns (p, Hi, Hm, eta, hu, hv)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny + 5, 0:nx + 5]
    Hi = (30 + 10 * np.sin(x / 3.0) + 8 * np.cos(y / 2.5)).astype(np.float32)
    if shallow:
        w = (nx + 5) // 3
        Hi[:, :w] = (0.05 + 0.02 * rng.random((ny + 5, w))).astype(np.float32)
    if dyadic:
        # multiples of 2^-12: sums of four corners are exact in float32 and
        # float64, so a ghost cell's depth equals its mirror image's in both
        # precisions and wall-face upwind ties resolve the same way in both
        Hi = (np.round(Hi.astype(np.float64) * 4096.0) / 4096.0).astype(np.float32)
    if land:
        Hi[ny // 2:ny // 2 + 3, nx // 2:nx // 2 + 2] = O.LAND_VALUE
    f = np.float32(1.2e-4)
    if fvar:
        f = (1e-4 + 2e-5 * rng.random((ny + 4, nx + 4))).astype(np.float32)
    nxv, nyv = np.float32(0.0), np.float32(1.0)
    if angvar:
        ang = (0.3 * rng.random((ny + 4, nx + 4))).astype(np.float32)
        nxv, nyv = np.sin(ang).astype(np.float32), np.cos(ang).astype(np.float32)
    p = O.Params(nx=nx, ny=ny, dx=dx, dy=dy, bc=bc, f=f, north_x=nxv, north_y=nyv, r=r,
                 wind_x=(wind[0], wind[0] * 1.5), wind_y=(wind[1], wind[1] * 0.5))
    Hi = O.bathymetry_halo(Hi, p)
    Hm = O.init_bm(Hi)
    eta = (0.3 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    eta = np.maximum(eta, -Hm)
    hu = (2 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    hv = (2 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    if shallow:
        w = (nx + 4) // 3
        eta[:, :w] = -Hm[:, :w] + (0.08 * rng.random((ny + 4, w))).astype(np.float32)
        hu[:, :w] *= 0.001
        hv[:, :w] *= 0.001
    O.periodic_bc(eta, hu, hv, p)
    return p, Hi, Hm, eta, hu, hv


def hi64(Hi):
    """Corner depths in float64 with land kept at exactly the land value: the
    float32 land marker 1e20 widens to 1.00000002e20, which the float64
    land test |H - 1e20| < 1e-3 would no longer recognise."""
    Hi = np.asarray(Hi)
    return np.where(Hi == np.float32(O.LAND_VALUE), float(O.LAND_VALUE), Hi.astype(np.float64))
