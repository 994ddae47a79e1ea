"""Batched CDKLM16 ensemble: all members (of any sizes) advance together.

The reference steps ensemble members one by one from Python
(OceanModelEnsemble.modelStep, gpu_ocean/SWESimulators/OceanModelEnsemble.py:
105-117; EnsembleFromFiles.stepToObservation, EnsembleFromFiles.py:302-325),
each step being several launches plus host work.  Here one launch per RK
substep covers every member's tiles (``gocean_ensemble_*`` in
``lib/libgocean_hip.so``), dt lives on the device per member, and a run of n
steps can be captured in a HIP graph and replayed without host work.

State layout (HBM): three flat float32 buffers (h0, h1, and h2 for RK3); member
m's eta, hu, hv are consecutive (ny+4) x (nx+4) planes at ``state_offset[m]``.
"""

from __future__ import annotations

import ctypes
import logging

import numpy as np

from . import Common
from . import _gocean as G


def _get(obj, name, default=None):
    if isinstance(obj, dict):
        return obj.get(name, default)
    return getattr(obj, name, default)


class CDKLM16Ensemble(object):
    """Ensemble of CDKLM16 domains stepped by batched HIP launches.

    ``members``: sequence of objects or dicts with ``eta``/``eta0``, ``hu``/
    ``hu0``, ``hv``/``hv0`` ((ny+4, nx+4)), ``Hi``/``H`` ((ny+5, nx+5)), ``nx``,
    ``ny``, ``dx``, ``dy`` and optionally ``g``, ``f``, ``r``, ``bc``
    (north, east, south, west), ``wind_x``, ``wind_y``.
    """

    def __init__(self, gpu_ctx, members, rk_order=2, theta=1.3, courant_number=0.8,
                 dt=None, tile_config=None, flux_slope_eps=1.0e-1,
                 desingularization_eps=1.0e-1, depth_cutoff=1.0e-5):
        torch = Common._torch()
        self.logger = logging.getLogger(__name__)
        self.gpu_ctx = gpu_ctx
        self.lib = G.load()
        self.gpu_stream = gpu_ctx.new_stream()
        self.rk_order = int(rk_order)
        self.courant_number = courant_number
        self.n = len(members)
        dev = gpu_ctx.torch_device
        self.nx = [int(_get(m, "nx")) for m in members]
        self.ny = [int(_get(m, "ny")) for m in members]
        self.pitch = [nx + 4 for nx in self.nx]
        self.planes = [(ny + 4) * p for ny, p in zip(self.ny, self.pitch)]
        self.state_off = np.concatenate([[0], np.cumsum([3 * pl for pl in self.planes])]).astype(np.int64)
        hi_sizes = [(ny + 5) * (nx + 5) for nx, ny in zip(self.nx, self.ny)]
        self.hi_off = np.concatenate([[0], np.cumsum(hi_sizes)]).astype(np.int64)
        self.hm_off = np.concatenate([[0], np.cumsum(self.planes)]).astype(np.int64)
        total = int(self.state_off[-1])
        arr = (G.Member * self.n)()
        st = self.gpu_stream
        with torch.cuda.stream(st):
            self.state = [torch.zeros(total, dtype=torch.float32, device=dev) for _ in range(2)]
            self.state2 = torch.zeros(total, dtype=torch.float32, device=dev) if self.rk_order == 3 else None
            self.Hi = torch.zeros(int(self.hi_off[-1]), dtype=torch.float32, device=dev)
            self.Hm = torch.zeros(int(self.hm_off[-1]), dtype=torch.float32, device=dev)
            self.dt = torch.zeros(self.n, dtype=torch.float32, device=dev)
        self.bcs = []
        for k, m in enumerate(members):
            nx, ny = self.nx[k], self.ny[k]
            bc = _get(m, "bc", (1, 1, 1, 1))
            if isinstance(bc, Common.BoundaryConditions):
                bc = bc.codes()
            self.bcs.append(tuple(int(b) for b in bc))
            d = Common.make_domain(nx, ny, _get(m, "dx"), _get(m, "dy"), _get(m, "g", 9.81),
                                   _get(m, "r", 0.0), theta, self.rk_order, bc,
                                   pitch=self.pitch[k], hi_pitch=nx + 5, f=_get(m, "f", 0.0),
                                   desing_eps=desingularization_eps, flux_slope_eps=flux_slope_eps,
                                   depth_cutoff=depth_cutoff)
            arr[k].domain = d
            arr[k].state_offset = int(self.state_off[k])
            arr[k].hi_offset = int(self.hi_off[k])
            arr[k].hm_offset = int(self.hm_off[k])
            arr[k].field_offset = -1
            wx = float(np.float32(_get(m, "wind_x", 0.0)))
            wy = float(np.float32(_get(m, "wind_y", 0.0)))
            arr[k].wind.t, arr[k].wind.x0, arr[k].wind.x1 = 0.0, wx, wx
            arr[k].wind.y0, arr[k].wind.y1 = wy, wy
            Hi = np.asarray(_get(m, "Hi", _get(m, "H")), dtype=np.float32)
            assert Hi.shape == (ny + 5, nx + 5), (Hi.shape, (ny + 5, nx + 5))
            with torch.cuda.stream(st):
                self.Hi[self.hi_off[k]:self.hi_off[k + 1]].copy_(torch.from_numpy(Hi.ravel()))
        self.members_c = arr
        h = ctypes.c_void_p()
        G.check(self.lib.gocean_ensemble_create(arr, self.n, -1 if tile_config is None else int(tile_config),
                                                ctypes.byref(h)), "gocean_ensemble_create")
        self.handle = h
        G.check(self.lib.gocean_ensemble_bathymetry_init(self.handle, self.Hi.data_ptr(), self.Hm.data_ptr(),
                                                         G.stream_handle(st)), "ensemble bathymetry")
        st.synchronize()
        Hm_host = self.Hm.to("cpu").numpy()
        for k, m in enumerate(members):
            e0 = np.asarray(_get(m, "eta", _get(m, "eta0")), dtype=np.float32)
            hm = Hm_host[self.hm_off[k]:self.hm_off[k + 1]].reshape(self.ny[k] + 4, self.pitch[k])
            e0 = np.maximum(e0, -hm)                      # CDKLM16.py:237-238
            self.upload(k, e0, _get(m, "hu", _get(m, "hu0")), _get(m, "hv", _get(m, "hv0")))
        # boundary conditions at t == 0 (CDKLM16.py:502-505) on both buffers
        for s in self.state:
            G.check(self.lib.gocean_ensemble_periodic_bc(self.handle, s.data_ptr(), G.stream_handle(st)),
                    "ensemble periodic bc")
        self.t = np.zeros(self.n)
        self.num_iterations = 0
        if dt is None:
            self.update_dt()
        else:
            self.set_dt(dt)

    # ------------------------------------------------------------------
    def views(self, k, buf=0):
        """(eta, hu, hv) device views of member k in state buffer ``buf``."""
        s = self.state[buf] if buf < 2 else self.state2
        o, pl = int(self.state_off[k]), self.planes[k]
        shape = (self.ny[k] + 4, self.pitch[k])
        return tuple(s[o + i * pl:o + (i + 1) * pl].view(shape) for i in range(3))

    def upload(self, k, eta, hu, hv):
        torch = Common._torch()
        with torch.cuda.stream(self.gpu_stream):
            for buf in (0, 1):
                for v, a in zip(self.views(k, buf), (eta, hu, hv)):
                    v.copy_(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)))

    def download(self, k, interior_domain_only=False):
        torch = Common._torch()
        with torch.cuda.stream(self.gpu_stream):
            out = [v.to("cpu").numpy() for v in self.views(k)]
        if interior_domain_only:
            out = [a[2:-2, 2:-2] for a in out]
        return tuple(out)

    def set_dt(self, dt):
        torch = Common._torch()
        d = np.broadcast_to(np.asarray(dt, dtype=np.float32), (self.n,)).copy()
        with torch.cuda.stream(self.gpu_stream):
            self.dt.copy_(torch.from_numpy(d))

    def update_dt(self, courant_number=None):
        """dt[m] = courant * CFL minimum of member m, on the device."""
        c = self.courant_number if courant_number is None else courant_number
        G.check(self.lib.gocean_ensemble_max_dt(self.handle, self.state[0].data_ptr(), self.Hm.data_ptr(),
                                                float(c), self.dt.data_ptr(), G.stream_handle(self.gpu_stream)),
                "gocean_ensemble_max_dt")

    def dt_host(self):
        self.gpu_stream.synchronize()
        return self.dt.cpu().numpy()

    def _ptr2(self):
        return G.ptr(self.state2) if self.state2 is not None else None

    def step_n(self, n_steps):
        """``n_steps`` RK steps of every member with its own device dt."""
        G.check(self.lib.gocean_ensemble_rk_steps(
            self.handle, int(n_steps), self.state[0].data_ptr(), self.state[1].data_ptr(), self._ptr2(),
            self.Hi.data_ptr(), None, self.dt.data_ptr(), G.stream_handle(self.gpu_stream)),
            "gocean_ensemble_rk_steps")
        self.num_iterations += int(n_steps)

    def capture(self, n_steps):
        """Capture ``n_steps`` RK steps of all members in a HIP graph."""
        self.gpu_stream.synchronize()
        G.check(self.lib.gocean_ensemble_capture(
            self.handle, int(n_steps), self.state[0].data_ptr(), self.state[1].data_ptr(), self._ptr2(),
            self.Hi.data_ptr(), None, self.dt.data_ptr()), "gocean_ensemble_capture")
        self.graph_steps = int(n_steps)

    def replay(self):
        G.check(self.lib.gocean_ensemble_replay(self.handle, G.stream_handle(self.gpu_stream)),
                "gocean_ensemble_replay")
        self.num_iterations += self.graph_steps

    def cells(self):
        return int(sum(nx * ny for nx, ny in zip(self.nx, self.ny)))

    def cleanUp(self):
        if getattr(self, "handle", None) is not None:
            self.lib.gocean_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.cleanUp()
        except Exception:
            pass
