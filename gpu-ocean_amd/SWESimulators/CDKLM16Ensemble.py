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
        self.tile_config = -1 if tile_config is None else int(tile_config)
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
        self.noise = None
        self.drifters = None
        self.f = [float(np.float32(_get(m, "f", 0.0))) for m in members]
        self.g = [float(np.float32(_get(m, "g", 9.81))) for m in members]
        self.dxy = [(float(np.float32(_get(m, "dx"))), float(np.float32(_get(m, "dy")))) for m in members]
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

    def _rk(self, n_steps):
        G.check(self.lib.gocean_ensemble_rk_steps(
            self.handle, int(n_steps), self.state[0].data_ptr(), self.state[1].data_ptr(), self._ptr2(),
            self.Hi.data_ptr(), None, self.dt.data_ptr(), G.stream_handle(self.gpu_stream)),
            "gocean_ensemble_rk_steps")

    def step_n(self, n_steps, model_error=False):
        """``n_steps`` RK steps of every member with its own device dt; with
        ``model_error`` each step is followed by the batched perturbation and
        the ghost-cell exchange (CDKLM16.step with a small-scale
        perturbation, CDKLM16.py:580-586)."""
        if self.rk_order == 1 and (int(n_steps) % 2 or model_error or self.drifters is not None):
            raise ValueError("an rk_order 1 ensemble advances in pairs of Euler steps (the two state "
                             "buffers trade roles every step): use an even n_steps, without per-step "
                             "model error or drifters")
        if not model_error and self.drifters is None:
            self._rk(n_steps)
        else:
            for _ in range(int(n_steps)):
                self._rk(1)
                if model_error:
                    self.perturb()
                if self.drifters is not None:
                    self.drift()
        self.num_iterations += int(n_steps)

    # ------------------------------------------------------------------
    # model error (OceanStateNoise for every member in batched launches)
    # ------------------------------------------------------------------
    # ------------------------------------------------------------------
    # time-targeted steps (CDKLM16.step / dataAssimilationStep for every member)
    # ------------------------------------------------------------------
    def _need_single_steps(self, what):
        # members that finish early sit out a launch with dt 0 and leave the
        # other buffer untouched, so single Euler steps cannot swap buffers
        if self.rk_order == 1:
            raise NotImplementedError(f"{what} needs rk_order 2 or 3: an rk_order 1 ensemble advances "
                                      "in pairs of Euler steps")

    def _dt_sequences(self, dt64, t_span):
        """Per member, the local time steps of CDKLM16.step(t_span) with the
        member's dt (CDKLM16.py:508-593): local_dt = float32(min(dt,
        float32(t_span - t_now))), t_now accumulated in float64."""
        seqs = []
        for d in dt64:
            t_now, seq = 0.0, []
            while t_now < t_span:
                ld = np.float32(min(d, np.float32(t_span - t_now)))
                seq.append(ld)
                t_now += np.float64(ld)
            seqs.append(seq)
        return seqs

    def update_dt_host(self, courant_number=None):
        """Per-member dt as the simulator's updateDt sets it: courant times
        the float32 CFL minimum, in float64 (CDKLM16.py:712-739); the device
        dt gets the float32 value."""
        c = self.courant_number if courant_number is None else courant_number
        G.check(self.lib.gocean_ensemble_max_dt(self.handle, self.state[0].data_ptr(), self.Hm.data_ptr(),
                                                0.0, self.dt.data_ptr(), G.stream_handle(self.gpu_stream)),
                "gocean_ensemble_max_dt")
        raw = self.dt_host().astype(np.float64)
        self.dt64 = float(c) * raw
        self.set_dt(self.dt64.astype(np.float32))
        return self.dt64

    def step_to(self, t_span, dt64=None):
        """Advance every member by ``t_span`` seconds as CDKLM16.step(t_span)
        without model error (CDKLM16.py:494-598): each member's own dt
        sequence, the last step shortened, members that are done skipped
        (dt 0 in their column of the launch's dt row).  Drifters, when
        enabled, move with the same local steps."""
        torch = Common._torch()
        self._need_single_steps("step_to")
        if dt64 is None:
            if getattr(self, "dt64", None) is None:
                self.update_dt_host()
            dt64 = self.dt64
        dt64 = np.asarray(dt64, dtype=np.float64)
        seqs = self._dt_sequences(dt64, float(t_span))
        steps = max(len(q) for q in seqs)
        table = np.zeros((steps, self.n), dtype=np.float32)
        for m, q in enumerate(seqs):
            table[:len(q), m] = q
        st = G.stream_handle(self.gpu_stream)
        with torch.cuda.stream(self.gpu_stream):
            d_table = torch.from_numpy(table).to(self.state[0].device)
        for i in range(steps):
            row = d_table[i]
            G.check(self.lib.gocean_ensemble_rk_steps(
                self.handle, 1, self.state[0].data_ptr(), self.state[1].data_ptr(), self._ptr2(),
                self.Hi.data_ptr(), None, row.data_ptr(), st), "gocean_ensemble_rk_steps")
            if self.drifters is not None:
                G.check(self.lib.gocean_drift_batch_step(self.drifters, self.state[0].data_ptr(),
                                                         self.Hm.data_ptr(), self.drifter_positions.data_ptr(),
                                                         row.data_ptr(), st), "gocean_drift_batch_step")
        for m, q in enumerate(seqs):
            for ld in q:
                self.t[m] += np.float64(ld)
        self.num_iterations += steps
        return self.t

    def dataAssimilationStep(self, observation_time, model_time_step, model_error_final_step=True,
                             courant_number=0.8):
        """Every member to ``observation_time`` in model time steps, the
        leftover first, the model error after each model step (with
        q0_scale sqrt(leftover / model_time_step) after the leftover, and
        optionally not after the last), dt refreshed every 5 model steps:
        CDKLM16.dataAssimilationStep (CDKLM16.py:650-704) for every member at
        once, as EnsembleFromFiles.stepToObservation runs it member by member
        (EnsembleFromFiles.py:302-325).  Members share their time."""
        self._need_single_steps("dataAssimilationStep")
        t0 = float(self.t[0])
        assert np.all(self.t == t0), "members must share their time"
        mts = float(model_time_step)
        full = int(round(observation_time - t0) / mts)
        leftover = observation_time - t0 - full * mts
        if leftover / mts < 0.1 and full > 1:
            leftover += mts
            full -= 1
        if leftover / mts < 0.00001:
            leftover = 0
        assert full > 0, "There is less than model_time_step until the observation"
        if not hasattr(self, "total_time_steps"):
            self.total_time_steps = 0
        self.update_dt_host(courant_number)
        for i in range(full + 1):
            if i == 0 and leftover == 0:
                continue
            if i == 0:
                self.step_to(leftover)
                if self.noise is not None:
                    self.perturb(q0_scale=np.sqrt(leftover / mts))
            else:
                self.step_to(mts)
                if self.noise is not None and ((i < full) or model_error_final_step):
                    self.perturb()
            self.total_time_steps += 1
            if self.total_time_steps % 5 == 0:
                self.update_dt_host(courant_number)
        assert round(observation_time) == round(float(self.t[0])), \
            "the members' time is not the observation time after dataAssimilationStep"
        return self.t

    def enable_model_error(self, soar_q0=None, soar_L=None, interpolation_factor=1, use_lcg=False,
                           seed=None, beta=0.0):
        """Set up the batched model error of every member: the member-wise
        OceanStateNoise (OceanStateNoise.py:49-170) with one device table.
        ``seed`` seeds the LCG states (use_lcg) or keys the Philox generator."""
        torch = Common._torch()
        arr = (G.NoiseMember * self.n)()
        for k in range(self.n):
            nx, ny = self.nx[k], self.ny[k]
            dx, dy = self.dxy[k]
            f = int(interpolation_factor)
            d = arr[k].noise
            d.nx, d.ny, d.pitch, d.hi_pitch = nx, ny, self.pitch[k], nx + 5
            d.dx, d.dy = dx, dy
            d.interpolation_factor = f
            d.periodic_ns = int(self.bcs[k][0] == Common.PERIODIC)
            d.periodic_ew = int(self.bcs[k][1] == Common.PERIODIC)
            d.soar_q0 = float(np.float32(dx / 100000 if soar_q0 is None else soar_q0))
            d.soar_L = float(np.float32(0.75 * np.float32(dx * f) if soar_L is None else soar_L))
            d.g, d.f, d.beta = self.g[k], self.f[k], float(np.float32(beta))
            d.north_x, d.north_y = 0.0, 1.0
            d.land_value = float(np.float32(1.0e20))
            arr[k].state_offset = int(self.state_off[k])
            arr[k].hi_offset = int(self.hi_off[k])
            arr[k].field_offset = -1
        kind = G.RNG_LCG_NORMAL if use_lcg else G.RNG_PHILOX_NORMAL
        h = ctypes.c_void_p()
        G.check(self.lib.gocean_noise_batch_create(arr, self.n, kind, ctypes.byref(h)),
                "gocean_noise_batch_create")
        self.noise = h
        self.noise_members = arr
        self.noise_use_lcg = use_lcg
        offs = (ctypes.c_int64 * (3 * self.n))()
        tot = (ctypes.c_int64 * 3)()
        G.check(self.lib.gocean_noise_batch_layout(h, offs, tot), "gocean_noise_batch_layout")
        self.noise_offsets = np.array(list(offs), dtype=np.int64).reshape(self.n, 3)
        rs = np.random.RandomState(seed)
        dev = self.gpu_ctx.torch_device
        with torch.cuda.stream(self.gpu_stream):
            self.noise_seed = None
            if use_lcg:
                host = (rs.rand(int(tot[0])) * 2147483648.0).astype(np.uint64).astype(np.int64)
                self.noise_seed = torch.from_numpy(host).to(dev)
            self.noise_random = torch.zeros(max(1, int(tot[1])), dtype=torch.float32, device=dev)
            self.noise_coarse = torch.zeros(max(1, int(tot[2])), dtype=torch.float32, device=dev)
            self.noise_counter = torch.zeros(1, dtype=torch.int64, device=dev)
        self.noise_key = int(rs.randint(0, 2 ** 31)) | (int(rs.randint(0, 2 ** 31)) << 31)

    def perturb(self, q0_scale=1.0, perturbation_scale=1.0, update_random_field=True):
        """perturbSim for every member (OceanStateNoise.py:376-496) followed
        by the periodic ghost exchange, all on the device."""
        if self.noise is None:
            raise RuntimeError("enable_model_error() first")
        st = G.stream_handle(self.gpu_stream)
        G.check(self.lib.gocean_noise_batch_perturb(
            self.noise, self.state[0].data_ptr(), self.Hi.data_ptr(), None, float(np.float32(q0_scale)),
            float(np.float32(perturbation_scale)), int(bool(update_random_field)),
            G.ptr(self.noise_seed), self.noise_random.data_ptr(), self.noise_coarse.data_ptr(),
            self.noise_key, self.noise_counter.data_ptr(), st), "gocean_noise_batch_perturb")
        G.check(self.lib.gocean_ensemble_periodic_bc(self.handle, self.state[0].data_ptr(), st),
                "ensemble periodic bc")

    # ------------------------------------------------------------------
    # drifters (GPUDrifterCollection per member in batched launches)
    # ------------------------------------------------------------------
    def enable_drifters(self, num_drifters, positions=None, sensitivity=1.0, wind_drift_factor=0.0):
        """``num_drifters`` drifters (plus the observation slot) per member,
        moved after every step by one batched launch (drifterStep,
        CDKLM16.py:600-610, for every member).  positions: (n, num+1, 2)
        metres; default: the centres of a near-square grid of sub-domains."""
        torch = Common._torch()
        num = int(num_drifters)
        arr = (G.DriftMember * self.n)()
        for k in range(self.n):
            arr[k].domain = self.members_c[k].domain
            arr[k].state_offset = int(self.state_off[k])
            arr[k].hm_offset = int(self.hm_off[k])
            arr[k].positions_offset = k * (num + 1) * 2
            arr[k].observations_offset = k * num * 2
            arr[k].num_drifters = num
            arr[k].sensitivity = float(np.float32(sensitivity))
            arr[k].wind_drift_factor = float(np.float32(wind_drift_factor))
            arr[k].wind = self.members_c[k].wind
        h = ctypes.c_void_p()
        G.check(self.lib.gocean_drift_batch_create(arr, self.n, ctypes.byref(h)), "gocean_drift_batch_create")
        self.drifters = h
        self.drift_members = arr
        self.num_drifters = num
        if positions is None:
            positions = np.zeros((self.n, num + 1, 2))
            rows = max(1, int(np.round(np.sqrt(num))))
            cols = int(np.ceil(num / rows)) if num else 1
            kk = np.arange(num)
            for k in range(self.n):
                dx, dy = self.dxy[k]
                positions[k, :num, 0] = (kk % cols + 0.5) * self.nx[k] * dx / cols
                positions[k, :num, 1] = (kk // cols + 0.5) * self.ny[k] * dy / rows
        dev = self.gpu_ctx.torch_device
        with torch.cuda.stream(self.gpu_stream):
            self.drifter_positions = torch.from_numpy(
                np.ascontiguousarray(positions, dtype=np.float32).reshape(self.n, num + 1, 2)).to(dev)
            self.drifter_observations = torch.zeros((self.n, max(num, 1), 2), dtype=torch.float32, device=dev)

    def drift(self):
        """One drift step of every member's drifters with its own device dt."""
        G.check(self.lib.gocean_drift_batch_step(self.drifters, self.state[0].data_ptr(), self.Hm.data_ptr(),
                                                 self.drifter_positions.data_ptr(), self.dt.data_ptr(),
                                                 G.stream_handle(self.gpu_stream)), "gocean_drift_batch_step")

    def observe(self, positions=None):
        """(n, num, 2) device tensor of hu, hv observed by every member:
        at ``positions`` ((num, 2) or (num+1, 2) device tensor, e.g. the true
        state's drifters) when given, else at each member's own drifters
        (observeUnderlyingFlow per particle, BaseOceanStateEnsemble.py:105-136)."""
        shared = positions is not None
        pos = positions if shared else self.drifter_positions
        G.check(self.lib.gocean_observe_flow_batch(self.drifters, self.state[0].data_ptr(), pos.data_ptr(),
                                                   int(shared), self.drifter_observations.data_ptr(),
                                                   G.stream_handle(self.gpu_stream)), "gocean_observe_flow_batch")
        return self.drifter_observations

    def noise_views(self, k):
        """(random field, coarse buffer) device views of member k."""
        d = self.noise_members[k].noise
        f = d.interpolation_factor
        cnx, cny = d.nx // f, d.ny // f
        gx = 0 if d.periodic_ew else 4
        gy = 0 if d.periodic_ns else 4
        rnx, rny = cnx + 2 * gx, cny + 2 * gy
        so, ro, co = (int(v) for v in self.noise_offsets[k])
        return (self.noise_random[ro:ro + rnx * rny].view(rny, rnx),
                self.noise_coarse[co:co + (cnx + 4) * (cny + 4)].view(cny + 4, cnx + 4))

    def capture(self, n_steps, model_error=False):
        """Capture ``n_steps`` RK steps of all members in a HIP graph (with
        ``model_error``, each step followed by the batched perturbation; the
        generator state lives on the device, so every replay draws new
        numbers; with drifters enabled, each step also moves them)."""
        if self.rk_order == 1 and (int(n_steps) % 2 or model_error or self.drifters is not None):
            raise ValueError("an rk_order 1 ensemble graph needs an even n_steps, without per-step "
                             "model error or drifters")
        self.gpu_stream.synchronize()
        self.graph_steps = int(n_steps)
        if not model_error and self.drifters is None:
            G.check(self.lib.gocean_ensemble_capture(
                self.handle, int(n_steps), self.state[0].data_ptr(), self.state[1].data_ptr(), self._ptr2(),
                self.Hi.data_ptr(), None, self.dt.data_ptr()), "gocean_ensemble_capture")
            self.torch_graph = None
            return
        torch = Common._torch()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.gpu_stream):
            for _ in range(int(n_steps)):
                self._rk(1)
                if model_error:
                    self.perturb()
                if self.drifters is not None:
                    self.drift()
        self.torch_graph = g

    def replay(self):
        if getattr(self, "torch_graph", None) is not None:
            torch = Common._torch()
            with torch.cuda.stream(self.gpu_stream):
                self.torch_graph.replay()
        else:
            G.check(self.lib.gocean_ensemble_replay(self.handle, G.stream_handle(self.gpu_stream)),
                    "gocean_ensemble_replay")
        self.num_iterations += self.graph_steps

    def set_tile_config(self, tile_config):
        """Switch the batched substep's tile configuration (an index into
        ``autotune.tile_config_names()``); a captured graph must be captured
        again."""
        G.check(self.lib.gocean_ensemble_set_tile_config(self.handle, int(tile_config)),
                "gocean_ensemble_set_tile_config")
        self.tile_config = int(tile_config)
        self.torch_graph = None

    def cells(self):
        return int(sum(nx * ny for nx, ny in zip(self.nx, self.ny)))

    # ------------------------------------------------------------------
    # the getters DA code calls on the reference's ensembles
    # (BaseOceanStateEnsemble.py:192-296); members are all active here
    # ------------------------------------------------------------------
    @property
    def particlesActive(self):
        return [True] * self.n

    def getNumParticles(self):
        return np.int32(self.n)

    def getNumActiveParticles(self):
        return self.n

    def getNumDrifters(self):
        return np.int32(self.num_drifters if self.drifters is not None else 0)

    def getNx(self):
        return np.int32(self.nx[0])

    def getNy(self):
        return np.int32(self.ny[0])

    def getDx(self):
        return np.float32(self.dxy[0][0])

    def getDy(self):
        return np.float32(self.dxy[0][1])

    def getDt(self):
        return np.float32(self.dt_host()[0])

    def getDomainSizeX(self):
        return self.getNx() * self.getDx()

    def getDomainSizeY(self):
        return self.getNy() * self.getDy()

    def downloadParticleOceanState(self, particleNo):
        return self.download(particleNo)

    def getInnovations(self, observations):
        """(N, N_d, 2) y - H(x_m): the observed (hu, hv) of the true state
        minus every member's at the same positions (getInnovations /
        observeParticles(innovation=True), BaseOceanStateEnsemble.py:80-205).
        observations: (N_d, 4) [x, y, hu, hv]."""
        torch = Common._torch()
        obs = np.asarray(observations, dtype=np.float64).reshape(-1, 4)
        with torch.cuda.stream(self.gpu_stream):
            pos = torch.from_numpy(obs[:, :2].astype(np.float32)).to(self.state[0].device)
            HX = self.observe(pos)[:, :obs.shape[0], :].to("cpu").numpy().astype(np.float64)
        return obs[None, :, 2:4] - HX

    def getGaussianWeight(self, innovations, observation_cov, normalize=True, R_scale=1.0):
        """Particle weights from Gaussian observation errors
        (BaseOceanStateEnsemble.getGaussianWeight, :216-261)."""
        return gaussian_weights(innovations, observation_cov, normalize, R_scale)

    def cleanUp(self):
        if getattr(self, "handle", None) is not None:
            self.lib.gocean_ensemble_destroy(self.handle)
            self.handle = None
        if getattr(self, "noise", None) is not None:
            self.lib.gocean_noise_batch_destroy(self.noise)
            self.noise = None
        if getattr(self, "drifters", None) is not None:
            self.lib.gocean_drift_batch_destroy(self.drifters)
            self.drifters = None

    def __del__(self):
        try:
            self.cleanUp()
        except Exception:
            pass


def gaussian_weights(innovations, observation_cov, normalize=True, R_scale=1.0):
    """log w_p = -1/2 sum_d d^T (R R_scale)^-1 d per particle; normalised with
    the max-shift (exp-normalise) when asked, else scaled by the Gaussian
    density's constant ((2 pi)^N_d det(R)^(N_d/2)) (BaseOceanStateEnsemble.
    getGaussianWeight, :216-261).  innovations: (N, N_d, 2) or (N,) scalars
    with observation variance R[0, 0]."""
    inn = np.asarray(innovations, dtype=np.float64)
    R = np.asarray(observation_cov, dtype=np.float64).reshape(2, 2) * R_scale
    if inn.ndim == 1:
        var = R[0, 0]
        logw = -(inn ** 2 / (2 * var))
        if normalize:
            return np.exp(logw - logw.max()) / np.exp(logw - logw.max()).sum()
        return (1.0 / np.sqrt(2 * np.pi * var)) * np.exp(logw)
    Rinv = np.linalg.inv(R)
    logw = -0.5 * np.einsum("pdi,ij,pdj->p", inn, Rinv, inn)
    if normalize:
        return np.exp(logw - logw.max()) / np.exp(logw - logw.max()).sum()
    nd = inn.shape[1]
    return (1.0 / ((2 * np.pi) ** nd * np.linalg.det(R) ** (nd / 2.0))) * np.exp(logw)
