"""CDKLM16 simulator on MI355X (gfx950).

Keeps the constructor, ``step``, ``download``, ``updateDt``,
``dataAssimilationStep`` surface of the reference class
(gpu_ocean/SWESimulators/CDKLM16.py:44-794) so that ensemble, drifter and
data-assimilation code can use it unchanged.  Each RK substep is one launch of
the HIP kernel in ``lib/libgocean_hip.so`` (``gocean_cdklm_substep``); wall and
periodic boundaries are applied inside that kernel, so the reference's
per-substep ghost-cell launches disappear.

Differences from the reference, by design:
  * Coriolis/angle fields are used at full resolution (evaluated at cell
    centres) instead of a subsampled, bilinearly filtered texture;
  * runs of equal time steps are issued with one library call
    (``gocean_rk_steps``); ``step_n`` and ``capture_graph`` expose fixed-dt
    stepping without any host synchronisation;
  * open boundaries (3, 4) are not part of this build and raise
    ``NotImplementedError``; gridded wind is read with exact bilinear weights
    where the reference's texture unit rounds them to 8 fractional bits;
    NetCDF output is NetCDF-3 (scipy.io) since netCDF4 is not available.
"""

from __future__ import annotations

import gc
import logging

import numpy as np

from . import Common, OceanStateNoise, SimReader, SimWriter, Simulator, WindStress
from . import _gocean as G


def calcCoriolisParams(lat):
    """f and beta at latitude ``lat`` in radians
    (OceanographicUtilities.py:205-219)."""
    omega = 7.2921 * 1.0e-5
    f = 2.0 * omega * np.sin(lat)
    a = 6371008.8
    beta = 2.0 * omega * np.cos(lat) / a
    return [f, beta]


class CDKLM16(Simulator.Simulator):
    """Well-balanced central-upwind scheme with Coriolis (Chertock, Dudzinski,
    Kurganov & Lukacova-Medvidova 2016) for the shallow-water equations."""

    def __init__(self, gpu_ctx, eta0, hu0, hv0, H, nx, ny, dx, dy, dt, g, f, r,
                 subsample_f=10, angle=np.array([[0]], dtype=np.float32),
                 subsample_angle=10, latitude=None, t=0.0, theta=1.3, rk_order=2,
                 coriolis_beta=0.0, max_wind_direction_perturbation=0,
                 wind_stress=WindStress.WindStress(),
                 boundary_conditions=Common.BoundaryConditions(),
                 boundary_conditions_data=Common.BoundaryConditionsData(),
                 small_scale_perturbation=False, small_scale_perturbation_amplitude=None,
                 small_scale_perturbation_interpolation_factor=1, model_time_step=None,
                 reportGeostrophicEquilibrium=False, use_lcg=False, xorwow_seed=None,
                 write_netcdf=False, comm=None, local_particle_id=0, super_dir_name=None,
                 netcdf_filename=None, ignore_ghostcells=False, courant_number=0.8,
                 offset_x=0, offset_y=0, flux_slope_eps=1.0e-1, desingularization_eps=1.0e-1,
                 depth_cutoff=1.0e-5, block_width=12, block_height=32, num_threads_dt=256,
                 block_width_model_error=16, block_height_model_error=16,
                 tile_config=None):
        self.logger = logging.getLogger(__name__)
        assert rk_order in (1, 2, 3), "Only 1st, 2nd and 3rd order Runge Kutta supported"
        if rk_order == 3:
            assert r == 0.0, "3rd order Runge Kutta supported only without friction"
        bc_codes = boundary_conditions.codes()
        if any(b not in (Common.WALL, Common.PERIODIC) for b in bc_codes):
            raise NotImplementedError("open boundary conditions (3: flow relaxation, "
                                      "4: linear interpolation) are not part of this build")
        if max_wind_direction_perturbation > 0:
            raise NotImplementedError("wind direction perturbation needs the legacy wind classes")
        if not wind_stress.is_uniform():
            shapes = {np.asarray(a).shape for a in list(wind_stress.X) + list(wind_stress.Y)}
            if len(shapes) != 1:
                raise ValueError("gridded wind stress: every X and Y field must have the same shape, got "
                                 + str(sorted(shapes)))

        ghost_cells_x = 2
        ghost_cells_y = 2
        y_zero_reference_cell = ghost_cells_y
        self.boundary_conditions = boundary_conditions
        angle = np.asarray(angle)
        # f compensation for the reference cell (CDKLM16.py:146-148)
        north = np.array([np.sin(angle.flat[0]), np.cos(angle.flat[0])])
        if np.ndim(f) == 0:
            f = f - coriolis_beta * (ghost_cells_x * dx * north[0] + ghost_cells_y * dy * north[1])
        super(CDKLM16, self).__init__(gpu_ctx, nx, ny, ghost_cells_x, ghost_cells_y, dx, dy, dt,
                                      g, f, r, None, t, theta, rk_order, coriolis_beta, 0,
                                      wind_stress, write_netcdf, ignore_ghostcells,
                                      offset_x, offset_y, comm, block_width, block_height,
                                      local_particle_id=local_particle_id)
        self.interior_domain_indices = np.array([-2, -2, 2, 2])
        self.flux_slope_eps = flux_slope_eps
        self.desingularization_eps = desingularization_eps
        self.depth_cutoff = depth_cutoff
        self.tile_config = -1 if tile_config is None else int(tile_config)
        self.lib = G.load()

        # bathymetry (CDKLM16.py:232-238)
        self.bathymetry = Common.Bathymetry(gpu_ctx, self.gpu_stream, nx, ny,
                                            ghost_cells_x, ghost_cells_y, H, boundary_conditions)
        Hm = self.downloadBathymetry()[1]
        eta0 = np.maximum(eta0, -Hm)
        self.gpu_data = Common.SWEDataArakawaA(self.gpu_stream, nx, ny, ghost_cells_x,
                                               ghost_cells_y, eta0, hu0, hv0,
                                               device=gpu_ctx.torch_device)
        self.courant_number = courant_number
        self.constant_equilibrium_depth = np.max(H)
        self.reportGeostrophicEquilibrium = np.int32(reportGeostrophicEquilibrium)
        self.geoEq_uxpvy = self.geoEq_Kx = self.geoEq_Ly = None
        self.num_threads_dt = num_threads_dt

        # Coriolis parameter and north vector (CDKLM16.py:274-351): constant,
        # or per-cell fields evaluated at cell centres (no subsampling)
        shape = (int(ny) + 4, int(nx) + 4)
        coriolis_f = None
        if latitude is not None:
            if self.f != 0.0:
                raise RuntimeError("Cannot specify both latitude and f. Make your mind up.")
            coriolis_f, _ = calcCoriolisParams(latitude)
            coriolis_f = np.asarray(coriolis_f, dtype=np.float32)
        elif self.coriolis_beta != 0.0:
            if angle.size != 1:
                raise RuntimeError("non-constant angle cannot be combined with beta plane model (makes no sense)")
            x = np.linspace((-ghost_cells_x + 0.5) * self.dx, (nx + ghost_cells_x - 0.5) * self.dx, nx + 2 * ghost_cells_x)
            y = np.linspace((-ghost_cells_y + 0.5) * self.dy, (ny + ghost_cells_y - 0.5) * self.dy, ny + 2 * ghost_cells_y)
            x, y = np.meshgrid(x, y)
            n = x * np.sin(angle.flat[0]) + y * np.cos(angle.flat[0])
            coriolis_f = (self.f + self.coriolis_beta * n).astype(np.float32)
        elif np.ndim(self.f) != 0 and np.asarray(self.f).size > 1:
            if np.asarray(self.f).shape != shape:
                raise RuntimeError("The shape of f should match up with eta or be scalar.")
            coriolis_f = np.asarray(self.f, dtype=np.float32)
        self._fields_host = None
        if coriolis_f is not None and np.asarray(coriolis_f).size > 1 or angle.size > 1:
            fF = np.broadcast_to(np.asarray(coriolis_f if coriolis_f is not None else self.f,
                                            dtype=np.float32), shape)
            if angle.size > 1:
                assert angle.shape == shape, "angle must be scalar or have the shape of eta"
                ang = angle.astype(np.float32).astype(np.float64)
                nxF = np.sin(ang).astype(np.float32)
                nyF = np.cos(ang).astype(np.float32)
            else:
                a32 = float(np.float32(angle.flat[0]))
                nxF = np.full(shape, np.float32(np.sin(a32)), dtype=np.float32)
                nyF = np.full(shape, np.float32(np.cos(a32)), dtype=np.float32)
            self._fields_host = (np.ascontiguousarray(fF), nxF, nyF)
            self.f_const = 0.0
        else:
            self.f_const = float(np.float32(coriolis_f if coriolis_f is not None else self.f))
        a32 = float(np.float32(angle.flat[0]))
        self.north_const = (float(np.float32(np.sin(a32))), float(np.float32(np.cos(a32))))
        self._make_device_descriptors()

        # model error (CDKLM16.py:352-362)
        self.small_scale_perturbation = small_scale_perturbation
        self.small_scale_model_error = None
        self.small_scale_perturbation_interpolation_factor = small_scale_perturbation_interpolation_factor
        if small_scale_perturbation:
            self.small_scale_model_error = OceanStateNoise.OceanStateNoise.fromsim(
                self, soar_q0=small_scale_perturbation_amplitude,
                interpolation_factor=small_scale_perturbation_interpolation_factor,
                use_lcg=use_lcg, xorwow_seed=xorwow_seed,
                block_width=block_width_model_error, block_height=block_height_model_error)
        self.model_time_step = model_time_step
        self.total_time_steps = 0
        if model_time_step is None:
            self.model_time_step = self.dt
        if self.write_netcdf:
            # NetCDF output after each step() (CDKLM16.py:375-377)
            self.sim_writer = SimWriter.SimNetCDFWriter(self, super_dir_name=super_dir_name,
                                                        filename=netcdf_filename,
                                                        ignore_ghostcells=self.ignore_ghostcells,
                                                        offset_x=self.offset_x, offset_y=self.offset_y)
        if self.dt <= 0:
            self.updateDt()

    # ------------------------------------------------------------------
    def _make_device_descriptors(self):
        torch = Common._torch()
        dev = self.gpu_ctx.torch_device
        self.domain = Common.make_domain(
            self.nx, self.ny, self.dx, self.dy, self.g, self.r, self.theta, self.rk_order,
            self.boundary_conditions.codes(), pitch=self.gpu_data.h0.pitch_elems,
            hi_pitch=self.bathymetry.Bi.pitch_elems, f=self.f_const,
            north_x=self.north_const[0], north_y=self.north_const[1],
            desing_eps=self.desingularization_eps, flux_slope_eps=self.flux_slope_eps,
            depth_cutoff=self.depth_cutoff, land_value=self.bathymetry.mask_value)
        self.fields = G.Fields()
        self.field_tensors = None
        if self._fields_host is not None:
            with torch.cuda.stream(self.gpu_stream):
                self.field_tensors = tuple(torch.from_numpy(a).to(dev) for a in self._fields_host)
            self.fields.f, self.fields.north_x, self.fields.north_y = (t.data_ptr() for t in self.field_tensors)
        with torch.cuda.stream(self.gpu_stream):
            # dt reduction scratch: [min, courant*min, partial minima...]
            self.dt_scratch = torch.zeros(2 + 1024, dtype=torch.float32, device=dev)
            self.h2 = None
        # texture handles of the reference; the fields above replace them
        self.angle_texref = self.field_tensors[1:] if self.field_tensors else None
        self.coriolis_texref = self.field_tensors[0] if self.field_tensors else None

    def _stream(self):
        return G.stream_handle(self.gpu_stream)

    def _wind(self):
        """The wind of the current time (Simulator.update_wind_stress,
        Simulator.py:145-223): uniform values, or for gridded fields the two
        bracketing samples in one device buffer (X0, X1, Y0, Y1), uploaded
        when the bracketing interval changes."""
        t0, t1, wt = self.wind_stress_interpolation()
        w = G.Wind()
        w.t = float(np.float32(wt))
        ws = self.wind_stress
        if ws.is_uniform():
            w.x0 = float(np.asarray(ws.X[t0], np.float32).flat[0])
            w.x1 = float(np.asarray(ws.X[t1], np.float32).flat[0])
            w.y0 = float(np.asarray(ws.Y[t0], np.float32).flat[0])
            w.y1 = float(np.asarray(ws.Y[t1], np.float32).flat[0])
            return w
        torch = Common._torch()
        shape = np.asarray(ws.X[0]).shape
        key = (id(ws), t0, t1)
        if getattr(self, "_wind_key", None) != key:
            host = np.stack([np.asarray(a, np.float32) for a in (ws.X[t0], ws.X[t1], ws.Y[t0], ws.Y[t1])])
            with torch.cuda.stream(self.gpu_stream):
                if getattr(self, "_wind_grid", None) is None or tuple(self._wind_grid.shape) != host.shape:
                    self._wind_grid = torch.empty(host.shape, dtype=torch.float32,
                                                  device=self.gpu_ctx.torch_device)
                self._wind_grid.copy_(torch.from_numpy(host))
            self._wind_key = key
        w.grid = self._wind_grid.data_ptr()
        w.grid_ny, w.grid_nx = int(shape[0]), int(shape[1])
        return w

    def _ensure_h2(self):
        if self.h2 is None:
            torch = Common._torch()
            with torch.cuda.stream(self.gpu_stream):
                self.h2 = tuple(a.tensor.clone() for a in (self.gpu_data.h1, self.gpu_data.hu1, self.gpu_data.hv1))
        return self.h2

    def _rk_steps(self, n, dt, wind, dt_dev=None):
        gd = self.gpu_data
        h2 = self._ensure_h2() if self.rk_order == 3 else (None, None, None)
        if self.rk_order == 1 and n % 2 == 1:
            # Euler swaps buffers each step (CDKLM16.py:555-560)
            G.check(self.lib.gocean_cdklm_substep(
                self.domain, float(dt), G.ptr(dt_dev), 0,
                gd.h0.data.gpudata, gd.hu0.data.gpudata, gd.hv0.data.gpudata, None, None, None,
                gd.h1.data.gpudata, gd.hu1.data.gpudata, gd.hv1.data.gpudata,
                self.bathymetry.Bi.data.gpudata,
                self.fields, wind, self.tile_config, self._stream()), "gocean_cdklm_substep")
            gd.swap()
            n -= 1
            if n == 0:
                return
        G.check(self.lib.gocean_rk_steps(
            self.domain, int(n), float(dt), G.ptr(dt_dev),
            gd.h0.data.gpudata, gd.hu0.data.gpudata, gd.hv0.data.gpudata,
            gd.h1.data.gpudata, gd.hu1.data.gpudata, gd.hv1.data.gpudata,
            G.ptr(h2[0]), G.ptr(h2[1]), G.ptr(h2[2]),
            self.bathymetry.Bi.data.gpudata, self.fields, wind, self.tile_config,
            self._stream()), "gocean_rk_steps")

    def _periodic_bc(self, eta, hu, hv):
        G.check(self.lib.gocean_periodic_bc(self.domain, eta.data.gpudata, hu.data.gpudata,
                                            hv.data.gpudata, self._stream()), "gocean_periodic_bc")

    # ------------------------------------------------------------------
    def cleanUp(self):
        if getattr(self, "sim_writer", None) is not None:
            self.closeNetCDF()
        if getattr(self, "gpu_data", None) is not None:
            self.gpu_data.release()
        if getattr(self, "bathymetry", None) is not None:
            self.bathymetry.release()
        self.field_tensors = None
        self.dt_scratch = None
        self.h2 = None
        self.gpu_ctx = None
        gc.collect()

    def closeNetCDF(self):
        if self.write_netcdf and self.sim_writer is not None:
            self.sim_writer.close()
            self.write_netcdf = False

    @classmethod
    def fromfilename(cls, gpu_ctx, filename, cont_write_netcdf=True, use_lcg=False,
                     xorwow_seed=None, new_netcdf_filename=None, time0=None):
        """Hot start from the last (or the ``time0``) record of a file written
        by SimNetCDFWriter, with its parameters (CDKLM16.py:410-490)."""
        reader = SimReader.SimNetCDFReader(filename, ignore_ghostcells=False)
        sim_name = str(reader.get('simulator_short'))
        assert sim_name == cls.__name__, \
            "Trying to initialize a " + cls.__name__ + " simulator with netCDF file based on " + sim_name + " results."
        H = reader.getH()
        if time0 is None:
            eta0, hu0, hv0, time0 = reader.getLastTimeStep()
        else:
            eta0, hu0, hv0, time0 = reader.getStateAtTime(time0)
        if len(H.shape) == 3:
            H = H[0, :, :]
        params = dict(gpu_ctx=gpu_ctx, eta0=eta0, hu0=hu0, hv0=hv0, H=H,
                      nx=reader.get("nx"), ny=reader.get("ny"), dx=reader.get("dx"), dy=reader.get("dy"),
                      dt=reader.get("dt"), g=reader.get("g"), f=reader.get("coriolis_force"),
                      r=reader.get("bottom_friction_r"), t=float(time0), theta=reader.get("minmod_theta"),
                      rk_order=reader.get("time_integrator"), coriolis_beta=reader.get("coriolis_beta"),
                      write_netcdf=cont_write_netcdf, use_lcg=use_lcg, xorwow_seed=xorwow_seed,
                      netcdf_filename=new_netcdf_filename, wind_stress=WindStress.WindStress(),
                      boundary_conditions=reader.getBC())
        if reader.has('small_scale_perturbation'):
            params['small_scale_perturbation'] = reader.get('small_scale_perturbation') == 'True'
            if params['small_scale_perturbation']:
                params['small_scale_perturbation_amplitude'] = reader.get('small_scale_perturbation_amplitude')
                params['small_scale_perturbation_interpolation_factor'] = reader.get(
                    'small_scale_perturbation_interpolation_factor')
        if reader.has('model_time_step'):
            params['model_time_step'] = reader.get('model_time_step')
        reader.close()
        return cls(**params)

    def step(self, t_end=0.0, apply_stochastic_term=True, write_now=True, update_dt=False):
        """Advance by ``t_end`` seconds (CDKLM16.py:494-598).  Same loop:
        ``local_dt = float32(min(dt, float32(t_end - t_now)))``, ``t``
        accumulated in float64, boundary conditions applied at t == 0."""
        if self.t == 0:
            self._periodic_bc(self.gpu_data.h0, self.gpu_data.hu0, self.gpu_data.hv0)
            self._periodic_bc(self.gpu_data.h1, self.gpu_data.hu1, self.gpu_data.hv1)
        perturb = self.small_scale_perturbation and apply_stochastic_term
        groupable = (not update_dt) and (not self.hasDrifters) and len(self.wind_stress.t) == 1 \
            and not perturb
        t_now = 0.0
        if groupable:
            # the loop's dt sequence does not depend on the state: issue each
            # run of equal local_dt as one library call
            dts = []
            while t_now < t_end:
                local_dt = np.float32(min(self.dt, np.float32(t_end - t_now)))
                dts.append(local_dt)
                t_now += np.float64(local_dt)
            wind = self._wind()
            k = 0
            while k < len(dts):
                m = k
                while m < len(dts) and dts[m] == dts[k]:
                    m += 1
                self._rk_steps(m - k, dts[k], wind)
                for q in range(k, m):
                    self.t += np.float64(dts[q])
                    self.num_iterations += 1
                k = m
            if self.write_netcdf and write_now:
                self.sim_writer.writeTimestep(self)
            return self.t
        while t_now < t_end:
            if update_dt:
                self.updateDt()
            local_dt = np.float32(min(self.dt, np.float32(t_end - t_now)))
            self._rk_steps(1, local_dt, self._wind())
            if perturb:
                # model error, then the ghost cells (CDKLM16.py:580-586)
                self.small_scale_model_error.perturbSim(self)
                self._periodic_bc(self.gpu_data.h0, self.gpu_data.hu0, self.gpu_data.hv0)
            self.drifterStep(local_dt)
            self.t += np.float64(local_dt)
            t_now += np.float64(local_dt)
            self.num_iterations += 1
        # NetCDF record after the step (CDKLM16.py:595-596)
        if self.write_netcdf and write_now:
            self.sim_writer.writeTimestep(self)
        return self.t

    def step_n(self, n_steps, dt=None, dt_dev=None):
        """``n_steps`` fixed-dt RK steps with one library call and no host
        synchronisation.  ``dt_dev`` (a device float tensor) overrides dt."""
        d = np.float32(self.dt if dt is None else dt)
        self._rk_steps(int(n_steps), d, self._wind(), dt_dev=dt_dev)
        self.t += float(n_steps) * np.float64(d)
        self.num_iterations += int(n_steps)
        return self.t

    def capture_graph(self, n_steps, dt=None):
        """Capture ``n_steps`` fixed-dt steps in a HIP graph (torch.cuda.CUDAGraph)
        and return it; ``graph.replay()`` advances the state without host work.
        The simulator's time is not advanced by replays."""
        torch = Common._torch()
        d = np.float32(self.dt if dt is None else dt)
        wind = self._wind()
        self._ensure_h2() if self.rk_order == 3 else None
        if self.rk_order == 1 and n_steps % 2:
            raise ValueError("rk_order 1 graphs need an even number of steps")
        graph = torch.cuda.CUDAGraph()
        self.gpu_stream.synchronize()
        with torch.cuda.graph(graph, stream=self.gpu_stream):
            self._rk_steps(int(n_steps), d, wind)
        return graph

    def drifterStep(self, dt):
        if self.hasDrifters:
            self.drifters.drift(self.gpu_data.h0, self.gpu_data.hu0, self.gpu_data.hv0,
                                self.bathymetry.Bm, self.nx, self.ny, self.t, self.dx, self.dy,
                                dt, np.int32(2), np.int32(2))
            self.drifter_t += dt
            return self.drifter_t

    def perturbState(self, q0_scale=1):
        if self.small_scale_perturbation:
            self.small_scale_model_error.perturbSim(self, q0_scale=q0_scale)
            self._periodic_bc(self.gpu_data.h0, self.gpu_data.hu0, self.gpu_data.hv0)

    def applyBoundaryConditions(self):
        self._periodic_bc(self.gpu_data.h0, self.gpu_data.hu0, self.gpu_data.hv0)

    def dataAssimilationStep(self, observation_time, model_error_final_step=True,
                             write_now=True, courant_number=0.8):
        """Step to ``observation_time`` in model time steps, leftover first
        (CDKLM16.py:650-704)."""
        full_model_time_steps = int(round(observation_time - self.t) / self.model_time_step)
        leftover_step_size = observation_time - self.t - full_model_time_steps * self.model_time_step
        if leftover_step_size / self.model_time_step < 0.1 and full_model_time_steps > 1:
            leftover_step_size += self.model_time_step
            full_model_time_steps -= 1
        if leftover_step_size / self.model_time_step < 0.00001:
            leftover_step_size = 0
        assert full_model_time_steps > 0, "There is less than CDKLM16.model_time_step until the observation"
        self.updateDt(courant_number=courant_number)
        for i in range(full_model_time_steps + 1):
            if i == 0 and leftover_step_size == 0:
                continue
            elif i == 0:
                self.step(leftover_step_size, apply_stochastic_term=False, write_now=False)
                self.perturbState(q0_scale=np.sqrt(leftover_step_size / self.model_time_step))
            else:
                self.step(self.model_time_step, apply_stochastic_term=False, write_now=False)
                if (i < full_model_time_steps) or model_error_final_step:
                    self.perturbState()
            self.total_time_steps += 1
            if self.total_time_steps % 5 == 0:
                self.updateDt(courant_number=courant_number)
        if self.write_netcdf and write_now:
            self.sim_writer.writeTimestep(self)
        assert round(observation_time) == round(self.t), \
            'The simulation time is not the same as observation time after dataAssimilationStep! \n' + \
            '(self.t, observation_time, diff): ' + str((self.t, observation_time, self.t - observation_time))

    def writeState(self):
        if self.write_netcdf:
            self.sim_writer.writeTimestep(self)

    def max_dt_device(self, courant_number=None):
        """Launch the CFL reduction; returns the scratch tensor whose [0] is
        the raw minimum and [1] courant*min (float32), without syncing."""
        c = self.courant_number if courant_number is None else courant_number
        gd = self.gpu_data
        G.check(self.lib.gocean_max_dt(self.domain, gd.h0.data.gpudata, gd.hu0.data.gpudata,
                                       gd.hv0.data.gpudata, self.bathymetry.Bm.data.gpudata,
                                       float(c), G.ptr(self.dt_scratch), self._stream()),
                "gocean_max_dt")
        return self.dt_scratch

    def updateDt(self, courant_number=None):
        """dt = courant * CFL minimum (CDKLM16.py:712-739).  The product is
        taken in double, as NumPy 1.x evaluated ``float * float32``."""
        if courant_number is None:
            courant_number = self.courant_number
        s = self.max_dt_device(courant_number)
        self.gpu_stream.synchronize()
        dt_host = np.float32(s[0].item())
        self.dt = float(courant_number) * float(dt_host)

    def _getMaxTimestepHost(self, courant_number=0.8):
        """Host closed form of the CFL step (CDKLM16.py:741-759)."""
        eta, hu, hv = self.download(interior_domain_only=True)
        Hm = self.downloadBathymetry()[1][2:-2, 2:-2]
        h = eta + Hm
        gravityWaves = np.sqrt(self.g * h)
        u = hu / h
        v = hv / h
        max_dt = 0.25 * float(min(self.dx / np.max(np.abs(u) + gravityWaves),
                                  self.dy / np.max(np.abs(v) + gravityWaves)))
        return float(courant_number) * max_dt

    def downloadBathymetry(self, interior_domain_only=False):
        Bi, Bm = self.bathymetry.download(self.gpu_stream)
        if interior_domain_only:
            Bi = Bi[self.interior_domain_indices[2]:self.interior_domain_indices[0] + 1,
                    self.interior_domain_indices[3]:self.interior_domain_indices[1] + 1]
            Bm = Bm[self.interior_domain_indices[2]:self.interior_domain_indices[0],
                    self.interior_domain_indices[3]:self.interior_domain_indices[1]]
        return [Bi, Bm]

    def getLandMask(self, interior_domain_only=True):
        if self.gpu_data.h0.mask is None:
            return None
        if interior_domain_only:
            return self.gpu_data.h0.mask[2:-2, 2:-2]
        return self.gpu_data.h0.mask

    def downloadDt(self):
        return self.dt_scratch.cpu().numpy()
