"""Base class of the simulators (gpu_ocean/SWESimulators/Simulator.py:43-349):
parameters, stream, wind-stress time interpolation, download/upload/copy."""

from __future__ import annotations

import logging
from abc import ABCMeta, abstractmethod

import numpy as np


class Simulator(object):
    __metaclass__ = ABCMeta

    def __init__(self, gpu_ctx, nx, ny, ghost_cells_x, ghost_cells_y, dx, dy, dt,
                 g, f, r, A, t, theta, rk_order, coriolis_beta, y_zero_reference_cell,
                 wind_stress, write_netcdf, ignore_ghostcells, offset_x, offset_y,
                 comm, block_width, block_height, local_particle_id=0):
        self.gpu_stream = gpu_ctx.new_stream()
        self.logger = logging.getLogger(__name__)
        # same dtypes as the reference keeps (Simulator.py:78-91)
        self.gpu_ctx = gpu_ctx
        self.nx = np.int32(nx)
        self.ny = np.int32(ny)
        self.ghost_cells_x = np.int32(ghost_cells_x)
        self.ghost_cells_y = np.int32(ghost_cells_y)
        self.dx = np.float32(dx)
        self.dy = np.float32(dy)
        self.dt = dt
        self.g = np.float32(g)
        self.f = np.float32(f) if np.ndim(f) == 0 else np.asarray(f, dtype=np.float32)
        self.r = np.float32(r)
        self.coriolis_beta = np.float32(coriolis_beta)
        self.wind_stress = wind_stress
        self.y_zero_reference_cell = np.float32(y_zero_reference_cell)
        self.offset_x = offset_x
        self.offset_y = offset_y
        self.t = t
        self.num_iterations = 0
        self.A = 'NA' if A is None else np.float32(A)
        self.theta = 'NA' if theta is None else np.float32(theta)
        self.rk_order = 'NA' if rk_order is None else np.int32(rk_order)
        self.hasDrifters = False
        self.drifters = None
        self.write_netcdf = write_netcdf
        self.ignore_ghostcells = ignore_ghostcells
        self.sim_writer = None
        self.comm = comm
        self.local_particle_id = local_particle_id
        self.local_size = (block_width, block_height, 1)
        self.global_size = (int(np.ceil(self.nx / float(self.local_size[0]))),
                            int(np.ceil(self.ny / float(self.local_size[1]))))

    def wind_stress_interpolation(self):
        """Bracketing wind fields and the time weight for self.t
        (Simulator.update_wind_stress, Simulator.py:145-223)."""
        ws = self.wind_stress
        t_max_index = len(ws.t) - 1
        t0_index = max(0, int(np.searchsorted(ws.t, self.t)) - 1)
        t1_index = min(t_max_index, int(np.searchsorted(ws.t, self.t)))
        new_t0 = ws.t[t0_index]
        new_t1 = ws.t[t1_index]
        elapsed_since_t0 = (self.t - new_t0)
        time_interval = max(1.0e-10, (new_t1 - new_t0))
        wind_stress_t = max(0.0, min(1.0, elapsed_since_t0 / time_interval))
        return t0_index, t1_index, wind_stress_t

    @abstractmethod
    def step(self, t_end=0.0):
        pass

    def __del__(self):
        try:
            self.cleanUp()
        except Exception:
            pass

    @abstractmethod
    def cleanUp(self):
        pass

    def closeNetCDF(self):
        if self.write_netcdf and self.sim_writer is not None:
            self.sim_writer.__exit__(0, 0, 0)
            self.write_netcdf = False

    def attachDrifters(self, drifters):
        self.drifters = drifters
        self.hasDrifters = True
        if hasattr(self.drifters, "setGPUStream"):
            self.drifters.setGPUStream(self.gpu_stream)
        self.drifter_t = 0.0

    def download(self, interior_domain_only=False):
        """Latest state (eta, hu, hv) as host arrays (Simulator.py:272-285)."""
        eta, hu, hv = self.gpu_data.download(self.gpu_stream)
        if interior_domain_only:
            s = (slice(self.interior_domain_indices[2], self.interior_domain_indices[0]),
                 slice(self.interior_domain_indices[3], self.interior_domain_indices[1]))
            return eta[s], hu[s], hv[s]
        return eta, hu, hv

    def downloadPrevTimestep(self):
        return self.gpu_data.downloadPrevTimestep(self.gpu_stream)

    def copyState(self, otherSim):
        """Copy the ocean state and wind of another simulator
        (Simulator.py:294-322)."""
        assert type(otherSim) is type(self), \
            "A simulator can only copy the state from another simulator of the same class."
        assert (self.ny, self.nx) == (otherSim.ny, otherSim.nx), \
            "Simulators differ in computational domain."
        for a, b in ((self.gpu_data.h0, otherSim.gpu_data.h0), (self.gpu_data.hu0, otherSim.gpu_data.hu0),
                     (self.gpu_data.hv0, otherSim.gpu_data.hv0), (self.gpu_data.h1, otherSim.gpu_data.h1),
                     (self.gpu_data.hu1, otherSim.gpu_data.hu1), (self.gpu_data.hv1, otherSim.gpu_data.hv1)):
            a.copyBuffer(self.gpu_stream, b)
        self.wind_stress = otherSim.wind_stress
        if otherSim.hasDrifters and self.hasDrifters:
            self.drifters.setDrifterPositions(otherSim.drifters.getDrifterPositions())
            self.drifters.setObservationPosition(otherSim.drifters.getObservationPosition())

    def upload(self, eta0, hu0, hv0, eta1=None, hu1=None, hv1=None):
        """Re-initialise the ocean state (Simulator.py:326-341)."""
        self.gpu_data.h0.upload(self.gpu_stream, eta0)
        self.gpu_data.hu0.upload(self.gpu_stream, hu0)
        self.gpu_data.hv0.upload(self.gpu_stream, hv0)
        if eta1 is None:
            self.gpu_data.h1.upload(self.gpu_stream, eta0)
            self.gpu_data.hu1.upload(self.gpu_stream, hu0)
            self.gpu_data.hv1.upload(self.gpu_stream, hv0)
        else:
            self.gpu_data.h1.upload(self.gpu_stream, eta1)
            self.gpu_data.hu1.upload(self.gpu_stream, hu1)
            self.gpu_data.hv1.upload(self.gpu_stream, hv1)
