"""Drifters on the device (gpu_ocean/SWESimulators/GPUDrifterCollection.py:33-256).

Positions live in one ``(num + 1) x 2`` float32 device array (drifters, then
the observation position).  ``drift`` advances all of them with one call of
``gocean_drift`` in ``lib/libgocean_hip.so``; wind is the uniform wind of the
simulator's ``WindStress`` lerped in time, as the reference's 1x1 textures.
"""

from __future__ import annotations

import numpy as np

from . import BaseDrifterCollection, Common, WindStress
from . import _gocean as G


def _drift_bc(bc):
    """Wall/periodic codes for the drift kernel: periodic stays periodic,
    anything else is a closed side."""
    return tuple(Common.PERIODIC if b == Common.PERIODIC else Common.WALL for b in bc.codes())


class GPUDrifterCollection(BaseDrifterCollection.BaseDrifterCollection):

    def __init__(self, gpu_ctx, numDrifters, observation_variance=0.01,
                 wind=WindStress.WindStress(), wind_drift_factor=0.0,
                 boundaryConditions=Common.BoundaryConditions(),
                 initialization_cov_drifters=None, domain_size_x=1.0, domain_size_y=1.0,
                 gpu_stream=None, block_width=64):
        super(GPUDrifterCollection, self).__init__(numDrifters, observation_variance=observation_variance,
                                                   boundaryConditions=boundaryConditions,
                                                   domain_size_x=domain_size_x, domain_size_y=domain_size_y)
        torch = Common._torch()
        self.lib = G.load()
        self.gpu_ctx = gpu_ctx
        self.block_width = block_width
        self.block_height = 1
        self.wind = wind
        self.wind_drift_factor = np.float32(wind_drift_factor)
        if self.wind_drift_factor and not wind.is_uniform():
            raise NotImplementedError("wind drift needs uniform (1x1) wind fields")
        self.gpu_stream = gpu_stream
        if self.gpu_stream is None:
            self.gpu_stream = gpu_ctx.new_stream() if gpu_ctx is not None else torch.cuda.Stream()
        self.sensitivity = 1.0
        self.driftersHost = np.zeros((self.getNumDrifters() + 1, 2), dtype=np.float32)
        dev = gpu_ctx.torch_device if gpu_ctx is not None else None
        self.driftersDevice = Common.DeviceArray2D(self.gpu_stream, 2, self.getNumDrifters() + 1, 0, 0,
                                                   self.driftersHost, device=dev)
        self.uniformly_distribute_drifters(initialization_cov_drifters=initialization_cov_drifters)

    def copy(self):
        c = GPUDrifterCollection(self.gpu_ctx, self.getNumDrifters(),
                                 observation_variance=self.observation_variance,
                                 boundaryConditions=self.boundaryConditions,
                                 domain_size_x=self.domain_size_x, domain_size_y=self.domain_size_y,
                                 gpu_stream=self.gpu_stream, block_width=self.block_width)
        c.setDrifterPositions(self.getDrifterPositions())
        c.setObservationPosition(self.getObservationPosition())
        return c

    def update_wind(self, t):
        """(wind_t, gocean_wind_t) for time t: the bracketing wind fields and
        the clamped interpolation weight (GPUDrifterCollection.py:116-176)."""
        tt = self.wind.t
        t_max = len(tt) - 1
        i0 = max(0, int(np.searchsorted(tt, t)) - 1)
        i1 = min(t_max, int(np.searchsorted(tt, t)))
        interval = max(1.0e-10, tt[i1] - tt[i0])
        wind_t = max(0.0, min(1.0, (t - tt[i0]) / interval))
        w = G.Wind()
        w.t = float(np.float32(wind_t))
        w.x0 = float(np.asarray(self.wind.X[i0], np.float32).flat[0])
        w.x1 = float(np.asarray(self.wind.X[i1], np.float32).flat[0])
        w.y0 = float(np.asarray(self.wind.Y[i0], np.float32).flat[0])
        w.y1 = float(np.asarray(self.wind.Y[i1], np.float32).flat[0])
        return wind_t, w

    def setDrifterPositions(self, newDrifterPositions):
        all_ = np.concatenate((np.asarray(newDrifterPositions, dtype=np.float64).reshape(-1, 2),
                               np.array([self.getObservationPosition()])), axis=0)
        self.driftersDevice.upload(self.gpu_stream, all_)

    def setObservationPosition(self, newObservationPosition):
        all_ = np.concatenate((self.getDrifterPositions(),
                               np.asarray(newObservationPosition, dtype=np.float64).reshape(1, 2)))
        self.driftersDevice.upload(self.gpu_stream, all_)

    def setSensitivity(self, sensitivity):
        self.sensitivity = sensitivity

    def getDrifterPositions(self):
        return self.driftersDevice.download(self.gpu_stream)[:-1, :]

    def getObservationPosition(self):
        return self.driftersDevice.download(self.gpu_stream)[self.obs_index, :]

    def _domain(self, nx, ny, dx, dy, pitch):
        return Common.make_domain(nx, ny, dx, dy, 9.81, 0.0, 1.3, 2, _drift_bc(self.boundaryConditions),
                                  pitch=pitch)

    def drift(self, eta, hu, hv, Hm, nx, ny, t, dx, dy, dt, x_zero_ref, y_zero_ref):
        """Move every position (observation included) by the interpolated
        velocity over dt (GPUDrifterCollection.py:212-232)."""
        wind = None
        if self.wind_drift_factor:
            _, wind = self.update_wind(t)
        d = self._domain(nx, ny, dx, dy, eta.pitch_elems)
        G.check(self.lib.gocean_drift(d, int(self.getNumDrifters()), self.driftersDevice.data.gpudata,
                                      eta.data.gpudata, hu.data.gpudata, hv.data.gpudata, Hm.data.gpudata,
                                      float(np.float32(dt)), float(np.float32(self.sensitivity)), wind,
                                      float(self.wind_drift_factor), int(x_zero_ref), int(y_zero_ref),
                                      G.stream_handle(self.gpu_stream)), "gocean_drift")

    def observe(self, hu, hv, nx, ny, dx, dy, out=None, positions=None):
        """hu, hv at the drifters' cells (observeUnderlyingFlow,
        observationKernels.cu:29-82) as a (num, 2) device tensor."""
        torch = Common._torch()
        n = self.getNumDrifters()
        if out is None:
            with torch.cuda.stream(self.gpu_stream):
                out = torch.zeros((n, 2), dtype=torch.float32, device=self.driftersDevice.tensor.device)
        pos = self.driftersDevice.data.gpudata if positions is None else G.ptr(positions)
        d = self._domain(nx, ny, dx, dy, hu.pitch_elems)
        G.check(self.lib.gocean_observe_flow(d, n, pos, hu.data.gpudata, hv.data.gpudata, out.data_ptr(),
                                             2, 2, G.stream_handle(self.gpu_stream)), "gocean_observe_flow")
        return out

    def setGPUStream(self, gpu_stream):
        self.gpu_stream = gpu_stream

    def cleanUp(self):
        if getattr(self, "driftersDevice", None) is not None:
            self.driftersDevice.release()
            self.driftersDevice = None
        self.gpu_ctx = None

    def enforceBoundaryConditions(self):
        bc = self.boundaryConditions
        if bc.isPeriodicNorthSouth() or bc.isPeriodicEastWest():
            G.check(self.lib.gocean_drifters_enforce_bc(
                float(np.float32(self.domain_size_x)), float(np.float32(self.domain_size_y)),
                int(bc.isPeriodicNorthSouth()), int(bc.isPeriodicEastWest()), int(self.numDrifters),
                self.driftersDevice.data.gpudata, G.stream_handle(self.gpu_stream)),
                "gocean_drifters_enforce_bc")
