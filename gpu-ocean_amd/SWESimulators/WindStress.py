"""Wind-stress forcing (gpu_ocean/SWESimulators/WindStress.py:34-60).

A time series of wind-stress fields X, Y (tau/rho_water, m^2/s^2).  1x1
fields are uniform wind (the reference's 1x1 textures return the single value
everywhere); larger fields are gridded wind, which the simulator reads at every
cell like the reference's bilinear textures.
"""

from __future__ import annotations

import numpy as np


class WindStress():

    def __init__(self, source_filename=None, t=None, X=None, Y=None):
        self.source_filename = source_filename
        self.t = [0]
        self.X = [np.zeros((1, 1), dtype=np.float32, order='C')]
        self.Y = [np.zeros((1, 1), dtype=np.float32, order='C')]
        self.numWindSteps = 1
        if t is not None:
            assert X is not None, "missing wind forcing X"
            assert Y is not None, "missing wind forcing Y"
            assert len(t) == len(X), str(len(t)) + " vs " + str(len(X))
            assert len(t) == len(Y), str(len(t)) + " vs " + str(len(Y))
            self.numWindSteps = len(t)
            for i in range(len(X)):
                assert X[i].dtype == 'float32', "Wind data needs to be of type np.float32"
                assert Y[i].dtype == 'float32', "Wind data needs to be of type np.float32"
            self.t = t
            self.X = X
            self.Y = Y

    def is_uniform(self):
        return all(np.asarray(a).size == 1 for a in list(self.X) + list(self.Y))


def uniform_wind_stress(tau, direction_deg, rho_water=1025.0):
    """Uniform stress ``tau`` (N/m^2) blowing ``direction_deg`` counter-
    clockwise from +x, as a WindStress of 1x1 fields in m^2/s^2."""
    a = np.deg2rad(direction_deg)
    X = np.array([[tau * np.cos(a) / rho_water]], dtype=np.float32)
    Y = np.array([[tau * np.sin(a) / rho_water]], dtype=np.float32)
    return WindStress(t=[0.0], X=[X], Y=[Y])
