"""Reading simulator output back (gpu_ocean/SWESimulators/SimReader.py:36-200):
attributes, the time axis, states per record, bathymetry and the boundary
conditions, from the files SimWriter writes (NetCDF-3, scipy.io)."""

from __future__ import annotations

import numpy as np
from scipy.io import netcdf_file

from . import Common


def _native(a):
    """Copy in native byte order (NetCDF stores big-endian values)."""
    a = np.asarray(a)
    return a.astype(a.dtype.newbyteorder('='))


def _decode(v):
    if isinstance(v, bytes):
        return v.decode()
    if isinstance(v, np.ndarray) and v.size == 1:
        return v.reshape(()).item()
    return v


class SimNetCDFReader:

    def __init__(self, filename, ignore_ghostcells=True):
        self.filename = filename
        self.ignore_ghostcells = ignore_ghostcells
        self.ncfile = netcdf_file(filename, 'r', mmap=False)
        self.ghostCells = [self.get('ghost_cells_north'), self.get('ghost_cells_east'),
                           self.get('ghost_cells_south'), self.get('ghost_cells_west')]
        self.staggered_grid = str(self.get('staggered_grid')) == 'True'

    def close(self):
        if self.ncfile is not None:
            self.ncfile.close()
            self.ncfile = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, attr):
        if attr in self.ncfile._attributes:
            return _decode(self.ncfile._attributes[attr])
        return "not found"

    def has(self, fieldname):
        return fieldname in self.ncfile._attributes or fieldname in self.ncfile.variables

    def printVariables(self):
        for var in self.ncfile.variables:
            print(var)

    def printAttributes(self):
        for attr in self.ncfile._attributes:
            print(attr + "\t--> " + str(self.get(attr)))

    def getNumTimeSteps(self):
        return self.ncfile.variables['time'].shape[0]

    def getBC(self):
        return Common.BoundaryConditions.fromstring(self.get("boundary_conditions"))

    def getTimes(self):
        return _native(self.ncfile.variables['time'][:])

    def _interior(self, a):
        g = self.ghostCells
        return a[g[2]:-g[0], g[3]:-g[1]]

    def getLastTimeStep(self):
        return self.getTimeStep(-1)

    def getTimeStep(self, index):
        time = self.ncfile.variables['time']
        eta, hu, hv = (_native(self.ncfile.variables[v][index, :, :]) for v in ('eta', 'hu', 'hv'))
        if self.ignore_ghostcells:
            eta, hu, hv = (self._interior(a) for a in (eta, hu, hv))
        return eta, hu, hv, np.float32(time[index])

    def getH(self):
        for name in ('H', 'Hm' if self.staggered_grid else 'Hi'):
            if name in self.ncfile.variables:
                return _native(self.ncfile.variables[name][:, :])
        raise KeyError("no bathymetry in " + self.filename)

    def getHm(self):
        return _native(self.ncfile.variables['Hm'][:, :])

    def getStateAtTime(self, time):
        time = np.round(time)
        times = self.getTimes()
        hits = np.nonzero(times == time)[0]
        if len(hits) == 0:
            raise RuntimeError('Time ' + str(time) + ' not in NetCDF file ' + self.filename)
        return self.getStateAtTimeStep(int(hits[0]))

    def getStateAtTimeStep(self, index, etaOnly=False):
        time = self.ncfile.variables['time']
        eta = _native(self.ncfile.variables['eta'][index, :, :])
        if self.ignore_ghostcells:
            eta = self._interior(eta)
        if etaOnly:
            return eta, time[index]
        hu = _native(self.ncfile.variables['hu'][index, :, :])
        hv = _native(self.ncfile.variables['hv'][index, :, :])
        if self.ignore_ghostcells:
            hu, hv = self._interior(hu), self._interior(hv)
        return eta, hu, hv, time[index]

    def getEtaAtTimeStep(self, index):
        return self.getStateAtTimeStep(index, etaOnly=True)

    def getAxis(self):
        x = _native(self.ncfile.variables['x'][:])
        y = _native(self.ncfile.variables['y'][:])
        if self.ignore_ghostcells:
            g = self.ghostCells
            x = x[g[3]:-g[1]]
            y = y[g[2]:-g[0]]
        return x, y

    def getEtaXSlice(self, t, y):
        y_index = int(y) + int(self.get('ghost_cells_south'))
        return _native(self.ncfile.variables['eta'][t, y_index, self.ghostCells[3]:-self.ghostCells[1]])
