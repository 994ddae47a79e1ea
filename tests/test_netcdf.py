"""NetCDF output and reading back (SWESimulators/SimWriter.py, SimReader.py),
on the CPU with a stand-in simulator object: the reference writer's
attributes, dimensions and variables (gpu_ocean/SWESimulators/SimWriter.py:
161-441), record round trips, and the boundary-condition string round trip
the hot start relies on (SimReader.getBC, SimReader.py:64-66)."""

import numpy as np
import pytest
from scipy.io import netcdf_file

from SWESimulators import Common, SimReader, SimWriter, WindStress

# global attributes the reference writes for a CDKLM16 run (SimWriter.py:161-212)
REF_ATTRS = ["Conventions", "git_hash", "ignore_ghostcells", "num_layers", "staggered_grid",
             "simulator_short", "boundary_conditions", "boundary_conditions_sponge_mr", "time_integrator",
             "minmod_theta", "coriolis_force", "coriolis_beta", "y_zero_reference_cell",
             "eddy_viscosity_coefficient", "g", "nx", "ny", "dx", "dy", "dt", "auto_dt", "bottom_friction_r",
             "ghost_cells_north", "ghost_cells_east", "ghost_cells_south", "ghost_cells_west",
             "model_time_step", "small_scale_perturbation"]
REF_VARS = ["time", "x", "y", "x_Hi", "y_Hi", "projection_stere", "land_binary_mask", "Hm", "Hi",
            "eta", "hu", "hv"]


class _Bathy:
    def __init__(self, Hi, Hm):
        self.Hi, self.Hm = Hi, Hm

    def download(self, stream):
        return self.Hi, self.Hm


class CDKLM16:
    """Stand-in with the attributes the writer reads (named as the class it
    stands in for, since the file records the simulator's class name)."""

    def __init__(self, nx=12, ny=9, bc=(2, 1, 2, 1)):
        rs = np.random.RandomState(0)
        self.nx, self.ny, self.dx, self.dy = np.int32(nx), np.int32(ny), np.float32(200.0), np.float32(150.0)
        self.ghost_cells_x = self.ghost_cells_y = np.int32(2)
        self.gpu_stream = None
        self.boundary_conditions = Common.BoundaryConditions(*bc)
        self.rk_order, self.theta, self.f, self.coriolis_beta = np.int32(2), np.float32(1.3), np.float32(1e-4), np.float32(0)
        self.y_zero_reference_cell, self.A, self.g, self.r, self.dt = np.float32(2), 'NA', np.float32(9.81), np.float32(0), 0.5
        self.wind_stress = WindStress.WindStress()
        self.model_time_step, self.small_scale_perturbation = 0.5, False
        self.bathymetry = _Bathy(rs.rand(ny + 5, nx + 5).astype(np.float32), rs.rand(ny + 4, nx + 4).astype(np.float32))
        self.t = 0.0
        self.states = []

    def download(self):
        rs = np.random.RandomState(len(self.states) + 1)
        s = tuple(rs.randn(self.ny + 4, self.nx + 4).astype(np.float32) for _ in range(3))
        self.states.append((self.t, s))
        return s


def test_layout_and_round_trip(tmp_path):
    sim = CDKLM16()
    w = SimWriter.SimNetCDFWriter(sim, filename=str(tmp_path / "run"))
    for k in range(3):
        sim.t += 0.5
        w.writeTimestep(sim)
    w.close()
    f = netcdf_file(w.output_file_name, 'r', mmap=False)
    for a in REF_ATTRS:
        assert a in f._attributes, a
    for v in REF_VARS:
        assert v in f.variables, v
    assert f.dimensions['x'] == 16 and f.dimensions['y'] == 13 and f.dimensions['x_Hi'] == 17
    assert f.variables['eta'].shape == (4, 13, 16)
    f.close()
    r = SimReader.SimNetCDFReader(w.output_file_name, ignore_ghostcells=False)
    assert r.getNumTimeSteps() == 4
    np.testing.assert_array_equal(r.getTimes(), [0.0, 0.5, 1.0, 1.5])
    for k, (t, (e, u, v)) in enumerate(sim.states):
        ge, gu, gv, gt = r.getTimeStep(k)
        assert gt == np.float32(t)
        assert np.array_equal(ge, e) and np.array_equal(gu, u) and np.array_equal(gv, v)
    assert np.array_equal(r.getH(), sim.bathymetry.Hi) and np.array_equal(r.getHm(), sim.bathymetry.Hm)
    assert r.get("simulator_short") == "CDKLM16" and r.get("nx") == 12 and r.get("dx") == 200.0
    assert r.get("eddy_viscosity_coefficient") == "NA" and r.get("nothing") == "not found"
    assert r.getBC().codes() == (2, 1, 2, 1)
    ri = SimReader.SimNetCDFReader(w.output_file_name, ignore_ghostcells=True)
    e, u, v, t = ri.getStateAtTime(1.0)
    assert np.array_equal(e, sim.states[2][1][0][2:-2, 2:-2])
    x, y = ri.getAxis()
    np.testing.assert_allclose(x, (np.arange(12) + 0.5) * 200.0, rtol=1e-6)
    with pytest.raises(RuntimeError):
        ri.getStateAtTime(7.0)


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2), (2, 1, 2, 1), (1, 2, 1, 2)])
def test_bc_string_round_trip(bc):
    b = Common.BoundaryConditions(*bc)
    assert Common.BoundaryConditions.fromstring(str(b)).codes() == bc


def test_reader_returns_native_byte_order(tmp_path):
    sim = CDKLM16()
    w = SimWriter.SimNetCDFWriter(sim, filename=str(tmp_path / "bo"))
    r = SimReader.SimNetCDFReader(w.output_file_name, ignore_ghostcells=False)
    e, u, v, t = r.getTimeStep(0)
    for a in (e, u, v, r.getH(), r.getHm(), r.getTimes()):
        assert a.dtype.isnative
