"""NetCDF output of a running CDKLM16 simulator and the hot start from it
(CDKLM16.fromfilename, gpu_ocean/SWESimulators/CDKLM16.py:410-490): every
step's record equals the downloaded state, and a simulator started from the
file continues bit for bit like the one that wrote it."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_write_and_hot_start(tmp_path):
    from SWESimulators import CDKLM16, Common, SimReader
    nx, ny, dx, dy = 40, 32, 400.0, 400.0
    j, i = np.mgrid[0:ny + 5, 0:nx + 5]
    Hi = (40 + 5 * np.sin(i / 5.0) + 3 * np.cos(j / 7.0)).astype(np.float32)
    eta = (0.3 * np.exp(-((i[:-1, :-1] - nx / 2) ** 2 + (j[:-1, :-1] - ny / 2) ** 2) / 40.0)).astype(np.float32)
    z = np.zeros((ny + 4, nx + 4), dtype=np.float32)
    ctx = Common.HIPContext(0)
    bc = Common.BoundaryConditions(2, 1, 2, 1)
    sim = CDKLM16.CDKLM16(ctx, eta, z, z, Hi, nx, ny, dx, dy, 2.0, 9.81, 1e-4, 0.0, boundary_conditions=bc,
                          write_netcdf=True, netcdf_filename=str(tmp_path / "run"))
    states = [sim.download()]
    for k in range(4):
        sim.step(6.0)
        states.append(sim.download())
    fn = sim.sim_writer.output_file_name
    r = SimReader.SimNetCDFReader(fn, ignore_ghostcells=False)
    assert r.getNumTimeSteps() == 5
    for k, s in enumerate(states):
        e, u, v, t = r.getTimeStep(k)
        assert np.array_equal(e, s[0]) and np.array_equal(u, s[1]) and np.array_equal(v, s[2]), k
        assert t == np.float32(6.0 * k)
    assert r.getBC().codes() == (2, 1, 2, 1)
    r.close()
    hot = CDKLM16.CDKLM16.fromfilename(ctx, fn, cont_write_netcdf=False)
    assert hot.t == 24.0 and hot.dt == 2.0 and hot.boundary_conditions.codes() == (2, 1, 2, 1)
    for k in range(3):
        sim.step(6.0)
        hot.step(6.0)
        for a, b in zip(sim.download(), hot.download()):
            assert np.array_equal(a, b), k
    sim.cleanUp()
    hot.cleanUp()
