"""HIP-graph capture of the single-domain step loop (``CDKLM16.capture_graph``):
replays of a captured run of fixed-dt steps equal the same steps launched
eagerly, bit for bit, with the program's own eager work and synchronisations
between the replays; for the column-marching and the tile kernel, and with
the marching kernel's gridded-wind variant."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _sim(inp, ctx, tile, gridded):
    from SWESimulators import CDKLM16, Common, WindStress
    bc = Common.BoundaryConditions(*inp.bc)
    if gridded:
        rng = np.random.default_rng(5)
        X = [(1e-4 * rng.standard_normal((6, 7))).astype(np.float32)]
        Y = [(1e-4 * rng.standard_normal((6, 7))).astype(np.float32)]
    else:
        X = [np.array([[inp.wind_x]], np.float32)]
        Y = [np.array([[inp.wind_y]], np.float32)]
    wind = WindStress.WindStress(t=[0.0], X=X, Y=Y)
    return CDKLM16.CDKLM16(ctx, inp.eta, inp.hu, inp.hv, inp.Hi, inp.nx, inp.ny, inp.dx, inp.dy, 2.0,
                           inp.g, inp.f, inp.r, boundary_conditions=bc, wind_stress=wind, tile_config=tile)


@pytest.mark.parametrize("tile,gridded", [(9, False), (9, True), (1, False)])
def test_graph_replays_equal_eager_steps(tile, gridded):
    from SWESimulators import BathymetryAndICs as B, Common
    print("graph replay vs eager, tile", tile, "gridded wind", gridded, flush=True)
    ctx = Common.HIPContext(0)
    inp = B.config2(nx=200, ny=150)
    a = _sim(inp, ctx, tile, gridded)
    b = _sim(inp, ctx, tile, gridded)
    g = b.capture_graph(3)
    torch.cuda.synchronize()
    state = lambda s: [t.tensor.clone() for t in (s.gpu_data.h0, s.gpu_data.hu0, s.gpu_data.hv0)]
    start = state(b)
    for r in range(3):
        a.step_n(3)
        torch.cuda.synchronize()
        with torch.cuda.stream(b.gpu_stream):
            g.replay()
        torch.cuda.synchronize()
        _ = float(b.gpu_data.h0.tensor.sum())          # eager work between replays
        for x, y, name in zip(state(a), state(b), ("eta", "hu", "hv")):
            assert torch.equal(x, y), (r, name)
    assert not torch.equal(start[0], state(b)[0])
