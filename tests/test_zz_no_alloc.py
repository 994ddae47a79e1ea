"""The entry point that allocates nothing (gocean_max_dt_static: partial
minima in a library-owned static device array) against gocean_max_dt (partial
minima in the caller's buffer): the same two floats, bit for bit, on calls in
a row with shrinking domains, through the C ABI, with a two-float output."""

import numpy as np
import pytest
import torch

import helpers as H
from oracle import cdklm16_np as O

pytestmark = pytest.mark.gpu


def _run(nx, ny, seed):
    from SWESimulators import _gocean as G
    lib = G.load()
    p, Hi, Hm, eta, hu, hv = H.random_case(nx, ny, (1, 2, 1, 2), seed, land=True)

    class Inp:
        pass
    inp = Inp()
    inp.nx, inp.ny, inp.dx, inp.dy, inp.g, inp.r = nx, ny, p.dx, p.dy, p.g, p.r
    inp.bc, inp.f, inp.wind_x, inp.wind_y, inp.Hi = p.bc, float(p.f), 0.0, 0.0, Hi
    D = H.DeviceDomain(inp)
    D.bathymetry_init()
    h = D.state(eta, hu, hv)
    want = D.max_dt(h, courant=0.7)
    out = torch.full((2,), -1.0, dtype=torch.float32, device="cuda")
    G.check(lib.gocean_max_dt_static(D.dom, h[0].data_ptr(), h[1].data_ptr(), h[2].data_ptr(),
                                     D.Hm.data_ptr(), 0.7, out.data_ptr(), None), "max_dt_static")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    print(f"   {nx}x{ny}: static {got}, caller buffer {want}", flush=True)
    assert np.array_equal(got, want)
    assert got[0] == O.max_dt(eta, hu, hv, D.Hm[:, :nx + 4].cpu().numpy(), p)


def test_max_dt_static_matches_caller_buffer_on_shrinking_domains():
    print("no-alloc max_dt: 300x1500, 1100x900, 37x29", flush=True)
    _run(300, 1500, 2)     # 1024 partial minima (the array's size)
    _run(1100, 900, 3)     # 900
    _run(37, 29, 4)        # 29: the array keeps the earlier calls' tail
