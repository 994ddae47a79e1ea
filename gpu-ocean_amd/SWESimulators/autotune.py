"""Tile-shape autotuner for the CDKLM16 substep kernels.

Times every tile configuration the library is built with on the caller's own
domain or ensemble (a few fixed-dt steps each, HIP events on its stream),
restores the state afterwards, and returns the fastest.  Results are cached
per (sizes, boundary conditions) for the life of the process.
"""

from __future__ import annotations

import ctypes

from . import _gocean as G

_cache = {}


def tile_config_names():
    lib = G.load()
    names = (ctypes.c_char_p * 64)()
    n = lib.gocean_tile_configs(names, 64)
    return [names[i].decode() for i in range(n)]


def tune_simulator(sim, steps=8, warmup=2, configs=None, use_cache=True):
    """Return (best_config, {config: ms_per_step}) for ``sim`` (a CDKLM16).
    The simulator's state, time and iteration count are left unchanged; its
    ``tile_config`` is set to the winner."""
    import torch
    key = (int(sim.nx), int(sim.ny), tuple(sim.boundary_conditions.codes()),
           sim.field_tensors is not None, (sim.north_const[0], sim.north_const[1]))
    if use_cache and key in _cache:
        sim.tile_config = _cache[key][0]
        return _cache[key]
    if configs is None:
        configs = list(range(len(tile_config_names())))
    gd = sim.gpu_data
    arrays = [gd.h0, gd.hu0, gd.hv0, gd.h1, gd.hu1, gd.hv1]
    st = sim.gpu_stream
    with torch.cuda.stream(st):
        saved = [a.tensor.clone() for a in arrays]
    t_saved, it_saved = sim.t, sim.num_iterations
    times = {}
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    for c in configs:
        sim.tile_config = c
        sim.step_n(warmup)
        ev0.record(st)
        sim.step_n(steps)
        ev1.record(st)
        ev1.synchronize()
        times[c] = ev0.elapsed_time(ev1) / steps
    with torch.cuda.stream(st):
        for a, s in zip(arrays, saved):
            a.tensor.copy_(s)
    st.synchronize()
    sim.t, sim.num_iterations = t_saved, it_saved
    best = min(times, key=times.get)
    sim.tile_config = best
    _cache[key] = (best, times)
    return best, times


def tune_ensemble(ens, steps=4, warmup=1, configs=None, use_cache=True):
    """Return (best_config, {config: ms_per_step}) for a ``CDKLM16Ensemble``:
    every tile configuration of the batched substep timed on the ensemble's
    own members (one launch per substep covers them all).  The ensemble's
    state and iteration count are left unchanged and its tile configuration
    is set to the winner; a captured graph must be captured again."""
    import torch
    key = ("ensemble", tuple(ens.nx), tuple(ens.ny), tuple(tuple(b) for b in ens.bcs), ens.rk_order)
    if use_cache and key in _cache:
        ens.set_tile_config(_cache[key][0])
        return _cache[key]
    if configs is None:
        configs = list(range(len(tile_config_names())))
    st = ens.gpu_stream
    bufs = list(ens.state) + ([ens.state2] if getattr(ens, "state2", None) is not None else [])
    with torch.cuda.stream(st):
        saved = [b.clone() for b in bufs]
    times = {}
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    for c in configs:
        ens.set_tile_config(c)
        # plain RK steps only: no model error, drifters or iteration count
        ens._rk(warmup)
        ev0.record(st)
        ens._rk(steps)
        ev1.record(st)
        ev1.synchronize()
        times[c] = ev0.elapsed_time(ev1) / steps
        with torch.cuda.stream(st):
            for b, s in zip(bufs, saved):
                b.copy_(s)
    st.synchronize()
    best = min(times, key=times.get)
    ens.set_tile_config(best)
    _cache[key] = (best, times)
    return best, times


def tune_adjoint_record(adj, q, dt, steps=4, configs=None):
    """Pick the fastest tile configuration of the recording forward for the
    RK2Adjoint ``adj``: time ``steps`` recorded steps from the (3, ny+4,
    pitch) state ``q`` per configuration (the record is the same for every
    configuration).  Sets ``adj.record_tile_config``; returns (best,
    {config: ms_per_step})."""
    import torch
    from . import _gocean as G
    lib = adj.lib
    if configs is None:
        configs = list(range(len(tile_config_names())))
    nw = int(lib.gocean_adjoint_segment_work_floats(adj.d, steps))
    st = adj.stream
    with torch.cuda.stream(st):
        w = torch.empty(nw, dtype=torch.float32, device=adj.dev)
        start = [a.clone() for a in q.unbind(0)]
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    times = {}

    def record():
        G.check(lib.gocean_adjoint_record(adj.d, adj.wind, steps, float(dt), None, *(G.ptr(a) for a in start),
                                          G.ptr(adj.Hi), G.ptr(w), nw, G.stream_handle(st)),
                "gocean_adjoint_record")

    for c in configs:
        G.check(lib.gocean_adjoint_set_record_tile_config(c), "gocean_adjoint_set_record_tile_config")
        record()
        ev0.record(st)
        record()
        ev1.record(st)
        ev1.synchronize()
        times[c] = ev0.elapsed_time(ev1) / steps
    best = min(times, key=times.get)
    adj.record_tile_config = best
    G.check(lib.gocean_adjoint_set_record_tile_config(best), "gocean_adjoint_set_record_tile_config")
    return best, times
