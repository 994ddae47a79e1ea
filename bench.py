"""Benchmark of the CDKLM16 step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]

Headline workload (printed as ONE JSON line): plain CDKLM16, 2048 x 2048 cells,
dx=dy=500 m, random depth, periodic E/W + walls N/S, 20 balanced eddies,
uniform wind, fp32, fixed dt from CFL 0.7; a "step" is one full SSP-RK2 step
(two substep launches) over the whole domain with the state resident in HBM.
value = Mcell-steps/s = nx*ny*K / wall_s / 1e6 (the reference's Mcells/s,
gpu_ocean/papers/testCasesDemos/run_benchmark.py:278-288).

roofline: HBM-bound.  Algorithmic bytes are 68 B per cell-step (rk0 reads
Q^n 12 B + H 4 B and writes Q* 12 B; rk1 reads Q* 12 + Q^n 12 + H 4 B and
writes 12 B), i.e. 34 B per cell per substep launch; achieved = 34 B x cells
/ average substep-launch duration (HIP events on the launch stream).
cpu_baseline: the NumPy oracle (single core) on a bounded sample of the same
workload (one RK2 step of the 2048^2 domain).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gpu-ocean_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STEP = 68.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_sim(inp, ctx, tile_config=None, dt=None):
    from SWESimulators import CDKLM16, Common, WindStress
    bc = Common.BoundaryConditions(*inp.bc)
    wind = WindStress.WindStress(t=[0.0], X=[np.array([[inp.wind_x]], np.float32)],
                                 Y=[np.array([[inp.wind_y]], np.float32)])
    sim = CDKLM16.CDKLM16(ctx, inp.eta, inp.hu, inp.hv, inp.Hi, inp.nx, inp.ny, inp.dx, inp.dy,
                          1.0 if dt is None else dt, inp.g, inp.f, inp.r,
                          boundary_conditions=bc, wind_stress=wind, tile_config=tile_config)
    if dt is None:
        # dt = courant * CFL minimum, computed once on the device and held fixed
        s = sim.max_dt_device(courant_number=inp.courant)
        sim.gpu_stream.synchronize()
        sim.dt = float(s[1].item())
    return sim


def time_isolated_step(sim, reps=20):
    """Median time of one RK2 step (two launches) bracketed by its own pair
    of events and followed by a synchronisation, as a profiler's per-dispatch
    timestamps see it; time_plain's back-to-back average is the throughput.
    A short spin kernel ahead of the first event keeps the host's launch
    latency out of the bracket."""
    import torch
    st = sim.gpu_stream
    out = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(st):
            torch.cuda._sleep(200000)
        e0.record(st)
        sim.step_n(1)
        e1.record(st)
        e1.synchronize()
        out.append(e0.elapsed_time(e1) / 1000.0)
    return float(np.median(out))


def time_plain(sim, steps, warmup):
    """Time ``steps`` fixed-dt RK2 steps; returns (seconds, launches)."""
    import torch
    sim.step_n(warmup)
    st = sim.gpu_stream
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    sim.step_n(steps)
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = e0.elapsed_time(e1) / 1000.0
    return wall, ev, 2 * steps


def cpu_baseline(inp, n_steps=1):
    """NumPy oracle, one core, on ``n_steps`` RK2 steps of the same input."""
    from oracle import cdklm16_np as O
    p = O.Params(nx=inp.nx, ny=inp.ny, dx=inp.dx, dy=inp.dy, g=inp.g, r=inp.r, bc=inp.bc,
                 f=inp.f, wind_x=(inp.wind_x, inp.wind_x), wind_y=(inp.wind_y, inp.wind_y))
    sim = O.OracleSim(inp.eta, inp.hu, inp.hv, inp.Hi, p, dt=-1)
    dt = np.float32(inp.courant * sim.dt / 0.8)
    t0 = time.perf_counter()
    sim.step_n(n_steps, dt)
    s = time.perf_counter() - t0
    return inp.nx * inp.ny * n_steps / s / 1e6, s


def _cpu_member_steps(args):
    """Pool worker: ``n_steps`` RK2 steps of one member with the NumPy oracle
    (oracle setup included: depth halo, cell depth, CFL dt)."""
    eta, hu, hv, Hi, nx, ny, dx, dy, f, bc, courant, n_steps = args
    from oracle import cdklm16_np as O
    p = O.Params(nx=nx, ny=ny, dx=dx, dy=dy, f=f, bc=bc)
    sim = O.OracleSim(eta, hu, hv, Hi, p, dt=-1)
    sim.step_n(n_steps, np.float32(courant * sim.dt / 0.8))
    return nx * ny * n_steps


def _cpu_pool_ready(_):
    from oracle import cdklm16_np  # noqa: F401
    return os.getpid()


def cpu_baseline_ensemble(sample_members=32, n=512, steps=2, procs=None):
    """C3 members split across a process pool of the host's cores (at most
    16), each process running the NumPy oracle on whole members: a bounded
    sample of ``sample_members`` members x ``steps`` RK2 steps."""
    import multiprocessing as mp
    from SWESimulators import BathymetryAndICs as B
    procs = procs or min(16, os.cpu_count() or 1)
    inps = B.config3(members=sample_members, n=n)
    args = [(i.eta, i.hu, i.hv, i.Hi, i.nx, i.ny, i.dx, i.dy, i.f, i.bc, i.courant, steps) for i in inps]
    # spawned workers: fresh interpreters that never touch the GPU
    with mp.get_context("spawn").Pool(procs) as pool:
        pool.map(_cpu_pool_ready, range(procs))
        t0 = time.perf_counter()
        done = sum(pool.map(_cpu_member_steps, args, chunksize=1))
        s = time.perf_counter() - t0
    return {"value": round(done / s / 1e6, 3), "unit": "Mcell-steps/s", "cores": procs, "kind": "port",
            "sample": f"NumPy fp32 oracle, {sample_members} C3 members of {n}x{n} x {steps} RK2 steps in a "
                      f"pool of {procs} spawned processes ({s:.1f} s, oracle setup per member included)"}


def cpu_baseline_adjoint(steps=2, n=512, n_obs=64, seed=44):
    """C4 on the host: the float64 torch restatement of the forward
    (oracle/cdklm16_torch.py) over ``steps`` RK2 steps of the 512^2 domain and
    the gradient of the 64-point misfit by torch autograd, timed together."""
    import torch
    from oracle import cdklm16_np as O, cdklm16_torch as T
    from SWESimulators import BathymetryAndICs as B
    inp = B.config4(n=n)
    p = O.Params(nx=inp.nx, ny=inp.ny, dx=inp.dx, dy=inp.dy, f=inp.f, bc=inp.bc)
    Hi = O.bathymetry_halo(inp.Hi, p).astype(np.float64)
    Hm = O.init_bm(Hi, dtype=np.float64)
    sim = O.OracleSim(inp.eta, inp.hu, inp.hv, inp.Hi, p, dt=-1)
    dt = float(np.float32(inp.courant * sim.dt / 0.8))
    rng = np.random.default_rng(seed)
    ox = rng.uniform(2.0, n - 2.0, n_obs) * inp.dx
    oy = rng.uniform(2.0, n - 2.0, n_obs) * inp.dy
    fx = torch.from_numpy(ox / inp.dx - 0.5)
    fy = torch.from_numpy(oy / inp.dy - 0.5)
    i0, j0 = fx.floor().long(), fy.floor().long()
    wx, wy = fx - i0, fy - j0
    y = torch.zeros(n_obs, dtype=torch.float64)
    q = [torch.from_numpy(np.asarray(a, np.float64)).requires_grad_(True) for a in (inp.eta, inp.hu, inp.hv)]
    Hit, Hmt = torch.from_numpy(Hi), torch.from_numpy(Hm)
    t0 = time.perf_counter()
    out = T.rk2_steps(*q, Hit, Hmt, p, dt, steps)
    e = out[0]
    r, c = j0 + 2, i0 + 2
    est = (1 - wy) * ((1 - wx) * e[r, c] + wx * e[r, c + 1]) + wy * ((1 - wx) * e[r + 1, c] + wx * e[r + 1, c + 1])
    J = 0.5 * ((est - y) ** 2).sum()
    J.backward()
    s = time.perf_counter() - t0
    return {"value": round(n * n * steps / s / 1e6, 4), "unit": "Mcell-steps/s (forward + adjoint)",
            "cores": torch.get_num_threads(), "kind": "port",
            "sample": f"float64 torch CPU restatement of the forward + autograd gradient of the 64-point "
                      f"misfit, {steps} RK2 steps of the {n}x{n} C4 domain ({s:.1f} s, "
                      f"{torch.get_num_threads()} torch threads)"}


def bench_ensemble(ctx, members=100, n=512, steps=50, warmup=5, per_member_steps=5):
    """C3: ``members`` domains of n^2 advanced by the batched launch (eager
    and HIP-graph replay), and the same members stepped one member per launch
    from a Python loop (the reference's ensemble pattern)."""
    import torch
    from SWESimulators import BathymetryAndICs as B, CDKLM16Ensemble
    from SWESimulators import autotune
    inps = B.config3(members=members, n=n)
    cells = sum(i.nx * i.ny for i in inps)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, inps, courant_number=0.7)
    best, tune_ms = autotune.tune_ensemble(ens)
    names = autotune.tile_config_names()
    st = ens.gpu_stream
    ens.step_n(warmup)
    st.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    ens.step_n(steps)
    e1.record(st)
    st.synchronize()
    wall = time.perf_counter() - t0
    eager = cells * steps / wall / 1e6
    kernel_s = e0.elapsed_time(e1) / 1000.0 / (2 * steps)
    ens.capture(steps)
    ens.replay()
    st.synchronize()
    t0 = time.perf_counter()
    ens.replay()
    st.synchronize()
    gwall = time.perf_counter() - t0
    graph = cells * steps / gwall / 1e6
    ens.cleanUp()
    del ens
    # one member per launch (plain drop-in class, Python loop over members),
    # with the tile the single-domain autotuner picks for these members
    sims = [make_sim(i, ctx) for i in inps]
    pbest, _ = autotune.tune_simulator(sims[0])
    for sm in sims:
        sm.tile_config = pbest
        sm.step_n(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(per_member_steps):
        for sm in sims:
            sm.step_n(1)
    torch.cuda.synchronize()
    pwall = time.perf_counter() - t0
    plain = cells * per_member_steps / pwall / 1e6
    for sm in sims:
        sm.cleanUp()
    return {"workload": f"{members} members x {n}x{n}, walls, random depth, CFL 0.7 per member, SSP-RK2",
            "cells": cells,
            "batched_eager_mcell_steps_s": round(eager, 1),
            "batched_graph_mcell_steps_s": round(graph, 1),
            "plain_per_member_mcell_steps_s": round(plain, 1),
            "batched_vs_plain": round(max(eager, graph) / plain, 2),
            "batched_tile": names[best], "plain_tile": names[pbest],
            "batched_substep_launch_us": round(kernel_s * 1e6, 1),
            "batched_hbm_frac_68B": round(max(eager, graph) * 1e6 * BYTES_PER_CELL_STEP / (HBM_PEAK_GBS * 1e9), 4),
            "steps": steps, "per_member_steps": per_member_steps}


def bench_adjoint(ctx, n=512, steps=500, K=50, n_obs=64, seed=44):
    """C4: gradient of J = 0.5 sum_k (eta(x_k, T) - y_k)^2 (64 point
    observations, bilinear) with respect to the initial eta, hu, hv of one
    n^2 domain after ``steps`` fixed-dt RK2 steps, checkpoint every K.  The
    observations come from a twin run started from a perturbed eta plus
    N(0, 0.01 m) noise.  Reports the forward time (plain step loop) and the
    forward + adjoint time of the autograd wrapper, and their ratio."""
    import torch
    from SWESimulators import BathymetryAndICs as B, adjoint as A
    rng = np.random.default_rng(seed)
    from SWESimulators import autotune
    inp = B.config4(n=n)
    sim = make_sim(inp, ctx)
    autotune.tune_simulator(sim)       # the forward it is compared with: the fastest tile
    dt = sim.dt
    st = sim.gpu_stream
    ox = rng.uniform(2.0, n - 2.0, n_obs) * inp.dx
    oy = rng.uniform(2.0, n - 2.0, n_obs) * inp.dy
    q0 = [a.tensor.clone() for a in (sim.gpu_data.h0, sim.gpu_data.hu0, sim.gpu_data.hv0)]
    adj = A.simulator_adjoint(sim, checkpoint_every=K)
    # the recording forward on its own fastest tile (every tile records the same)
    rec_best, rec_times = autotune.tune_adjoint_record(adj, torch.stack(q0), dt, steps=4)
    # twin run: perturbed initial eta
    pert = 0.05 * B.gaussian_random_field((n + 4, n + 4), 16.0, rng).astype(np.float32)
    twin0 = [q0[0] + torch.from_numpy(pert).to(q0[0].device), q0[1], q0[2]]
    # the observation operator as one bilinear stencil gather over the eta plane
    sidx, swt = A.bilinear_stencil(0, q0[0].shape[1], ox, oy, inp.dx, inp.dy)
    sidx = torch.from_numpy(sidx).to(q0[0].device)
    swt = torch.from_numpy(swt).to(q0[0].device)
    with torch.no_grad():
        tw = adj.forward(twin0, steps, dt)
        y = (tw[0].reshape(-1)[sidx] * swt).sum(1) + torch.from_numpy(
            rng.normal(0.0, 0.01, n_obs).astype(np.float32)).to(q0[0].device)
    # forward alone: the plain step loop
    sim.step_n(5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step_n(steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0

    def fwd_adj(a):
        q = [x.clone().requires_grad_(True) for x in q0]
        out = A.CDKLM16RK2.apply(*q, a, steps, dt)
        J = A.stencil_misfit(out[0].reshape(-1), sidx, swt, y)
        J.backward()
        return J, q

    def timed(a):
        fwd_adj(a)                  # warm-up (allocations, first launches)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, q = fwd_adj(a)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, J, q

    # records kept from the forward run (default when they fit the budget)
    t_tot, J, q = timed(adj)
    kept = adj.records is not None
    rec_gb = adj.records_bytes(steps) / 1e9
    adj.records = None
    # checkpointed: the backward recomputes each segment
    adj_rc = A.simulator_adjoint(sim, checkpoint_every=K)
    adj_rc.keep_records = False
    adj_rc.record_tile_config = rec_best
    t_rc, J_rc, q_rc = timed(adj_rc)
    same = all(torch.equal(a.grad, b.grad) for a, b in zip(q, q_rc))
    cells = n * n
    g = [float(x.grad.norm()) for x in q]
    sim.cleanUp()
    return {"workload": f"{n}x{n} walls, random depth, 8 eddies, fixed dt {dt:.3f} s (CFL 0.7), "
                        f"{steps} RK2 steps, checkpoint every {K}, {n_obs} point observations",
            "fwd_s": round(t_fwd, 4), "fwd_adj_s": round(t_tot, 4),
            "fwd_adj_over_fwd": round(t_tot / t_fwd, 2), "records_kept": kept,
            "records_gb": round(rec_gb, 2),
            "record_tile": autotune.tile_config_names()[rec_best],
            "record_ms_per_step_by_tile": {autotune.tile_config_names()[c]: round(v, 4)
                                           for c, v in rec_times.items()},
            "fwd_adj_recompute_s": round(t_rc, 4), "fwd_adj_recompute_over_fwd": round(t_rc / t_fwd, 2),
            "kept_vs_recompute_grad_identical": bool(same),
            "fwd_mcell_steps_s": round(cells * steps / t_fwd / 1e6, 1),
            "fwd_adj_mcell_steps_s": round(cells * steps / t_tot / 1e6, 1),
            "J": float(J), "grad_norms_eta_hu_hv": [round(v, 6) for v in g]}


def bench_c5(ctx, steps=200, K=25, n_obs=16, seed=55, plain_members=4):
    """C5: 32 domains with sides from {256, 384, 512, 768}, walls, own random
    depth; per-member fixed dt (CFL 0.7); forward + adjoint of 200 RK2 steps
    with a checkpoint every 25, in fused launches over all members; a
    16-point eta misfit per domain (observations: twin run from a perturbed
    eta plus N(0, 0.01 m) noise) and its per-domain gradient.  Also times the
    single-domain reverse mode member by member on the first
    ``plain_members`` members, for the batched-vs-plain ratio."""
    import torch
    from SWESimulators import BathymetryAndICs as B, CDKLM16Ensemble, adjoint as A
    rng = np.random.default_rng(seed)
    inps = B.config5()
    cells = sum(i.nx * i.ny for i in inps)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, inps, courant_number=0.7)
    from SWESimulators import autotune
    autotune.tune_ensemble(ens)
    # the records of this run take ~113 GB: allowed up to 60 % of free HBM
    ea = A.EnsembleAdjoint(ens, checkpoint_every=K, record_budget=0.6)
    state0 = ens.state[0].clone()
    obs = []
    for k, i in enumerate(inps):
        ox = rng.uniform(2.0, i.nx - 2.0, n_obs) * i.dx
        oy = rng.uniform(2.0, i.ny - 2.0, n_obs) * i.dy
        obs.append((ox, oy))
    # twin run
    twin0 = state0.clone()
    for k, i in enumerate(inps):
        o = int(ens.state_off[k])
        pert = 0.05 * B.gaussian_random_field((i.ny + 4, i.nx + 4), 16.0, rng).astype(np.float32)
        twin0[o:o + ens.planes[k]] += torch.from_numpy(pert.ravel()).to(twin0.device)
    # one bilinear stencil over every domain's eta plane: the misfit is a
    # gather, a product and two sums for all 32 x n_obs points
    st = [A.bilinear_stencil(int(ens.state_off[k]), ens.pitch[k], obs[k][0], obs[k][1], i.dx, i.dy)
          for k, i in enumerate(inps)]
    dev = state0.device
    sidx = torch.from_numpy(np.concatenate([a for a, _ in st])).to(dev)
    swt = torch.from_numpy(np.concatenate([b for _, b in st])).to(dev)
    with torch.no_grad():
        tw = ea.forward(twin0, steps)
        ys = (tw[sidx] * swt).sum(1) + torch.from_numpy(
            rng.normal(0.0, 0.01, len(inps) * n_obs).astype(np.float32)).to(dev)

    def misfit(out):
        return A.stencil_misfit(out, sidx, swt, ys)

    # forward alone: batched eager steps
    ens.state[0].copy_(state0)
    ens.step_n(2)
    torch.cuda.synchronize()
    ens.state[0].copy_(state0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step_n(steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0

    def fwd_adj(a):
        out = a.forward(state0, steps).requires_grad_(True)
        J = misfit(out)
        lam = torch.autograd.grad(J, out)[0]
        return J, a.backward(lam)

    def timed(a):
        fwd_adj(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, grad = fwd_adj(a)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, J, grad

    # records kept from the forward run (default when they fit the budget)
    t_tot, J, grad = timed(ea)
    kept = ea.records is not None
    rec_gb = ea.records_bytes(steps) / 1e9
    ea.records = None
    # checkpointed: the backward recomputes each segment
    ea_rc = A.EnsembleAdjoint(ens, checkpoint_every=K, keep_records=False)
    t_rc, J_rc, grad_rc = timed(ea_rc)
    same = bool(torch.equal(grad, grad_rc))
    del ea_rc
    gnorm = [round(float(grad[int(ens.state_off[k]):int(ens.state_off[k]) + ens.planes[k]].norm()), 6)
             for k in range(ens.n)]
    # single-domain reverse mode, member by member (first plain_members members)
    from SWESimulators import Common, _gocean as G
    dts = ens.dt_host()
    plain_cells = 0
    plain_s = 0.0
    for k in range(plain_members):
        i = inps[k]
        d = Common.make_domain(i.nx, i.ny, i.dx, i.dy, i.g, 0.0, 1.3, 2, i.bc, pitch=ens.pitch[k],
                               hi_pitch=i.nx + 5, f=i.f)
        Hi_k = ens.Hi[int(ens.hi_off[k]):int(ens.hi_off[k + 1])].view(i.ny + 4 + 1, i.nx + 5)
        adj = A.RK2Adjoint(d, Hi_k, checkpoint_every=K)
        o, pl = int(ens.state_off[k]), ens.planes[k]
        q0 = tuple(state0[o + f * pl:o + (f + 1) * pl].view(i.ny + 4, ens.pitch[k]) for f in range(3))

        def one():
            qs = [t.clone().requires_grad_(True) for t in q0]
            out = A.CDKLM16RK2.apply(*qs, adj, steps, float(dts[k]))
            A.point_misfit(out[0], obs[k][0], obs[k][1], ys[k * n_obs:(k + 1) * n_obs], i.dx, i.dy).backward()
        one()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one()
        torch.cuda.synchronize()
        plain_s += time.perf_counter() - t0
        plain_cells += i.nx * i.ny
    batched_rate = cells * steps / t_tot / 1e6
    plain_rate = plain_cells * steps / plain_s / 1e6
    ens.cleanUp()
    return {"workload": f"32 domains, sides {{256,384,512,768}} (seeded), walls, own random depth, "
                        f"per-member dt (CFL 0.7), {steps} RK2 steps, checkpoint every {K}, "
                        f"{n_obs} point observations per domain",
            "cells": cells, "sides": [i.nx for i in inps],
            "fwd_batched_s": round(t_fwd, 4), "fwd_adj_batched_s": round(t_tot, 4),
            "fwd_adj_over_fwd": round(t_tot / t_fwd, 2), "records_kept": kept,
            "records_gb": round(rec_gb, 2),
            "fwd_adj_recompute_s": round(t_rc, 4), "fwd_adj_recompute_over_fwd": round(t_rc / t_fwd, 2),
            "kept_vs_recompute_grad_identical": same,
            "fwd_batched_mcell_steps_s": round(cells * steps / t_fwd / 1e6, 1),
            "fwd_adj_batched_mcell_steps_s": round(batched_rate, 1),
            "fwd_adj_plain_per_member_mcell_steps_s": round(plain_rate, 1),
            "fwd_adj_batched_vs_plain": round(batched_rate / plain_rate, 2),
            "plain_members_timed": plain_members,
            "J": float(J), "grad_norm_per_domain": gnorm}


def _event_time(st, fn, reps):
    """Average seconds per call of fn() on stream st (HIP events)."""
    import torch
    fn()
    st.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    st.synchronize()
    return e0.elapsed_time(e1) / 1000.0 / reps


def bench_model_error(ctx, inp, reps=50, ens_members=100, ens_n=512, ens_steps=20):
    """Model error (OceanStateNoise.perturbSim) on the headline domain: each
    kernel timed with HIP events against its algorithmic bytes, the step with
    and without per-step model error, and the batched ensemble step with and
    without it.  Algorithmic bytes: LCG generator 12 B per number (seed read
    and write 16 B per pair, two 4 B numbers); Philox 4 B per number; SOAR 8 B
    per coarse cell (read xi, write d_eta); balance 32 B per cell at factor 1
    (d_eta, Hi, eta/hu/hv read, eta/hu/hv written), 28 B + 4/factor^2 B at
    factor > 1."""
    import torch
    from SWESimulators import BathymetryAndICs as B, CDKLM16Ensemble, OceanStateNoise
    from SWESimulators import _gocean as G
    out = {}
    lib = G.load()
    inp3 = B.config2(nx=inp.nx - inp.nx % 3, ny=inp.ny - inp.ny % 3)    # sides divisible by 3
    for factor in (1, 3):
        fin = inp if factor == 1 else inp3
        sim = make_sim(fin, ctx)
        st = sim.gpu_stream
        cells = fin.nx * fin.ny
        for use_lcg in (True, False):
            nz = OceanStateNoise.OceanStateNoise.fromsim(sim, soar_q0=1e-4, interpolation_factor=factor,
                                                         use_lcg=use_lcg)
            sh = G.stream_handle(st)
            rnum = int(nz.rand_nx) * int(nz.rand_ny)
            ccells = (int(nz.coarse_nx) + 4) * (int(nz.coarse_ny) + 4)
            if use_lcg:
                gen = lambda: G.check(lib.gocean_noise_generate(
                    nz.desc, G.RNG_LCG_NORMAL, nz.seed.data_ptr(), nz.random_numbers.data.gpudata, 0, None, sh))
            else:
                gen = lambda: G.check(lib.gocean_noise_generate(
                    nz.desc, G.RNG_PHILOX_NORMAL, None, nz.random_numbers.data.gpudata, nz.rng.key,
                    nz.rng.counter.data_ptr(), sh))
            t_gen = _event_time(st, gen, reps)
            nz.perturbSim(sim)        # sets the state pitches of the descriptor
            t_soar = _event_time(st, lambda: G.check(lib.gocean_noise_soar(
                nz.desc, float(nz.soar_q0), 1.0, nz.random_numbers.data.gpudata,
                nz.coarse_buffer.data.gpudata, 0, sh)), reps)
            gd = sim.gpu_data
            t_bal = _event_time(st, lambda: G.check(lib.gocean_noise_balance(
                nz.desc, nz.coarse_buffer.data.gpudata, gd.h0.data.gpudata, gd.hu0.data.gpudata,
                gd.hv0.data.gpudata, sim.bathymetry.Bi.data.gpudata, nz.fields, 0, 0, sh)), reps)
            t_all = _event_time(st, lambda: nz.perturbSim(sim), reps)
            b_gen = (12.0 if use_lcg else 4.0) * rnum
            b_soar = 8.0 * ccells
            b_bal = (32.0 if factor == 1 else 28.0 + 4.0 / factor ** 2) * cells
            key = f"factor{factor}_{'lcg' if use_lcg else 'philox'}"
            out[key] = {
                "generate_us": round(t_gen * 1e6, 1), "generate_GBs": round(b_gen / t_gen / 1e9, 1),
                "soar_us": round(t_soar * 1e6, 1), "soar_GBs": round(b_soar / t_soar / 1e9, 1),
                "balance_us": round(t_bal * 1e6, 1), "balance_GBs": round(b_bal / t_bal / 1e9, 1),
                "perturb_us": round(t_all * 1e6, 1),
                "perturb_GBs_algorithmic": round((b_gen + b_soar + b_bal) / t_all / 1e9, 1),
                "domain": f"{fin.nx}x{fin.ny}"}
            nz.cleanUp()
        if factor == 3:
            sim.cleanUp()
    sim = make_sim(inp, ctx)
    st = sim.gpu_stream
    # RK2 step with and without per-step model error (factor 1, Philox)
    t_step = _event_time(st, lambda: sim.step_n(1), reps)
    nz = OceanStateNoise.OceanStateNoise.fromsim(sim, soar_q0=1e-4)
    t_step_me = _event_time(st, lambda: (sim.step_n(1), nz.perturbSim(sim), sim.applyBoundaryConditions()), reps)
    out["step_us"] = round(t_step * 1e6, 1)
    out["step_with_model_error_us"] = round(t_step_me * 1e6, 1)
    out["model_error_overhead"] = round(t_step_me / t_step - 1.0, 3)
    sim.cleanUp()
    # batched ensemble: step with and without model error
    inps = B.config3(members=ens_members, n=ens_n)
    ecells = sum(i.nx * i.ny for i in inps)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, inps, courant_number=0.7)
    from SWESimulators import autotune
    autotune.tune_ensemble(ens)
    ens.enable_model_error(soar_q0=1e-4, seed=1)
    est = ens.gpu_stream
    t_e = _event_time(est, lambda: ens.step_n(ens_steps), 3) / ens_steps
    t_eme = _event_time(est, lambda: ens.step_n(ens_steps, model_error=True), 3) / ens_steps
    ens.capture(ens_steps, model_error=True)
    t_g = _event_time(est, ens.replay, 3) / ens_steps
    out["ensemble"] = {
        "workload": f"{ens_members} members x {ens_n}x{ens_n}, walls, Philox, factor 1",
        "step_mcell_steps_s": round(ecells / t_e / 1e6, 1),
        "step_with_model_error_mcell_steps_s": round(ecells / t_eme / 1e6, 1),
        "step_with_model_error_graph_mcell_steps_s": round(ecells / t_g / 1e6, 1),
        "model_error_overhead": round(min(t_eme, t_g) / t_e - 1.0, 3)}
    ens.cleanUp()
    out["workload"] = (f"{inp.nx}x{inp.ny} headline domain, q0 1e-4, cutoff 2, "
                       f"interpolation factor 1 and 3, LCG and Philox generators")
    return out


def bench_etkf(ctx, members=100, n=512, n_drifters=64, reps=5):
    """Global ETKF analysis (ETKFOcean.py:74-301) of the C3 ensemble on the
    device: batched observation of every member at the true drifters, the
    ensemble-space algebra in float64 on the host, and the state update as one
    GEMM.  Algorithmic bytes: the state read twice (mean, update) and written
    once, 3 x N x 3 (ny+4)(nx+4) x 4 B."""
    import torch
    from SWESimulators import BathymetryAndICs as B, CDKLM16Ensemble, ETKFOcean
    inps = B.config3(members=members, n=n)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, inps, courant_number=0.7)
    ens.enable_drifters(n_drifters)
    rs = np.random.default_rng(9)
    pos = np.column_stack([rs.uniform(0, n * inps[0].dx, n_drifters), rs.uniform(0, n * inps[0].dy, n_drifters)])
    obs = np.column_stack([pos, 0.05 * rs.standard_normal((n_drifters, 2))])
    et = ETKFOcean.ETKFOcean(ens, np.diag([0.01, 0.01]))
    et.ETKF(obs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        et.ETKF(obs)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    state_bytes = members * 3 * (n + 4) * (n + 4) * 4
    ens.cleanUp()
    return {"workload": f"{members} members x {n}x{n}, {n_drifters} drifters (hu, hv observed)",
            "analysis_ms": round(t * 1e3, 2),
            "state_MB": round(state_bytes / 1e6, 1),
            "algorithmic_GBs": round(3 * state_bytes / t / 1e9, 1)}


def _double_jet_members(members, nx=500, ny=300, dx=2000.0, seed=12):
    """Members of the double-jet data-assimilation domain (500 x 300 cells,
    periodic east-west, walls north-south; DoubleJetCase.py:136-141) with
    synthetic random states; flat 1000 m depth."""
    rs = np.random.default_rng(seed)
    j, i = np.mgrid[0:ny + 4, 0:nx + 4]
    jet = np.exp(-((j - ny / 4) / 12.0) ** 2) - np.exp(-((j - 3 * ny / 4) / 12.0) ** 2)
    Hi = np.full((ny + 5, nx + 5), 1000.0, dtype=np.float32)
    out = []
    for _ in range(members):
        out.append(dict(nx=nx, ny=ny, dx=dx, dy=dx, g=9.81, f=1.2e-4, r=0.0, bc=(1, 2, 1, 2), Hi=Hi,
                        eta=(0.3 * jet + 0.02 * rs.standard_normal((ny + 4, nx + 4))).astype(np.float32),
                        hu=(50.0 * jet + 5.0 * rs.standard_normal((ny + 4, nx + 4))).astype(np.float32),
                        hv=(5.0 * rs.standard_normal((ny + 4, nx + 4))).astype(np.float32)))
    return out


def bench_letkf(ctx, members=100, spacing=(25, 50), r=15.0, reps=10, cpu=True):
    """Localised ETKF (ETKFOcean.LETKF, ETKFOcean.py:583-744) on the double
    jet data-assimilation setup of the reference's experiments: 100 members of
    500 x 300, static buoys every 25 cells ("all", 240 buoys in 4 groups) and
    every 50 cells ("sparse", 60 buoys), r = 15 (boxes of 47 x 47 cells).
    Algorithmic bytes per analysis: every cell with a nonzero weight, per
    group, read and written once for all members and fields
    (2 x 3 x N x 4 B).  The CPU baseline is the oracle (the reference's numpy
    algorithm: per-buoy eigh and solve on N x N matrices) on one thread."""
    import torch
    from SWESimulators import CDKLM16Ensemble, ETKFOcean
    nx, ny, dx = 500, 300, 2000.0
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, _double_jet_members(members), courant_number=0.7)
    rs = np.random.default_rng(21)
    out = {"workload": f"{members} members x {nx}x{ny} (double jet), r = {r}, buoys at cell centres"}
    for step in spacing:
        ys, xs = np.mgrid[0:ny:step, 0:nx:step]
        pos = np.column_stack([(xs.ravel() + 0.5) * dx, (ys.ravel() + 0.5) * dx])
        obs = np.column_stack([pos, 5.0 * rs.standard_normal((len(pos), 2))])
        et = ETKFOcean.ETKFOcean(ens, np.diag([1.0, 1.0]))
        et.LETKF(obs, r_factor=r)                       # geometry set-up and warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            et.LETKF(obs)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        S = et._letkf
        covered = 0
        jj, ii = np.nonzero(S["W_loc"])
        R = S["local_n"] // 2
        for g in S["groups"]:
            mask = np.zeros((ny, nx), dtype=bool)
            for o in g:
                bx = (int(np.round(pos[o, 0] / dx)) - R) % nx
                by = (int(np.round(pos[o, 1] / dx)) - R) % ny
                mask[(by + jj) % ny, (bx + ii) % nx] = True
            covered += int(mask.sum())
        alg = 2 * 3 * members * 4 * covered
        leg = {"buoys": len(pos), "groups": len(S["groups"]), "analysis_ms": round(t * 1e3, 3),
               "covered_cells": covered, "algorithmic_MB": round(alg / 1e6, 1),
               "algorithmic_GBs": round(alg / t / 1e9, 1)}
        if cpu:
            from threadpoolctl import threadpool_limits
            from oracle import etkf_np as E
            X = np.stack([np.stack(ens.download(k, interior_domain_only=True)) for k in range(members)])
            with threadpool_limits(1):
                t0 = time.perf_counter()
                E.letkf(X, obs, pos, np.diag([1.0, 1.0]), r, dx, dx)
                tc = time.perf_counter() - t0
            leg["cpu_baseline"] = {"value": round(tc * 1e3, 1), "unit": "ms per analysis", "cores": 1,
                                   "kind": "port", "sample": "the whole analysis"}
            leg["gpu_over_cpu"] = round(tc / t, 1)
        out[f"every_{step}_cells"] = leg
    ens.cleanUp()
    return out


def bench_iewpf(ctx, members=100, nx=500, ny=300, factor=5, n_drifters=64, reps=5, cpu=True, cpu_particles=2):
    """Two-stage IEWPF (IEWPFOcean.iewpf_2stage, IEWPFOcean.py:195-329) on
    100 particles of 500 x 300 with model error interpolated by 5 (the
    double-jet experiments' sizes), doubly periodic and of constant depth
    (the filter's assumptions), 64 drifters at random positions, synthetic
    states.  The Kalman-gain basis is built in the warm-up call.  The CPU
    baseline is the oracle (the reference's algorithm in numpy) on
    ``cpu_particles`` particles, per particle."""
    import torch
    from SWESimulators import CDKLM16Ensemble, IEWPFOcean
    dx = 2000.0
    rs = np.random.default_rng(31)
    Hi = np.full((ny + 5, nx + 5), 1000.0, dtype=np.float32)
    ms = [dict(nx=nx, ny=ny, dx=dx, dy=dx, g=9.81, f=1.2e-4, r=0.0, bc=(2, 2, 2, 2), Hi=Hi,
               eta=(0.05 * rs.standard_normal((ny + 4, nx + 4))).astype(np.float32),
               hu=(5.0 * rs.standard_normal((ny + 4, nx + 4))).astype(np.float32),
               hv=(5.0 * rs.standard_normal((ny + 4, nx + 4))).astype(np.float32)) for _ in range(members)]
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, courant_number=0.7)
    q0 = 5e-3
    ens.enable_model_error(soar_q0=q0, interpolation_factor=factor, seed=3)
    ens.enable_drifters(n_drifters)
    pos = np.column_stack([rs.uniform(0, nx * dx, n_drifters), rs.uniform(0, ny * dx, n_drifters)])
    R = np.eye(2) * 0.5 ** 2
    iw = IEWPFOcean.IEWPFOcean(ens, R)
    HX = iw.observeParticles(pos)
    obs = np.column_stack([pos, HX.mean(axis=0) + 0.5 * rs.standard_normal((n_drifters, 2))])
    t0 = time.perf_counter()
    iw.iewpf_2stage(obs)                                # builds the Kalman-gain basis
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(reps):
        iw.iewpf_2stage(obs)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    out = {"workload": f"{members} particles x {nx}x{ny}, model error interpolated by {factor}, "
                       f"{n_drifters} drifters, doubly periodic, constant depth",
           "analysis_ms": round(t * 1e3, 2), "per_particle_ms": round(t * 1e3 / members, 4),
           "first_call_with_basis_ms": round(t_setup * 1e3, 1)}
    if cpu:
        from oracle import iewpf_np as I
        p = I.Params(nx=nx, ny=ny, dx=dx, dy=dx, factor=factor, q0=q0, L=0.75 * np.float32(dx * factor),
                     f=1.2e-4, g=9.81, H=1000.0, R=R)
        states = np.stack([np.stack(ens.download(k)) for k in range(cpu_particles)]).astype(np.float64)
        cn = (nx // factor, ny // factor)
        xi = [rs.standard_normal((cn[1], cn[0])) for _ in range(cpu_particles)]
        nu = [rs.standard_normal((cn[1], cn[0])) for _ in range(cpu_particles)]
        inn = 0.5 * rs.standard_normal((cpu_particles, n_drifters, 2))
        t0 = time.perf_counter()
        I.iewpf_2stage(p, iw.S_host, iw.localSVD_host, states, Hi.astype(np.float64), pos, inn, xi, nu)
        tc = (time.perf_counter() - t0) / cpu_particles
        out["cpu_baseline"] = {"value": round(tc * 1e3, 1), "unit": "ms per particle", "cores": 1, "kind": "port",
                               "sample": f"the oracle on {cpu_particles} particles of the same workload"}
        out["gpu_over_cpu_per_particle"] = round(tc / (t / members), 1)
    ens.cleanUp()
    return out


def _profile_json():
    p = os.path.join(ROOT, "profiles", "cdklm16_substep_2048_counters.json")
    try:
        return json.load(open(p))
    except (OSError, ValueError):
        return {}


def measured_traffic(cells, tile):
    """HBM bytes per substep launch of tile configuration ``tile`` at this
    domain size, from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a plain copy of
    known size); None when no summary covers this tile and size."""
    d = _profile_json().get("tiles", {}).get(tile)
    if not d or d.get("cells") != cells:
        return None
    return d.get("bytes_per_launch")


def measured_valu(cells, tile, launch_s):
    """VALU work of the substep launch from the committed counter summary
    (SQ_INSTS_VALU per dispatch) and its issue rate against the nominal
    non-packed FP32 rate: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave
    instruction = 614.4 G wave instructions/s (packed FP32 is off in this
    build).  Moves, selects and bit operations of the kernel's mix do not
    all take an FP32 issue slot, so the fraction can exceed 1."""
    d = _profile_json().get("tiles", {}).get(tile)
    if not d or d.get("cells") != cells or "valu_per_launch" not in d:
        return None
    wi = d["valu_per_launch"]
    return {"valu_wave_instructions_per_launch": wi,
            "valu_lane_ops_per_cell": round(wi * 64 / cells, 1),
            "issue_rate_g_per_s": round(wi / launch_s / 1e9, 1),
            "issue_frac_of_614g": round(wi / launch_s / 614.4e9, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=None)
    ap.add_argument("--n", type=int, default=2048, help="domain side (headline: 2048)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-c5", action="store_true", help="skip the batched uneven ensemble adjoint")
    ap.add_argument("--no-model-error", action="store_true", help="skip the model-error leg")
    ap.add_argument("--letkf-only", action="store_true",
                    help="run only the LETKF leg and print its JSON (for profiling it)")
    ap.add_argument("--iewpf-only", action="store_true",
                    help="run only the IEWPF leg and print its JSON (for profiling it)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the CDKLM16 path "
              f"is built and measured on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch
    from SWESimulators import BathymetryAndICs as B, Common

    ctx = Common.HIPContext(0)
    log(f"[bench] device {torch.cuda.get_device_name(0)}")
    if a.letkf_only:
        print(json.dumps({"letkf": bench_letkf(ctx, cpu=not a.no_cpu)}), flush=True)
        return
    if a.iewpf_only:
        print(json.dumps({"iewpf": bench_iewpf(ctx, cpu=not a.no_cpu)}), flush=True)
        return
    inp = B.config2(nx=a.n, ny=a.n)
    log(f"[bench] workload {inp.name}: {inp.nx}x{inp.ny}, steps {a.steps}, warmup {a.warmup}")
    sim = make_sim(inp, ctx, tile_config=a.tile)
    if a.tile is None:
        from SWESimulators import autotune
        best, times = autotune.tune_simulator(sim)
        names = autotune.tile_config_names()
        log("[bench] autotune ms/step: " + ", ".join(f"{names[c]}={t:.4f}" for c, t in times.items())
            + f" -> {names[best]}")
    wall, ev, launches = time_plain(sim, a.steps, a.warmup)
    if a.calib_copy_mib:
        n = a.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {a.calib_copy_mib} MiB read + {a.calib_copy_mib} MiB written per launch")
    cells = inp.nx * inp.ny
    value = cells * a.steps / wall / 1e6
    avg_launch = ev / launches
    from SWESimulators import autotune
    from SWESimulators import _gocean as G
    cfg = sim.tile_config if sim.tile_config >= 0 else G.load().gocean_default_tile_config(inp.nx, inp.ny)
    tile_name = autotune.tile_config_names()[cfg]
    achieved = 0.5 * BYTES_PER_CELL_STEP * cells / avg_launch / 1e9
    log(f"[bench] plain: {value:.1f} Mcell-steps/s, event time/launch {avg_launch*1e6:.1f} us")

    iso = time_isolated_step(sim)
    log(f"[bench] one isolated RK2 step: {iso*1e6:.1f} us ({iso*0.5e6:.1f} us per launch)")
    extra = {"substep_launch_us": round(avg_launch * 1e6, 2),
             "isolated_step_launch_us": round(iso * 0.5e6, 2),
             "valu": measured_valu(cells, tile_name, avg_launch)}
    if not a.no_ensemble:
        log("[bench] batched ensemble C3")
        extra["ensemble"] = bench_ensemble(ctx)
        log(f"[bench] ensemble: {extra['ensemble']}")
        # many small members: one-member-per-launch is launch / host bound
        # there, which is where batching changes the picture
        log("[bench] batched ensemble, 1024 members of 64x64")
        extra["ensemble_small_members"] = bench_ensemble(ctx, members=1024, n=64, steps=50, per_member_steps=2)
        log(f"[bench] ensemble small: {extra['ensemble_small_members']}")

    if not a.no_adjoint:
        log("[bench] adjoint C4")
        extra["adjoint"] = bench_adjoint(ctx)
        log(f"[bench] adjoint: {extra['adjoint']}")
        if not a.no_c5:
            log("[bench] batched uneven ensemble with adjoint C5")
            extra["ensemble_adjoint"] = bench_c5(ctx)
            log(f"[bench] ensemble adjoint: {extra['ensemble_adjoint']}")

    if not a.no_ensemble:
        log("[bench] ETKF analysis on the C3 ensemble")
        extra["etkf"] = bench_etkf(ctx)
        log(f"[bench] etkf: {extra['etkf']}")
        log("[bench] LETKF analysis on the double-jet ensemble")
        extra["letkf"] = bench_letkf(ctx, cpu=not a.no_cpu)
        log(f"[bench] letkf: {extra['letkf']}")
        log("[bench] two-stage IEWPF on a 100-particle ensemble")
        extra["iewpf"] = bench_iewpf(ctx, cpu=not a.no_cpu)
        log(f"[bench] iewpf: {extra['iewpf']}")

    if not a.no_model_error:
        log("[bench] model error (OceanStateNoise) on the headline domain and the ensemble")
        extra["model_error"] = bench_model_error(ctx, inp)
        log(f"[bench] model error: {extra['model_error']}")

    cpu = None
    if not a.no_cpu and not a.no_ensemble:
        log("[bench] cpu baseline, C3 members over a process pool")
        extra["cpu_baseline_ensemble"] = cpu_baseline_ensemble()
        log(f"[bench] cpu ensemble: {extra['cpu_baseline_ensemble']}")
    if not a.no_cpu and not a.no_adjoint:
        log("[bench] cpu baseline, C4 forward + adjoint (float64 torch, autograd)")
        extra["cpu_baseline_adjoint"] = cpu_baseline_adjoint()
        log(f"[bench] cpu adjoint: {extra['cpu_baseline_adjoint']}")
    if not a.no_cpu:
        log("[bench] cpu baseline (oracle, 1 core, 1 RK2 step)")
        v, s = cpu_baseline(inp, 1)
        cpu = {"value": round(v, 3), "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
               "sample": f"NumPy fp32 oracle, 1 RK2 step of the {inp.nx}x{inp.ny} workload "
                         f"({s:.1f} s), {platform.processor() or platform.machine()}, "
                         f"numpy {np.__version__}"}

    out = {
        "metric": "Mcell-steps/s CDKLM16 fwd & fwd+adjoint, batched vs plain; % of HBM peak",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random depth, balanced eddies; no datasets)",
        "config": {"workload": f"CDKLM16 plain {inp.nx}x{inp.ny} periodic-EW walls-NS, "
                               f"dx=dy=500 m, fixed dt (CFL 0.7), SSP-RK2",
                   "nx": inp.nx, "ny": inp.ny, "cells": cells, "rk_order": 2,
                   "tile_config": tile_name},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(cells, tile_name)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
