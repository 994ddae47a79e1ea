"""The batched ensemble launch (members of uneven sizes and boundary
conditions in one launch per substep) against the oracle, member by member,
bit-exact; device dt; HIP-graph replays."""

import numpy as np
import pytest

from oracle import cdklm16_np as O
import helpers as H

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def members(seed=0, sizes=((37, 29), (64, 48), (50, 50), (33, 71)),
            bcs=((1, 1, 1, 1), (2, 2, 2, 2), (1, 2, 1, 2), (2, 1, 2, 1))):
    rng = np.random.default_rng(seed)
    out = []
    for k, ((nx, ny), bc) in enumerate(zip(sizes, bcs)):
        y, x = np.mgrid[0:ny + 5, 0:nx + 5]
        Hi = (35 + 8 * np.sin(x / (4.0 + k)) + 6 * np.cos(y / 3.0) + rng.random((ny + 5, nx + 5))).astype(np.float32)
        eta = (0.3 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
        hu = (2 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
        hv = (2 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
        out.append(dict(nx=nx, ny=ny, dx=300.0 + 10 * k, dy=250.0, g=9.81, f=1.0e-4 * (1 + k), r=0.0,
                        bc=bc, eta=eta, hu=hu, hv=hv, Hi=Hi, wind_x=1e-5 * k, wind_y=-2e-5))
    return out


def oracle_member(m, rk_order=2):
    p = O.Params(nx=m["nx"], ny=m["ny"], dx=m["dx"], dy=m["dy"], g=m["g"], r=m["r"], bc=m["bc"],
                 f=np.float32(m["f"]), rk_order=rk_order,
                 wind_x=(m["wind_x"], m["wind_x"]), wind_y=(m["wind_y"], m["wind_y"]))
    return O.OracleSim(m["eta"], m["hu"], m["hv"], m["Hi"], p, dt=1.0), p


@pytest.mark.parametrize("rk_order", [1, 2, 3])
@pytest.mark.parametrize("tile", [None, 0, 3])
def test_ensemble_matches_oracle(rk_order, tile):
    from SWESimulators import Common, CDKLM16Ensemble
    _say("ensemble vs oracle, rk", rk_order, "tile", tile)
    ms = members()
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=rk_order, tile_config=tile, courant_number=0.7)
    # device dt = 0.7 * CFL minimum per member (fp32)
    dts = ens.dt_host()
    sims = []
    for k, m in enumerate(ms):
        sim, p = oracle_member(m, rk_order)
        ref = np.float32(np.float32(0.7) * O.max_dt(*sim.h0, sim.Hm, p))
        assert dts[k] == ref, (k, dts[k], ref)
        sims.append(sim)
    n = 4
    report = []
    for step in range(n // (2 if rk_order == 1 else 1)):
        ens.step_n(2 if rk_order == 1 else 1)
        for k, sim in enumerate(sims):
            sim.step_n(2 if rk_order == 1 else 1, dts[k])
            got = ens.download(k)
            for g, r, name in zip(got, sim.h0, ("eta", "hu", "hv")):
                g, r = g[2:-2, 2:-2], r[2:-2, 2:-2]
                if not np.array_equal(g, r):
                    report.append((step, k, name, int(np.sum(g != r)), float(np.max(np.abs(g - r)))))
    ens.cleanUp()
    assert not report, report[:8]


def test_graph_replays_match_eager():
    """Replay a captured graph several times with eager work in between;
    every replay must equal the eager result."""
    import torch
    from SWESimulators import Common, CDKLM16Ensemble
    _say("ensemble graph replays")
    ms = members(seed=3)
    ctx = Common.HIPContext(0)
    a = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms)
    b = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms)
    b.capture(3)
    for rep in range(3):
        a.step_n(3)
        b.replay()
        x = torch.ones(1000, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        assert float(x) == 1000.0
        for k in range(len(ms)):
            for u, v in zip(a.download(k), b.download(k)):
                assert np.array_equal(u, v), f"replay {rep} member {k}"
    a.cleanUp()
    b.cleanUp()


def test_ensemble_matches_single_domain_path():
    """Member k of the ensemble equals the drop-in single-domain class."""
    from SWESimulators import Common, CDKLM16, CDKLM16Ensemble, WindStress
    _say("ensemble vs single-domain class")
    ms = members(seed=5)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms)
    dts = ens.dt_host()
    ens.step_n(5)
    for k, m in enumerate(ms):
        w = WindStress.WindStress(t=[0.0], X=[np.array([[m["wind_x"]]], np.float32)],
                                  Y=[np.array([[m["wind_y"]]], np.float32)])
        sim = CDKLM16.CDKLM16(ctx, m["eta"], m["hu"], m["hv"], m["Hi"], m["nx"], m["ny"], m["dx"], m["dy"],
                              float(dts[k]), m["g"], m["f"], m["r"],
                              boundary_conditions=Common.BoundaryConditions(*m["bc"]), wind_stress=w)
        sim.step_n(5)
        for u, v in zip(ens.download(k), sim.download()):
            assert np.array_equal(u[2:-2, 2:-2], v[2:-2, 2:-2])
    ens.cleanUp()


def test_ensemble_bathymetry_and_one_substep():
    """Hi halo, Hm and one substep of every member through the ensemble C ABI."""
    from SWESimulators import Common, CDKLM16Ensemble
    _say("ensemble bathymetry + one substep")
    ms = members(seed=11)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, dt=0.5)
    ens.gpu_stream.synchronize()
    Hi_all = ens.Hi.cpu().numpy()
    Hm_all = ens.Hm.cpu().numpy()
    for k, m in enumerate(ms):
        sim, p = oracle_member(m)
        nx, ny = m["nx"], m["ny"]
        Hi = Hi_all[ens.hi_off[k]:ens.hi_off[k + 1]].reshape(ny + 5, nx + 5)
        Hm = Hm_all[ens.hm_off[k]:ens.hm_off[k + 1]].reshape(ny + 4, nx + 4)
        assert np.array_equal(Hi, sim.Hi), f"member {k} Hi"
        assert np.array_equal(Hm, sim.Hm), f"member {k} Hm"
        e, u, v = ens.download(k)
        for a, b in zip((e, u, v), sim.h0):
            assert np.array_equal(a[2:-2, 2:-2], b[2:-2, 2:-2]), f"member {k} initial state"
    G = CDKLM16Ensemble.G
    G.check(ens.lib.gocean_ensemble_substep(ens.handle, 0, ens.state[0].data_ptr(), None,
                                            ens.state[1].data_ptr(), ens.Hi.data_ptr(), None,
                                            ens.dt.data_ptr(), G.stream_handle(ens.gpu_stream)), "substep")
    ens.gpu_stream.synchronize()
    for k, m in enumerate(ms):
        sim, p = oracle_member(m)
        O.periodic_bc(*sim.h0, p)          # the kernel wraps; the oracle reads ghosts
        ref = O.cdklm_substep(np.float32(0.5), 0, *sim.h0, *sim.h1, sim.Hi, sim.Hm, p)
        got = [x.cpu().numpy() for x in ens.views(k, 1)]
        for g, r, name in zip(got, ref, ("eta", "hu", "hv")):
            d = g[2:-2, 2:-2] != r[2:-2, 2:-2]
            assert not d.any(), f"member {k} {name}: {int(d.sum())} cells differ, first {np.argwhere(d)[:3].tolist()}"
    ens.cleanUp()


def test_tile_config_switch_and_tuner():
    """Every tile configuration of the batched substep gives the same bits
    (one ensemble switched between them with gocean_ensemble_set_tile_config),
    and the ensemble autotuner leaves the state unchanged and keeps the
    fastest configuration."""
    import torch
    from SWESimulators import Common, CDKLM16Ensemble, autotune
    _say("ensemble tile switch and tuner")
    ms = members(seed=11)
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=2, courant_number=0.7)
    start = [s.clone() for s in ens.state]
    ref = None
    for c in range(len(autotune.tile_config_names())):
        for s, s0 in zip(ens.state, start):
            s.copy_(s0)
        ens.set_tile_config(c)
        ens.step_n(3)
        torch.cuda.synchronize()
        if ref is None:
            ref = ens.state[0].clone()
        else:
            assert torch.equal(ens.state[0], ref), c
    before = ens.state[0].clone()
    best, times = autotune.tune_ensemble(ens, use_cache=False)
    torch.cuda.synchronize()
    assert torch.equal(ens.state[0], before)
    assert ens.tile_config == best == min(times, key=times.get)
    ens.capture(2)
    ens.replay()
    ens.step_n(0)
    torch.cuda.synchronize()
    ens.cleanUp()


def test_data_assimilation_step_matches_each_member():
    """CDKLM16Ensemble.dataAssimilationStep equals every member's own
    CDKLM16.dataAssimilationStep (CDKLM16.py:650-704) bit for bit: members of
    different sizes, depths and boundaries have different CFL dt, so their
    step counts per model step differ and finished members sit out launches
    (dt 0).  Two calls: the second crosses the dt refresh every 5 model
    steps."""
    from SWESimulators import Common, CDKLM16, CDKLM16Ensemble, WindStress
    _say("ensemble dataAssimilationStep vs members")
    ms = members(seed=7)
    for k, m in enumerate(ms):                  # depths, hence CFL dt, far apart
        m["Hi"] = (m["Hi"] * (1 + 2 * k)).astype(np.float32)
    mts = 10.0
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, courant_number=0.8)
    sims = []
    for m in ms:
        w = WindStress.WindStress(t=[0.0], X=[np.array([[m["wind_x"]]], np.float32)],
                                  Y=[np.array([[m["wind_y"]]], np.float32)])
        sims.append(CDKLM16.CDKLM16(ctx, m["eta"], m["hu"], m["hv"], m["Hi"], m["nx"], m["ny"], m["dx"], m["dy"],
                                    1.0, m["g"], m["f"], m["r"],
                                    boundary_conditions=Common.BoundaryConditions(*m["bc"]), wind_stress=w,
                                    model_time_step=mts))
    for obs_time in (25.0, 71.0):
        ens.dataAssimilationStep(obs_time, mts, courant_number=0.8)
        for sim in sims:
            sim.dataAssimilationStep(obs_time, courant_number=0.8)
        steps = [len(q) for q in ens._dt_sequences(ens.dt64, mts)]
        _say("  obs time", obs_time, "steps per model step", steps)
        assert len(set(steps)) > 1, "members should differ in step count"
        for k, sim in enumerate(sims):
            assert ens.t[k] == sim.t, (k, ens.t[k], sim.t)
            for u, v in zip(ens.download(k), sim.download()):
                assert np.array_equal(u[2:-2, 2:-2], v[2:-2, 2:-2]), k
    ens.cleanUp()


def test_data_assimilation_step_with_model_error_runs():
    """With the batched model error every member reaches the observation time
    and the perturbation changed the state (scale and order follow the
    reference's schedule; the random streams differ from a per-member run)."""
    from SWESimulators import Common, CDKLM16Ensemble
    ms = members(seed=8, bcs=((2, 2, 2, 2),) * 4)
    ctx = Common.HIPContext(0)
    a = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, courant_number=0.8)
    b = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, courant_number=0.8)
    a.enable_model_error(soar_q0=1e-3, seed=3)
    a.dataAssimilationStep(37.0, 10.0)
    b.dataAssimilationStep(37.0, 10.0)
    assert np.all(a.t == 37.0) or np.allclose(a.t, 37.0, rtol=0, atol=1e-9)
    assert np.array_equal(a.t, b.t)
    assert not np.array_equal(a.download(0)[0], b.download(0)[0])
    a.cleanUp()
    b.cleanUp()


def test_reference_getters_and_innovations():
    """The getters DA code calls on the reference's ensembles, and
    getInnovations = y - H(x_m) from the batched observation kernel."""
    from SWESimulators import Common, CDKLM16Ensemble
    from oracle import drifters_np as DR
    ms = members(seed=9, sizes=((40, 30),) * 3, bcs=((2, 2, 2, 2),) * 3)
    for m in ms:
        m["dx"] = 300.0
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms)
    ens.enable_drifters(4)
    assert ens.getNumParticles() == 3 and ens.getNumActiveParticles() == 3 and ens.getNumDrifters() == 4
    assert (ens.getNx(), ens.getNy(), ens.getDx(), ens.getDy()) == (40, 30, np.float32(300.0), np.float32(250.0))
    assert ens.getDomainSizeX() == 40 * 300.0 and ens.particlesActive == [True] * 3
    pos = np.array([[1000.0, 900.0], [5000.0, 3000.0], [11000.0, 7000.0], [300.0, 7400.0]])
    obs = np.column_stack([pos, np.arange(8.0).reshape(4, 2)])
    inn = ens.getInnovations(obs)
    for k in range(3):
        eta, hu, hv = ens.download(k)
        want = obs[:, 2:4] - DR.observe_flow(pos.astype(np.float32), hu, hv, 300.0, 250.0, 4)
        np.testing.assert_allclose(inn[k], want, rtol=0, atol=1e-6)
    w = ens.getGaussianWeight(inn, np.eye(2))
    assert abs(w.sum() - 1) < 1e-12
    ens.cleanUp()


def test_rk_order_1_ensemble_steps_in_pairs():
    """rk_order 1 advances in pairs of Euler steps; single steps (step_to,
    dataAssimilationStep, odd step_n, per-step model error) raise a clear
    error instead of failing inside the library or leaving stale buffers."""
    from SWESimulators import Common, CDKLM16Ensemble
    _say("rk_order 1 ensemble")
    ms = members()[:2]
    ctx = Common.HIPContext(0)
    ens = CDKLM16Ensemble.CDKLM16Ensemble(ctx, ms, rk_order=1, courant_number=0.7)
    with pytest.raises(ValueError):
        ens.step_n(3)
    with pytest.raises(ValueError):
        ens.step_n(2, model_error=True)
    with pytest.raises(NotImplementedError):
        ens.step_to(5.0)
    with pytest.raises(NotImplementedError):
        ens.dataAssimilationStep(20.0, 10.0)
    ens.step_n(2)
    sim, p = oracle_member(ms[0], 1)
    dts = ens.dt_host()
    sim.step_n(2, dts[0])
    for g, r in zip(ens.download(0), sim.h0):
        assert np.array_equal(g[2:-2, 2:-2], r[2:-2, 2:-2])
    ens.cleanUp()
