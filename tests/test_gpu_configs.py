"""The benchmark workloads and the reference's own regression runs on the GPU,
at their real sizes, through the drop-in classes.

* the 13 cases of gpu_ocean/tests/schemes/CDKLM16_test.py:154-421 through
  ``CDKLM16.step(50.0)`` (55 steps of 0.9 s and one of 0.5 s): bit-exact
  (every value of normal magnitude: the reference's --ftz=true build flushes
  subnormals, and the initial states have subnormal Gaussian tails) against
  the oracle's host loop (oracle/cdklm16_np.py ``OracleSim.step``) and
  within the reference's own tolerance of its golden files;
* 2048 x 2048 plain (periodic east/west, walls north/south, random depth,
  20 balanced eddies, wind): one step bit-exact at full size, 40 more steps
  checked by invariants, then a second bit-exact step from the developed state;
* 100 members of 512 x 512 in the batched ensemble with per-step model error:
  four members against the oracle step by step (the step bit-exact, the
  random numbers and the balanced perturbation to the model-error tolerances),
  every member against the single-domain path bit for bit;
* 512 x 512 adjoint with 64 point observations, two checkpoint segments of 50
  steps: the gradient against the float64 reverse pass along the same float32
  trajectory, 16 seeded tangent-linear directions by float64 central
  differences, and the eps = 1e-3 differences of the float64 run printed;
* 32 domains with sides from {256, 384, 512, 768}: the batched forward +
  adjoint (200 steps, checkpoint every 25) bit-identical to the single-domain
  adjoint of every domain, and two domains (256 and 768) against the float64
  reverse pass.

The float64 references run as torch operations (oracle/cdklm16_torch.py) on the
GPU, so that the full-size cases finish in seconds; they are the same
restatement the small-domain adjoint tests use on the CPU.
"""

import contextlib

import numpy as np
import pytest
import torch

from oracle import cdklm16_np as O
from oracle import cdklm16_torch as T
from oracle import ocean_noise_np as ON
import helpers as H

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _mods():
    from SWESimulators import BathymetryAndICs as B, CDKLM16, CDKLM16Ensemble, Common, adjoint as A
    from SWESimulators import _gocean as G
    return B, CDKLM16, CDKLM16Ensemble, Common, A, G


def _interior_equal(a, b):
    return np.array_equal(np.asarray(a)[2:-2, 2:-2], np.asarray(b)[2:-2, 2:-2])


# ---------------------------------------------------------------------------
# the reference's regression runs through the class's step loop
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", H.ref_cases(), ids=[c[0] for c in H.ref_cases()])
def test_reference_regression_case_through_class_step(case):
    """CDKLM16(...).step(50.0) as the reference's test runs it: the loop's
    float32 local_dt (55 x 0.9 s, then 0.5 s), runs of equal dt issued as one
    library call, the t == 0 ghost exchange; bit-exact against the oracle's
    host loop and within the reference's tolerance of the golden runs
    (checkResults, CDKLM16_test.py:114-150: signed max error < 5e-4 of
    max|ref|; the oracle itself is within 1e-4)."""
    B, CDKLM16, _, Common, _, _ = _mods()
    name, bc, ic, stem, f, beta, bathy = case
    _say("regression case through the class", name)
    eta, hu, hv, Hi = H.ref_initial(ic, bathy)
    ctx = Common.HIPContext(0)
    sim = CDKLM16.CDKLM16(ctx, eta, hu, hv, Hi, H.REF_NX, H.REF_NY, H.REF_DX, H.REF_DY, H.REF_DT,
                          9.81, f, 0.0, coriolis_beta=beta,
                          boundary_conditions=Common.BoundaryConditions(*bc))
    t = sim.step(H.REF_T)
    assert sim.num_iterations == 56
    assert t == pytest.approx(50.0, abs=1e-9)
    got = sim.download()

    if beta:
        # the class builds the current constructor's beta-plane field
        # (CDKLM16.py:146-148, :295-302); the oracle runs on the same field
        f_field = sim._fields_host[0]
        assert np.allclose(f_field, H.current_beta_field(f, beta), rtol=1e-6, atol=0)
        p = O.Params(nx=H.REF_NX, ny=H.REF_NY, dx=H.REF_DX, dy=H.REF_DY, bc=bc, f=f_field)
    else:
        p = O.Params(nx=H.REF_NX, ny=H.REF_NY, dx=H.REF_DX, dy=H.REF_DY, bc=bc, f=np.float32(f))
    ref = O.OracleSim(eta, hu, hv, Hi, p, dt=H.REF_DT)
    ref.step(H.REF_T)
    assert ref.num_iterations == 56
    # the initial bump's tails are subnormal: compare in the normal range
    # (H.normal_equal), as the reference's --ftz=true build sees them
    for g, r, k in zip(got, ref.download(), ("eta", "hu", "hv")):
        if bc == (1, 1, 1, 1):
            ok = H.normal_equal(g[2:-2, 2:-2], r[2:-2, 2:-2])      # wall ghosts are never rewritten
        else:
            ok = H.normal_equal(g, r)
        assert ok, (k, float(np.max(np.abs(g[2:-2, 2:-2] - r[2:-2, 2:-2]))))

    gold = H.load_goldens()
    err = H.golden_errors(got, gold, stem)
    _say("   vs golden (signed, max abs, L2) / max|ref|:", err)
    if beta:
        # the golden was written with an older Coriolis convention; the
        # current constructor's field misses it by ~4e-2 (the oracle test
        # shows the same)
        assert err["eta"][1] > 1e-2
    else:
        for k, (signed, absmax, l2) in err.items():
            assert abs(signed) < 5e-4, (k, err)
            assert absmax < 1e-4, (k, err)


# ---------------------------------------------------------------------------
# 2048 x 2048 plain
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("tile", [1, 9])
def test_config2_full_size(tile):
    """2048^2, periodic E/W + walls N/S, random depth, eddies, wind 0.1 N/m^2
    at 30 degrees, CFL 0.7, with the 2-D tile kernel (32x16/256) and the
    column-marching kernel with balanced rows per strip (march60xauto/256):
    the device dt equals the oracle's; one RK2 step bit-exact over the whole
    padded arrays; 40 more steps keep the periodic seam exact and the mass to
    rounding; one more step from that state is bit-exact again."""
    B, _, _, _, _, _ = _mods()
    _say("config 2 at 2048x2048, tile", tile)
    c2 = B.config2()
    p = H.oracle_params(c2)
    sim = O.OracleSim(c2.eta, c2.hu, c2.hv, c2.Hi, p, dt=-1)
    D = H.DeviceDomain(c2)
    D.bathymetry_init()
    assert np.array_equal(D.Hm[:, :c2.nx + 4].cpu().numpy(), sim.Hm)
    h0 = D.state(*sim.h0)
    h1 = D.state(*sim.h1)
    D.periodic_bc(h0)
    m = D.max_dt(h0, courant=0.7)
    dt = np.float32(np.float32(0.7) * np.float32(O.max_dt(*sim.h0, sim.Hm, p)))
    assert m[1] == dt
    _say("   dt", float(dt), "; oracle step (about 10 s on one core)")
    sim.step_n(1, dt)
    D.rk_steps(1, dt, h0, h1, tile=tile)
    for g, r, k in zip(D.host(h0), sim.h0, ("eta", "hu", "hv")):
        assert H.max_rel(g[2:-2, 2:-2], r[2:-2, 2:-2]) <= 1e-5
        assert H.normal_equal(g, r), (k, int(np.sum(g != r)), H.diff_report(g, r))

    nx, ny = c2.nx, c2.ny
    mass0 = float(np.sum(sim.h0[0][2:-2, 2:-2].astype(np.float64)))
    n_more = 40
    D.rk_steps(n_more, dt, h0, h1, tile=tile)
    e, u, v = D.host(h0)
    for a in (e, u, v):
        assert np.all(np.isfinite(a))
        # periodic east/west: the ghost columns are the images of the
        # columns nx, nx+1 and 2, 3, written by the substep itself
        assert np.array_equal(a[2:-2, 0:2], a[2:-2, nx:nx + 2])
        assert np.array_equal(a[2:-2, nx + 2:nx + 4], a[2:-2, 2:4])
    mass1 = float(np.sum(e[2:-2, 2:-2].astype(np.float64)))
    # no mass crosses a mirrored wall face or the periodic seam; what is left
    # is float32 rounding of every cell's update (a random walk)
    ulp = float(np.spacing(np.float32(np.max(np.abs(e[2:-2, 2:-2])))))
    bound = 10.0 * np.sqrt(nx * ny * n_more) * ulp
    _say("   mass change over", n_more, "steps:", mass1 - mass0, "bound", bound)
    assert abs(mass1 - mass0) <= bound
    assert np.max(np.abs(e[2:-2, 2:-2])) < 1.0

    p2 = H.oracle_params(c2)
    sim2 = O.OracleSim(e, u, v, c2.Hi, p2, dt=float(dt))
    sim2.step_n(1, dt)
    D.rk_steps(1, dt, h0, h1, tile=tile)
    for g, r, k in zip(D.host(h0), sim2.h0, ("eta", "hu", "hv")):
        assert H.normal_equal(g, r), (k, int(np.sum(g != r)), H.diff_report(g, r))


# ---------------------------------------------------------------------------
# 100 x 512 x 512 batched ensemble with per-step model error
# ---------------------------------------------------------------------------
def test_config3_ensemble_with_model_error():
    """All 100 members advance in batched launches with the model error after
    every step.  Members 0, 33, 66, 99: each RK2 step bit-exact against the
    oracle, the Philox normals within 4e-6 of the float64 generator and the
    balanced SOAR perturbation within 1e-5 (of its largest value) of the
    oracle fed with the device's random numbers.  Every member: the batched
    step equals the single-domain library path bit for bit.  A second
    ensemble with the same key through step_n(2, model_error=True) lands on
    the same state."""
    B, _, CE, Common, _, _ = _mods()
    _say("config 3: 100 members of 512x512 with model error")
    inps = B.config3()
    ctx = Common.HIPContext(0)
    ens = CE.CDKLM16Ensemble(ctx, inps, courant_number=0.7)
    ens.enable_model_error(soar_q0=1e-4, seed=1)
    dts = ens.dt_host()
    n = inps[0].nx
    check = (0, 33, 66, 99)
    Hi = {k: ens.Hi[int(ens.hi_off[k]):int(ens.hi_off[k + 1])].view(n + 5, n + 5).cpu().numpy() for k in check}
    cnx, cny, rnx, rny, *_ = ON.noise_sizes(n, n, 1, False, False)
    singles = {}

    def single(k):
        if k not in singles:
            D = H.DeviceDomain(inps[k])
            D.bathymetry_init()
            singles[k] = D
        return singles[k]

    for step in range(2):
        before = [ens.download(k) for k in range(ens.n)]
        counter = int(ens.noise_counter.item())
        ens.step_n(1)
        after_step = [ens.download(k) for k in range(ens.n)]
        for k in range(ens.n):
            # the batched step against the single-domain library path
            D = single(k) if k in check else H.DeviceDomain(inps[k])
            if k not in check:
                D.bathymetry_init()
            h0 = D.state(*before[k])
            h1 = D.state(*before[k])
            D.rk_steps(1, dts[k], h0, h1)
            for g, r in zip(after_step[k], D.host(h0)):
                assert _interior_equal(g, r), ("member", k, "step", step)
        for k in check:
            p = H.oracle_params(inps[k])
            sim = O.OracleSim(*before[k], inps[k].Hi, p, dt=float(dts[k]))
            sim.step_n(1, np.float32(dts[k]))
            for g, r in zip(after_step[k], sim.h0):
                assert H.normal_equal(g[2:-2, 2:-2], r[2:-2, 2:-2]), ("member", k, "oracle step", step)
        ens.perturb()
        after = {k: ens.download(k) for k in check}
        for k in check:
            rnd = ens.noise_views(k)[0].cpu().numpy()
            want = ON.philox_generate(rnx, rny, ens.noise_key, counter, member=k, normal=True)
            assert np.max(np.abs(rnd - want)) <= 4e-6
            d = ens.noise_members[k].noise
            coarse = ON.soar_coarse(rnd, cnx, cny, d.dx, d.dy, d.soar_q0, d.soar_L, 1.0, False, False)
            base = [np.asarray(a, np.float64) for a in after_step[k]]
            want = ON.geostrophic_balance(coarse, *base, Hi[k], n, n, d.dx, d.dy, d.g, d.f)
            I = (slice(2, -2), slice(2, -2))
            for name, g, w, b0 in zip(("eta", "hu", "hv"), after[k], want, base):
                # the perturbation to 1e-5 of its largest value, plus the
                # float32 rounding of the perturbed state (eta ~ 0.3 m
                # against a perturbation ~ 1e-3 m)
                dg, dw = g[I] - b0[I], w[I] - b0[I]
                err = float(np.max(np.abs(dg - dw)))
                bound = 1e-5 * float(np.max(np.abs(dw))) + float(np.spacing(np.max(np.abs(g[I]))))
                assert err <= bound, ("member", k, name, err, bound)
        _say("   step", step + 1, "checked")
    final = [ens.download(k) for k in range(ens.n)]

    # the same run through step_n(2, model_error=True) on a fresh ensemble
    ens2 = CE.CDKLM16Ensemble(ctx, inps, courant_number=0.7)
    ens2.enable_model_error(soar_q0=1e-4, seed=1)
    ens2.step_n(2, model_error=True)
    for k in range(ens.n):
        for a, b in zip(final[k], ens2.download(k)):
            assert np.array_equal(a, b), k
    ens.cleanUp()
    ens2.cleanUp()


# ---------------------------------------------------------------------------
# adjoints at the benchmark sizes
# ---------------------------------------------------------------------------
def _t64(a, dev="cuda"):
    if isinstance(a, torch.Tensor):
        return a.detach().to(dev, torch.float64)
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def _zero_ghosts(a):
    a = a.clone()
    a[:2] = 0
    a[-2:] = 0
    a[:, :2] = 0
    a[:, -2:] = 0
    return a


def _oracle_params(nx, ny, dx, dy, f, bc=(1, 1, 1, 1)):
    return O.Params(nx=nx, ny=ny, dx=dx, dy=dy, bc=bc, f=np.float32(f))


def _hip_trajectory(d, Hi_dev, q0, dt, n, lib, G):
    """(Q^n, Q*) of every RK2 step of the HIP forward (float32, bit-identical
    to the oracle's, which the parity tests check), on the device."""
    a = [x.clone() for x in q0]
    b = [x.clone() for x in q0]
    traj = []
    for _ in range(n):
        qn = [x.clone() for x in a]
        G.check(lib.gocean_rk_steps(d, 1, float(dt), None, *(G.ptr(x) for x in a), *(G.ptr(x) for x in b),
                                    None, None, None, G.ptr(Hi_dev), G.Fields(), G.Wind(), -1, None), "rk")
        traj.append((qn, [x.clone() for x in b]))
    return traj, a


class _Pins:
    """Branch decisions of every substep of a float32 HIP trajectory, recorded
    by the float32 torch restatement at the trajectory's own points.  Each
    recorded substep's output is checked against the HIP trajectory, so the
    decisions are those of the HIP forward (and of its adjoint, which replays
    the same float32 forward)."""

    def __init__(self, p, Hi_host, traj, end, dt):
        Hi32 = torch.from_numpy(np.ascontiguousarray(Hi_host, dtype=np.float32)).cuda()
        Hm32 = torch.from_numpy(O.init_bm(np.asarray(Hi_host, np.float32))).cuda()
        self.steps = []
        bad = 0
        for i, (qn, qs) in enumerate(traj):
            nxt = traj[i + 1][0] if i + 1 < len(traj) else end
            pins0, pins1 = T.BranchPins(), T.BranchPins()
            with torch.no_grad():
                with T.pinned(pins0):
                    o0 = T.substep(dt, 0, *qn, *qn, Hi32, Hm32, p)
                with T.pinned(pins1):
                    o1 = T.substep(dt, 1, *qs, *qn, Hi32, Hm32, p)
            bad += sum(_n_unequal(a, b) for a, b in zip(o0, qs))
            bad += sum(_n_unequal(a, b) for a, b in zip(o1, nxt))
            pins0.replay = pins1.replay = True
            self.steps.append((pins0, pins1))
        self.unequal = bad

    def ctx(self, i, rk):
        return T.pinned(self.steps[i][rk])


def _n_unequal(a, b):
    """Interior values that differ beyond the flush of subnormals."""
    a = a[2:-2, 2:-2]
    b = b[2:-2, 2:-2]
    tiny = float(H.TINY32)
    fa = torch.where(a.abs() < tiny, torch.zeros_like(a), a)
    fb = torch.where(b.abs() < tiny, torch.zeros_like(b), b)
    return int((fa != fb).sum())


def _nopin(i, rk):
    return contextlib.nullcontext()


def _chain_grad64(p, Hi64, Hm64, traj, lam_out, dt, pins=None, dtype=torch.float64):
    """float64 reverse pass through each RK2 step, linearised at the float32
    trajectory's points (torch autograd of the float64 restatement); with
    ``pins``, along the branches the float32 forward took.  With dtype
    float32 (and Hi64/Hm64 in float32) the same pass in float32: an
    independent float32 evaluation of the same linearisation."""
    ctx = pins.ctx if pins is not None else _nopin
    cv = lambda a: a.detach().to("cuda", dtype)
    lam = [_zero_ghosts(cv(a)) for a in lam_out]
    for i in reversed(range(len(traj))):
        qn, qs = traj[i]
        qn_t = [cv(a).requires_grad_(True) for a in qn]
        qs_t = [cv(a).requires_grad_(True) for a in qs]
        with ctx(i, 1):
            out = T.substep(dt, 1, *qs_t, *qn_t, Hi64, Hm64, p)
        L = sum((l[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for l, o in zip(lam, out))
        g = torch.autograd.grad(L, qs_t + qn_t, allow_unused=True)
        g_s = [_zero_ghosts(x) for x in g[:3]]
        g_n = [x if x is not None else torch.zeros_like(lam[0]) for x in g[3:]]
        qn2 = [cv(a).requires_grad_(True) for a in qn]
        with ctx(i, 0):
            out2 = T.substep(dt, 0, *qn2, *qn2, Hi64, Hm64, p)
        L2 = sum((a[2:-2, 2:-2] * o[2:-2, 2:-2]).sum() for a, o in zip(g_s, out2))
        g2 = torch.autograd.grad(L2, qn2)
        lam = [_zero_ghosts((a + b).detach()) for a, b in zip(g2, g_n)]
    return lam


def _fd_tangent64(p, Hi64, Hm64, traj, d0, dt, eps=1e-8, pins=None):
    """float64 central-difference tangent-linear run through the substeps of
    the float32 trajectory, each linearised at the trajectory's own points
    (the small-domain adjoint tests' construction, on the device); with
    ``pins``, every evaluation replays the float32 forward's branches."""
    ctx = pins.ctx if pins is not None else _nopin
    dq = [_zero_ghosts(_t64(a)) for a in d0]
    with torch.no_grad():
        for i, (qn, qs) in enumerate(traj):
            qn64 = [_t64(a) for a in qn]
            qs64 = [_t64(a) for a in qs]
            sc = max(float(a.abs().max()) for a in dq)
            u = [a / sc for a in dq]
            with ctx(i, 0):
                pl = T.substep(dt, 0, *[a + eps * b for a, b in zip(qn64, u)],
                               *[a + eps * b for a, b in zip(qn64, u)], Hi64, Hm64, p)
            with ctx(i, 0):
                mi = T.substep(dt, 0, *[a - eps * b for a, b in zip(qn64, u)],
                               *[a - eps * b for a, b in zip(qn64, u)], Hi64, Hm64, p)
            us = [_zero_ghosts((a - b) / (2 * eps)) for a, b in zip(pl, mi)]
            ss = max(float(a.abs().max()) for a in us)
            with ctx(i, 1):
                pl = T.substep(dt, 1, *[a + eps * b / ss for a, b in zip(qs64, us)],
                               *[a + eps * b / ss for a, b in zip(qn64, u)], Hi64, Hm64, p)
            with ctx(i, 1):
                mi = T.substep(dt, 1, *[a - eps * b / ss for a, b in zip(qs64, us)],
                               *[a - eps * b / ss for a, b in zip(qn64, u)], Hi64, Hm64, p)
            dq = [_zero_ghosts((a - b) / (2 * eps) * ss * sc) for a, b in zip(pl, mi)]
    return dq


# Branch ties.  The scheme is piecewise smooth: minmod picks a candidate,
# the tangential flux upwinds on the sign of u- + u+, the wave speeds are
# extrema, depths are clamped at zero.  A float64 reverse pass that takes
# those decisions from its own float64 evaluation at the float32
# trajectory's points leaves the HIP adjoint's pieces wherever a decision
# quantity sits within float32 rounding of its switch (u- + u+ at a wall face,
# u changing sign, near-equal minmod candidates).  On 512^2 and 768^2 eddy
# fields that happens within the first steps, and the relative L2 difference
# is 2e-4 to 2e-3.  The float64 references here therefore replay the float32
# forward's decisions (recorded by the float32 restatement, whose outputs are
# checked equal to the HIP trajectory), and the HIP gradient is held to
# PINNED_TOL against them; the unpinned figures are printed beside.  An
# independent float32 evaluation of the same pinned linearisation (torch
# autograd in float32) sits at 5e-7 to 2e-6 from it over 100 steps, and the
# HIP adjoint at or below that.
PINNED_TOL = 1e-5


def _rel(a, b):
    a = a[2:-2, 2:-2]
    b = b[2:-2, 2:-2]
    return float((a - b).norm() / b.norm())


def test_config4_adjoint_gradient():
    """512^2, walls, random depth, fixed dt from the initial CFL (0.7), 100
    RK2 steps in two checkpoint segments of 50, J = 0.5 sum over 64 seeded
    points of (bilinear eta - y)^2 with y from a twin run plus N(0, 0.01 m):

    * the float32 restatement, run at every substep of the HIP trajectory to
      record its branch decisions, reproduces that trajectory exactly;
    * the HIP gradient (records kept, and recomputed per segment: identical)
      against the float64 reverse pass along the same float32 trajectory and
      its branches, relative L2 <= PINNED_TOL for eta, hu and hv (the
      unpinned float64 figures are printed);
    * 16 seeded random directions d: the float64 central-difference
      (eps = 1e-8) tangent-linear run along that trajectory and its branches
      against <float64 reverse gradient, d> to 1e-6 of sum |grad * d| (the
      reverse pass is the transpose of the linearisation) and against
      <HIP gradient, d> to PINNED_TOL of it;
    * for 4 of them the eps = 1e-3 central difference of the float64 run's J
      is printed: it crosses kinks of the scheme (wave-speed maxima, limiter
      switches), so it is reported, not asserted."""
    B, CDKLM16, _, Common, A, G = _mods()
    _say("config 4 adjoint at 512x512")
    rng = np.random.default_rng(44)
    c4 = B.config4()
    n, steps, K, n_obs = c4.nx, 100, 50, 64
    lib = G.load()
    ctx = Common.HIPContext(0)
    sim = CDKLM16.CDKLM16(ctx, c4.eta, c4.hu, c4.hv, c4.Hi, n, n, c4.dx, c4.dy, -1.0, c4.g, c4.f, 0.0,
                          courant_number=0.7)
    sim.gpu_stream.synchronize()
    torch.cuda.synchronize()
    dt = float(np.float32(sim.dt))
    d = sim.domain
    Hi_dev = sim.bathymetry.Bi.tensor
    q0 = [a.tensor.clone() for a in (sim.gpu_data.h0, sim.gpu_data.hu0, sim.gpu_data.hv0)]
    ox = rng.uniform(2.0, n - 2.0, n_obs) * c4.dx
    oy = rng.uniform(2.0, n - 2.0, n_obs) * c4.dy
    # twin run from a perturbed eta, plus observation noise
    pert = torch.from_numpy((0.05 * B.gaussian_random_field((n + 4, n + 4), 16.0, rng)).astype(np.float32))
    twin, twin_end = _hip_trajectory(d, Hi_dev, [q0[0] + pert.cuda(), q0[1], q0[2]], dt, steps, lib, G)
    del twin
    yk = A.bilinear_eta(twin_end[0], ox, oy, c4.dx, c4.dy).detach().cpu().numpy().astype(np.float64)
    yk = yk + rng.normal(0.0, 0.01, n_obs)

    grads = []
    for keep in (True, False):
        adj = A.simulator_adjoint(sim, checkpoint_every=K)
        adj.keep_records = keep
        qd = [x.clone().requires_grad_(True) for x in q0]
        out = A.CDKLM16RK2.apply(*qd, adj, steps, dt)
        # the misfit in float64 (its residuals cancel: in float32 they would
        # carry 1e-6-relative rounding into the cotangent); the adjoint gets
        # its gradient rounded to float32
        A.point_misfit(out[0].double(), ox, oy, yk, c4.dx, c4.dy).backward()
        torch.cuda.synchronize()
        grads.append([x.grad.detach().to(torch.float64) for x in qd])
    for a, b in zip(*grads):
        assert torch.equal(a, b)
    g = grads[0]

    traj, end = _hip_trajectory(d, Hi_dev, q0, dt, steps, lib, G)
    for a, b in zip(end, out):
        assert torch.equal(a, b.detach())          # the adjoint's forward is the plain forward
    p = _oracle_params(n, n, c4.dx, c4.dy, c4.f)
    Hi_host = Hi_dev[:, :n + 5].cpu().numpy()
    Hi64 = _t64(H.hi64(Hi_host))
    Hm64 = _t64(O.init_bm(H.hi64(Hi_host), dtype=np.float64))
    e_end = _t64(end[0]).requires_grad_(True)
    A.point_misfit(e_end, ox, oy, yk, c4.dx, c4.dy).backward()
    lam_n = [e_end.grad.detach(), torch.zeros_like(e_end), torch.zeros_like(e_end)]
    pins = _Pins(p, Hi_host, traj, end, dt)
    _say("   float32 restatement vs HIP trajectory: unequal values", pins.unequal)
    assert pins.unequal == 0
    _say("   float64 reverse pass over", steps, "steps")
    free = _chain_grad64(p, Hi64, Hm64, traj, lam_n, dt)
    ref = _chain_grad64(p, Hi64, Hm64, traj, lam_n, dt, pins=pins)
    r32 = _chain_grad64(p, Hi64.float(), Hm64.float(), traj, lam_n, dt, pins=pins, dtype=torch.float32)
    errs = []
    for name, a, b, c, d in zip(("eta", "hu", "hv"), g, ref, free, r32):
        e = _rel(a, b)
        errs.append(e)
        _say("   vs fp64 reverse pass", name, "rel L2", e, "(own float64 branches:", _rel(a, c),
             "; float32 torch pass:", _rel(d.double(), b), ")")
    for name, e in zip(("eta", "hu", "hv"), errs):
        assert e <= PINNED_TOL, (name, e)

    drng = np.random.default_rng(4404)
    for k in range(16):
        D = [_zero_ghosts(torch.from_numpy(drng.standard_normal((n + 4, n + 4))).cuda()) for _ in range(3)]
        dq = _fd_tangent64(p, Hi64, Hm64, traj, D, dt, pins=pins)
        fd = float((lam_n[0] * dq[0]).sum())
        ad = sum(float((a * b).sum()) for a, b in zip(g, D))
        ad64 = sum(float((a * b).sum()) for a, b in zip(ref, D))
        scale = sum(float((a * b).abs().sum()) for a, b in zip(g, D))
        msg = (f"   direction {k}: tangent-linear {fd:.6e} adjoint {ad:.6e} rel {abs(ad - fd) / scale:.2e}"
               f" fp64 reverse {ad64:.6e} rel {abs(ad64 - fd) / scale:.2e}")
        if k < 4:
            # eps = 1e-3 central difference of the float64 run (own trajectory)
            Js = []
            with torch.no_grad():
                for sgn in (1.0, -1.0):
                    qe = [_t64(a) + sgn * 1e-3 * b for a, b in zip(q0, D)]
                    o = T.rk2_steps(*qe, Hi64, Hm64, p, dt, steps)
                    Js.append(float(A.point_misfit(o[0], ox, oy, yk, c4.dx, c4.dy)))
            msg += f"; eps=1e-3 fp64 run FD {(Js[0] - Js[1]) / 2e-3:.6e}"
        _say(msg)
        assert abs(ad64 - fd) <= 1e-6 * scale, (k, ad64, fd, scale)
        assert abs(ad - fd) <= PINNED_TOL * scale, (k, ad, fd, scale)
    sim.cleanUp()


def test_config5_batched_adjoint():
    """32 domains, sides drawn from {256, 384, 512, 768}, walls, own depth,
    per-domain dt (CFL 0.7), 200 RK2 steps with a checkpoint every 25, a
    16-point misfit per domain: the batched forward + adjoint equals every
    domain's single-domain adjoint bit for bit; two domains (the first 256^2
    and the first 768^2) over 50 steps against the float64 reverse pass along
    the float32 trajectory's branches (relative L2 <= PINNED_TOL; the unpinned
    figures are printed)."""
    B, _, CE, Common, A, G = _mods()
    _say("config 5: 32 uneven domains, forward + adjoint")
    rng = np.random.default_rng(55)
    inps = B.config5()
    ctx = Common.HIPContext(0)
    lib = G.load()
    ens = CE.CDKLM16Ensemble(ctx, inps, courant_number=0.7)
    dts = ens.dt_host()
    state0 = ens.state[0].clone()
    n_obs = 16
    obs = []
    for i in inps:
        obs.append((rng.uniform(2.0, i.nx - 2.0, n_obs) * i.dx, rng.uniform(2.0, i.ny - 2.0, n_obs) * i.dy,
                    0.05 * rng.standard_normal(n_obs)))

    def views(flat, k):
        o, pl = int(ens.state_off[k]), ens.planes[k]
        return [flat[o + f * pl:o + (f + 1) * pl].view(ens.ny[k] + 4, ens.pitch[k]) for f in range(3)]

    def run_batched(steps, K):
        ea = A.EnsembleAdjoint(ens, checkpoint_every=K)
        out = ea.forward(state0, steps).requires_grad_(True)
        J = sum(A.point_misfit(views(out, k)[0].double(), *obs[k][:2], obs[k][2], i.dx, i.dy)
                for k, i in enumerate(inps))
        lam = torch.autograd.grad(J, out)[0]
        grad = ea.backward(lam)
        torch.cuda.synchronize()
        return out.detach(), grad

    def domain(k):
        i = inps[k]
        d = Common.make_domain(i.nx, i.ny, i.dx, i.dy, i.g, 0.0, 1.3, 2, i.bc, pitch=ens.pitch[k],
                               hi_pitch=i.nx + 5, f=i.f)
        Hi_k = ens.Hi[int(ens.hi_off[k]):int(ens.hi_off[k + 1])].view(i.ny + 5, i.nx + 5)
        return d, Hi_k

    out, grad = run_batched(200, 25)
    for k, i in enumerate(inps):
        d, Hi_k = domain(k)
        adj = A.RK2Adjoint(d, Hi_k, checkpoint_every=25)
        qs = [t.clone().requires_grad_(True) for t in views(state0, k)]
        o = A.CDKLM16RK2.apply(*qs, adj, 200, float(dts[k]))
        A.point_misfit(o[0].double(), *obs[k][:2], obs[k][2], i.dx, i.dy).backward()
        torch.cuda.synchronize()
        for a, b in zip(views(out, k), o):
            assert torch.equal(a, b.detach()), ("forward", k)
        for a, b in zip(views(grad, k), qs):
            assert torch.equal(a, b.grad), ("gradient", k, i.nx)
    _say("   32 domains: batched == single-domain (forward and gradient)")

    steps = 50
    out, grad = run_batched(steps, 25)
    first = [next(k for k, i in enumerate(inps) if i.nx == side) for side in (256, 768)]
    for k in first:
        i = inps[k]
        d, Hi_k = domain(k)
        traj, end = _hip_trajectory(d, Hi_k, [t.clone() for t in views(state0, k)], float(dts[k]), steps, lib, G)
        for a, b in zip(end, views(out, k)):
            assert torch.equal(a, b)
        p = _oracle_params(i.nx, i.ny, i.dx, i.dy, i.f)
        Hi_host = Hi_k.cpu().numpy()
        Hi64 = _t64(H.hi64(Hi_host))
        Hm64 = _t64(O.init_bm(H.hi64(Hi_host), dtype=np.float64))
        e_end = _t64(end[0]).requires_grad_(True)
        A.point_misfit(e_end, *obs[k][:2], obs[k][2], i.dx, i.dy).backward()
        lam_n = [e_end.grad.detach(), torch.zeros_like(e_end), torch.zeros_like(e_end)]
        dt_k = float(np.float32(dts[k]))
        pins = _Pins(p, Hi_host, traj, end, dt_k)
        _say("   domain", k, "float32 restatement vs HIP trajectory: unequal values", pins.unequal)
        assert pins.unequal == 0
        free = _chain_grad64(p, Hi64, Hm64, traj, lam_n, dt_k)
        ref = _chain_grad64(p, Hi64, Hm64, traj, lam_n, dt_k, pins=pins)
        r32 = _chain_grad64(p, Hi64.float(), Hm64.float(), traj, lam_n, dt_k, pins=pins, dtype=torch.float32)
        del pins
        errs = []
        for name, a, b, c, d in zip(("eta", "hu", "hv"), views(grad, k), ref, free, r32):
            e = _rel(a.to(torch.float64), b)
            errs.append(e)
            _say("   domain", k, f"{i.nx}x{i.ny}", name, "vs fp64 reverse pass, rel L2", e,
                 "(own float64 branches:", _rel(a.to(torch.float64), c), "; float32 torch pass:",
                 _rel(d.double(), b), ")")
        for name, e in zip(("eta", "hu", "hv"), errs):
            assert e <= PINNED_TOL, (k, name, e)
    ens.cleanUp()
