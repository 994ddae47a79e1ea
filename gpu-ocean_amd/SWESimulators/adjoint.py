"""Reverse mode of the fixed-dt CDKLM16 RK2 step: checkpointed adjoint driver
and a ``torch.autograd.Function``.

The reference has no gradients; callers that do variational data
assimilation or parameter estimation around gpu-ocean's step
(CDKLM16.step, gpu_ocean/SWESimulators/CDKLM16.py:508-593 with
rk_order 2, CDKLM16.py:533-544) need d(misfit)/d(initial state).  The
Jacobian-transpose products are hand-written HIP kernels
(``gocean_adjoint_rk2_step`` in ``lib/libgocean_hip.so``); this module only
schedules them:

* forward: n steps of the HIP forward, a copy of the state kept every
  ``checkpoint_every`` steps (revolve-style uniform checkpoints);
* backward: for each segment, last first, one ``gocean_adjoint_segment``
  call re-runs the forward from its checkpoint in the recording variant
  (keeping Q^n, Q* and the branch record of every substep) and applies the
  substep adjoints in reverse.

Memory: (n / K + 1) checkpoints of 3 planes plus (16 K + 9) planes of segment
work, plane = (ny + 4) * pitch floats.  Work: one extra forward run plus the
adjoint kernels.

With ``keep_records`` the forward itself runs the recording variant
(``gocean_adjoint_record``, one work buffer per segment) and the backward
only applies the substep adjoints (``gocean_adjoint_reverse``): no second
forward run, for about 16 planes of memory per step.  ``"auto"`` (default)
keeps the records when they take at most ``record_budget`` of the free
device memory (free, plus what torch's allocator holds unused) at the time
of the forward run.

Scope (enforced by the library): north = (0, 1), constant f, r == 0,
rk_order 2, fixed dt, constant wind.  Ghost cells: an output ghost cell the
forward rewrites (periodic images) passes its adjoint to the interior cell it
copies; one it leaves untouched (wall sides) is the input's ghost value, so
its adjoint passes to the same input ghost cell.
"""

from __future__ import annotations

import numpy as np

from . import Common
from . import _gocean as G


def _torch():
    return Common._torch()


def _available_bytes(dev):
    """Device memory a new allocation can use: free in the driver plus what
    torch's allocator holds in reserve (cached blocks not in use)."""
    torch = _torch()
    free, _ = torch.cuda.mem_get_info(dev)
    return free + torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)


class RK2Adjoint(object):
    """Checkpointed forward / reverse runs of one domain.

    ``domain``: a ``gocean_domain_t`` (``Common.make_domain``), rk_order 2.
    ``Hi``: device tensor of corner depths, halo completed
    (``gocean_bathymetry_init``).  ``wind``: ``gocean_wind_t`` or None.
    """

    def __init__(self, domain, Hi, stream=None, wind=None, checkpoint_every=50, tile_config=-1,
                 keep_records="auto", record_budget=0.25, record_tile_config=-1):
        torch = _torch()
        self.lib = G.load()
        self.d = domain
        if domain.rk_order != 2:
            raise ValueError("the adjoint supports rk_order 2 only")
        self.Hi = Hi
        self.dev = Hi.device
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.dev)
        self.wind = wind if wind is not None else G.Wind()
        self.fields = G.Fields()
        self.K = max(1, int(checkpoint_every))
        self.tile_config = int(tile_config)
        self.record_tile_config = int(record_tile_config)
        if keep_records not in (True, False, "auto"):
            raise ValueError("keep_records must be True, False or 'auto'")
        self.keep_records = keep_records
        self.record_budget = float(record_budget)
        self.records = None
        self.shape = (domain.ny + 4, domain.pitch)
        n_scratch = int(self.lib.gocean_adjoint_scratch_floats(domain))
        with torch.cuda.stream(self.stream):
            self.scratch = torch.empty(n_scratch, dtype=torch.float32, device=self.dev)
        self._ghost_maps()
        self.checkpoints = None
        self.work = None

    # ------------------------------------------------------------------
    def _ghost_maps(self):
        """Index arrays for the output-ghost rules (see module docstring)."""
        torch = _torch()
        d = self.d
        nx, ny, pitch = d.nx, d.ny, d.pitch
        pNS = d.bc_north == G_BC_PERIODIC
        pEW = d.bc_east == G_BC_PERIODIC
        R = np.arange(ny + 4, dtype=np.int64)[:, None]
        C = np.arange(nx + 4, dtype=np.int64)[None, :]
        gr = (R < 2) | (R > ny + 1)
        gc = (C < 2) | (C > nx + 1)
        ghost = gr | gc
        written = ghost & (~gr | pNS) & (~gc | pEW)
        Ri = np.where(R < 2, R + ny, np.where(R > ny + 1, R - ny, R))
        Ci = np.where(C < 2, C + nx, np.where(C > nx + 1, C - nx, C))
        o = R * pitch + C
        mk = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=self.dev)
        self.wrap = [(mk(a), mk(b)) for a, b in _unique_dst_groups(o[written], (Ri * pitch + Ci)[written])]
        self.keep = mk(o[ghost & ~written])

    def _enter(self):
        """Order this driver's stream after the caller's current stream."""
        torch = _torch()
        cur = torch.cuda.current_stream(self.dev)
        if cur != self.stream:
            self.stream.wait_stream(cur)
        return cur

    def _leave(self, cur, outs):
        """Order the caller's stream after this driver's work; mark returned
        tensors as used there so their memory is not reused early."""
        if cur != self.stream:
            cur.wait_stream(self.stream)
            for t in outs:
                t.record_stream(cur)

    def _alloc3(self):
        torch = _torch()
        return tuple(torch.empty(self.shape, dtype=torch.float32, device=self.dev) for _ in range(3))

    def _rk_steps(self, q, tmp, n, dt):
        G.check(self.lib.gocean_rk_steps(
            self.d, int(n), float(dt), None, *(G.ptr(a) for a in q), *(G.ptr(a) for a in tmp),
            None, None, None, G.ptr(self.Hi), self.fields, self.wind, self.tile_config,
            G.stream_handle(self.stream)), "gocean_rk_steps")

    def _segments(self):
        n, K = self.n_steps, self.K
        return [min(K, n - s) for s in range(0, n, K)]

    def records_bytes(self, n_steps):
        """Device bytes the kept records of an ``n_steps`` run take."""
        n, K = int(n_steps), self.K
        return sum(4 * int(self.lib.gocean_adjoint_segment_work_floats(self.d, min(K, n - s)))
                   for s in range(0, n, K))

    def _use_records(self):
        if self.keep_records != "auto":
            return bool(self.keep_records)
        return self.records_bytes(self.n_steps) <= self.record_budget * _available_bytes(self.dev)

    def _forward_records(self, q):
        """Recording forward from the (3, ny + 4, pitch) state ``q``, one work
        buffer per segment; returns the final state as three new tensors.
        The last segment's final slot starts as a copy of
        ``q``, so the cells a step does not write (wall ghosts, pitch
        padding) keep the input's values, as in the in-place forward."""
        torch = _torch()
        pl = self.shape[0] * self.shape[1]
        self.records = []
        start = tuple(q.unbind(0))
        segs = self._segments()
        for n, k in enumerate(segs):
            nw = int(self.lib.gocean_adjoint_segment_work_floats(self.d, k))
            w = torch.empty(nw, dtype=torch.float32, device=self.dev)
            o = 3 * (k - 1) * pl
            last = w[o:o + 3 * pl].view(3, *self.shape)
            if n == len(segs) - 1:
                last.copy_(q)
            G.check(self.lib.gocean_adjoint_set_record_tile_config(self.record_tile_config),
                    "gocean_adjoint_set_record_tile_config")
            G.check(self.lib.gocean_adjoint_record(
                self.d, self.wind, k, self.dt, None, *(G.ptr(a) for a in start), G.ptr(self.Hi),
                G.ptr(w), nw, G.stream_handle(self.stream)), "gocean_adjoint_record")
            self.records.append((start, w, k))
            start = tuple(last.unbind(0))
        return tuple(a.clone() for a in start)

    # ------------------------------------------------------------------
    def forward(self, q0, n_steps, dt):
        """Run ``n_steps`` fixed-dt RK2 steps from padded (eta, hu, hv) device
        tensors; returns new tensors and keeps the checkpoints (or the
        records, see ``keep_records``)."""
        torch = _torch()
        self.n_steps, self.dt = int(n_steps), float(np.float32(dt))
        self.records = None
        self.checkpoints = None
        cur = self._enter()
        with torch.cuda.stream(self.stream):
            if self._use_records():
                q3 = torch.stack([a.detach().to(torch.float32) for a in q0])
                q = self._forward_records(q3)
                self.checkpoints = [r[0] for r in self.records]
                self.final = q
                self._leave(cur, q)
                return q
            q = tuple(a.detach().to(torch.float32).contiguous().clone() for a in q0)
            tmp = self._alloc3()
            self.checkpoints = []
            done = 0
            while done < self.n_steps:
                self.checkpoints.append(tuple(a.clone() for a in q))
                k = min(self.K, self.n_steps - done)
                self._rk_steps(q, tmp, k, self.dt)
                done += k
            self.final = q
        self._leave(cur, q)
        return q

    def backward(self, lam):
        """Adjoint of the last ``forward`` run: ``lam`` = d(loss)/d(output)
        (padded tensors); returns d(loss)/d(input)."""
        torch = _torch()
        if self.checkpoints is None:
            raise RuntimeError("backward() needs a preceding forward()")
        cur_stream = self._enter()
        with torch.cuda.stream(self.stream):
            # the three fields as one (3, ny + 4, pitch) tensor, so that each
            # ghost-cell rule is one indexing launch for all fields
            L = torch.stack([a.detach().to(torch.float32) for a in lam])
            Lf = L.view(3, -1)
            keep = Lf[:, self.keep]
            for src, dst in self.wrap:
                Lf.index_add_(1, dst, Lf[:, src])
            cur = tuple(L.unbind(0))
            nxt = tuple(torch.empty_like(L).unbind(0))
            n_seg = len(self.checkpoints)
            if self.records is not None:
                for start, w, k in reversed(self.records):
                    G.check(self.lib.gocean_adjoint_reverse(
                        self.d, k, self.dt, None, G.ptr(start[0]), *(G.ptr(a) for a in cur),
                        *(G.ptr(a) for a in nxt), G.ptr(self.Hi), G.ptr(w), w.numel(),
                        G.stream_handle(self.stream)), "gocean_adjoint_reverse")
                    cur, nxt = nxt, cur
                n_seg = 0
            if self.work is None and n_seg:
                nw = int(self.lib.gocean_adjoint_segment_work_floats(self.d, self.K))
                self.work = torch.empty(nw, dtype=torch.float32, device=self.dev)
            if n_seg:
                G.check(self.lib.gocean_adjoint_set_record_tile_config(self.record_tile_config),
                        "gocean_adjoint_set_record_tile_config")
            for s in range(n_seg - 1, -1, -1):
                k = min(self.K, self.n_steps - s * self.K)
                G.check(self.lib.gocean_adjoint_segment(
                    self.d, self.wind, k, self.dt, None, *(G.ptr(a) for a in self.checkpoints[s]),
                    *(G.ptr(a) for a in cur), *(G.ptr(a) for a in nxt), G.ptr(self.Hi),
                    G.ptr(self.work), self.work.numel(), G.stream_handle(self.stream)),
                    "gocean_adjoint_segment")
                cur, nxt = nxt, cur
            O = torch.stack(cur)
            O.view(3, -1)[:, self.keep] += keep
            out = tuple(O.unbind(0))
        self._leave(cur_stream, out)
        return out


G_BC_PERIODIC = 2


class CDKLM16RK2(object):
    """``torch.autograd.Function`` wrapper: ``CDKLM16RK2.apply(eta, hu, hv,
    adjoint, n_steps, dt)`` returns the state after ``n_steps`` RK2 steps and
    backpropagates through the hand-written HIP adjoint."""

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, eta, hu, hv, adj, n_steps, dt):
                    ctx.adj = adj
                    return adj.forward((eta, hu, hv), n_steps, dt)

                @staticmethod
                def backward(ctx, g_eta, g_hu, g_hv):
                    z = lambda g, like: g if g is not None else torch.zeros_like(like)
                    fin = ctx.adj.final
                    lam = ctx.adj.backward((z(g_eta, fin[0]), z(g_hu, fin[1]), z(g_hv, fin[2])))
                    return lam[0], lam[1], lam[2], None, None, None

            cls._fn = _Fn
        return cls._fn

    @classmethod
    def apply(cls, eta, hu, hv, adj, n_steps, dt):
        return cls.function().apply(eta, hu, hv, adj, n_steps, dt)


def simulator_adjoint(sim, checkpoint_every=50):
    """An ``RK2Adjoint`` on a CDKLM16 simulator's domain, bathymetry, stream
    and current wind."""
    return RK2Adjoint(sim.domain, sim.bathymetry.Bi.tensor, stream=sim.gpu_stream,
                      wind=sim._wind(), checkpoint_every=checkpoint_every,
                      tile_config=sim.tile_config)


def _unique_dst_groups(src, dst):
    """Split (src, dst) index pairs, in their order, into groups with distinct
    destinations.  On a doubly periodic domain an interior cell next to a
    corner is the image of three ghost cells; one index_add_ over all pairs
    would sum them in whatever order the device's atomics land, so the
    folded adjoint would change in its last bits from run to run.  Applied
    one group after another, every destination takes its adds in a fixed
    order."""
    groups = []
    for a, b in zip(np.asarray(src).tolist(), np.asarray(dst).tolist()):
        for g in groups:
            if b not in g[2]:
                g[0].append(a); g[1].append(b); g[2].add(b)
                break
        else:
            groups.append(([a], [b], {b}))
    return [(np.asarray(g[0], np.int64), np.asarray(g[1], np.int64)) for g in groups]


def _ghost_index(nx, ny, pitch, bc_north, bc_east):
    """Flat indices within one padded plane: output ghost cells the forward
    rewrites (periodic images) with the interior cell each copies, and the
    ghost cells it leaves untouched (see the module docstring)."""
    pNS = bc_north == G_BC_PERIODIC
    pEW = bc_east == G_BC_PERIODIC
    R = np.arange(ny + 4, dtype=np.int64)[:, None]
    C = np.arange(nx + 4, dtype=np.int64)[None, :]
    gr = (R < 2) | (R > ny + 1)
    gc = (C < 2) | (C > nx + 1)
    ghost = gr | gc
    written = ghost & (~gr | pNS) & (~gc | pEW)
    Ri = np.where(R < 2, R + ny, np.where(R > ny + 1, R - ny, R))
    Ci = np.where(C < 2, C + nx, np.where(C > nx + 1, C - nx, C))
    o = R * pitch + C
    return o[written], (Ri * pitch + Ci)[written], o[ghost & ~written]


class EnsembleAdjoint(object):
    """Reverse mode of a ``CDKLM16Ensemble`` (rk_order 2, constant f, no
    rotation, r == 0): the forward keeps a copy of the flat ensemble state
    every ``checkpoint_every`` steps; the backward issues one
    ``gocean_ensemble_adjoint_segment`` per checkpoint segment, each a few
    launches that cover every member.  Adjoints are flat tensors laid out
    like the ensemble state, so member k's gradient is ``ens.views``-shaped
    at ``ens.state_off[k]``.

    ``keep_records`` as in ``RK2Adjoint``: the forward runs the recording
    variant (``gocean_ensemble_adjoint_record``, one work buffer per segment)
    and the backward only the reverse passes
    (``gocean_ensemble_adjoint_reverse``)."""

    def __init__(self, ens, checkpoint_every=25, keep_records="auto", record_budget=0.25):
        torch = _torch()
        self.ens = ens
        self.lib = G.load()
        self.K = max(1, int(checkpoint_every))
        nw = int(self.lib.gocean_ensemble_adjoint_setup(ens.handle, self.K))
        if nw < 0:
            G.check(int(nw), "gocean_ensemble_adjoint_setup")
        self.stream = ens.gpu_stream
        dev = ens.state[0].device
        with torch.cuda.stream(self.stream):
            self.work = torch.empty(nw, dtype=torch.float32, device=dev)
        src, dst, keep = [], [], []
        for k in range(ens.n):
            a, b, c = _ghost_index(ens.nx[k], ens.ny[k], ens.pitch[k], ens.bcs[k][0], ens.bcs[k][1])
            for f in range(3):
                off = int(ens.state_off[k]) + f * ens.planes[k]
                src.append(a + off); dst.append(b + off); keep.append(c + off)
        mk = lambda parts: torch.as_tensor(np.concatenate(parts), device=dev)
        self.wrap = [(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev))
                     for a, b in _unique_dst_groups(np.concatenate(src), np.concatenate(dst))]
        self.keep = mk(keep)
        if keep_records not in (True, False, "auto"):
            raise ValueError("keep_records must be True, False or 'auto'")
        self.keep_records = keep_records
        self.record_budget = float(record_budget)
        self.checkpoints = None
        self.records = None

    def records_bytes(self, n_steps):
        """Device bytes the kept records of an ``n_steps`` run take."""
        n_seg = -(-int(n_steps) // self.K)
        return n_seg * 4 * (self.work.numel() + self.ens.state[0].numel())

    def _use_records(self):
        if self.keep_records != "auto":
            return bool(self.keep_records)
        return self.records_bytes(self.n_steps) <= self.record_budget * _available_bytes(self.work.device)

    def forward(self, state0, n_steps):
        """Run n_steps RK2 steps of every member (per-member device dt) from
        the flat state ``state0``; returns the final flat state."""
        torch = _torch()
        ens = self.ens
        self.n_steps = int(n_steps)
        self.records = None
        cur = torch.cuda.current_stream(state0.device)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            ens.state[0].copy_(state0.detach())
            if self._use_records():
                self.records = []
                start = ens.state[0].clone()
                self.checkpoints = [start]
                done = 0
                while done < self.n_steps:
                    k = min(self.K, self.n_steps - done)
                    w = torch.empty_like(self.work)
                    nxt = torch.empty_like(start)
                    G.check(self.lib.gocean_ensemble_adjoint_record(
                        ens.handle, k, G.ptr(ens.dt), G.ptr(start), G.ptr(ens.Hi), G.ptr(w), w.numel(),
                        G.ptr(nxt), G.stream_handle(self.stream)), "gocean_ensemble_adjoint_record")
                    self.records.append((start, w, k))
                    start = nxt
                    done += k
                ens.state[0].copy_(start)
                out = start
                cur.wait_stream(self.stream)
                out.record_stream(cur)
                return out
            self.checkpoints = []
            done = 0
            while done < self.n_steps:
                self.checkpoints.append(ens.state[0].clone())
                k = min(self.K, self.n_steps - done)
                ens.step_n(k)
                done += k
            out = ens.state[0].clone()
        cur.wait_stream(self.stream)
        out.record_stream(cur)
        return out

    def backward(self, lam):
        """Flat d(loss)/d(final state) -> flat d(loss)/d(state0)."""
        torch = _torch()
        if self.checkpoints is None:
            raise RuntimeError("backward() needs a preceding forward()")
        ens = self.ens
        cur_stream = torch.cuda.current_stream(lam.device)
        self.stream.wait_stream(cur_stream)
        with torch.cuda.stream(self.stream):
            a = lam.detach().to(torch.float32).contiguous().clone()
            kept = a[self.keep].clone()
            for src, dst in self.wrap:
                a.index_add_(0, dst, a[src])
            b = torch.empty_like(a)
            n_seg = len(self.checkpoints)
            if self.records is not None:
                for start, w, k in reversed(self.records):
                    G.check(self.lib.gocean_ensemble_adjoint_reverse(
                        ens.handle, k, G.ptr(ens.dt), G.ptr(start), G.ptr(a), G.ptr(b), G.ptr(ens.Hi),
                        G.ptr(w), w.numel(), G.stream_handle(self.stream)), "gocean_ensemble_adjoint_reverse")
                    a, b = b, a
                n_seg = 0
            for s in range(n_seg - 1, -1, -1):
                k = min(self.K, self.n_steps - s * self.K)
                G.check(self.lib.gocean_ensemble_adjoint_segment(
                    ens.handle, k, G.ptr(ens.dt), G.ptr(self.checkpoints[s]), G.ptr(a), G.ptr(b),
                    G.ptr(ens.Hi), G.ptr(self.work), self.work.numel(), G.stream_handle(self.stream)),
                    "gocean_ensemble_adjoint_segment")
                a, b = b, a
            a[self.keep] += kept
        cur_stream.wait_stream(self.stream)
        a.record_stream(cur_stream)
        return a


def bilinear_eta(eta, obs_x, obs_y, dx, dy):
    """eta bilinearly interpolated at physical points (x, y), with cell (i, j)
    of the interior centred at ((i + 0.5) dx, (j + 0.5) dy); ``eta`` is a
    padded (ny + 4, nx + 4 or pitch) tensor.  Differentiable torch ops; points
    must lie between interior cell centres."""
    torch = _torch()
    fx = torch.as_tensor(obs_x, dtype=eta.dtype, device=eta.device) / dx - 0.5
    fy = torch.as_tensor(obs_y, dtype=eta.dtype, device=eta.device) / dy - 0.5
    i0 = torch.floor(fx).to(torch.int64)
    j0 = torch.floor(fy).to(torch.int64)
    wx = fx - i0.to(eta.dtype)
    wy = fy - j0.to(eta.dtype)
    r, c = j0 + 2, i0 + 2
    e00, e01 = eta[r, c], eta[r, c + 1]
    e10, e11 = eta[r + 1, c], eta[r + 1, c + 1]
    return (1 - wy) * ((1 - wx) * e00 + wx * e01) + wy * ((1 - wx) * e10 + wx * e11)


def bilinear_stencil(offset, pitch, obs_x, obs_y, dx, dy):
    """Flat indices and weights, both (n, 4), of the bilinear interpolation
    of ``bilinear_eta`` for points (x, y) in the padded plane that starts at
    element ``offset`` of a flat buffer with row length ``pitch``: eta at the
    points is ``(flat[index] * weight).sum(1)``.  With the stencils of many
    domains concatenated, one gather covers every member of an ensemble."""
    fx = np.asarray(obs_x, dtype=np.float64) / dx - 0.5
    fy = np.asarray(obs_y, dtype=np.float64) / dy - 0.5
    i0 = np.floor(fx).astype(np.int64)
    j0 = np.floor(fy).astype(np.int64)
    wx = (fx - i0).astype(np.float32)
    wy = (fy - j0).astype(np.float32)
    base = int(offset) + (j0 + 2) * int(pitch) + (i0 + 2)
    index = np.stack([base, base + 1, base + pitch, base + pitch + 1], axis=1)
    weight = np.stack([(1 - wy) * (1 - wx), (1 - wy) * wx, wy * (1 - wx), wy * wx], axis=1)
    return index, weight.astype(np.float32)


def stencil_misfit(flat, index, weight, y):
    """J = 0.5 * sum_k ((flat[index_k] . weight_k) - y_k)^2 for device
    tensors from ``bilinear_stencil``: a few launches for any number of
    points and domains."""
    d = (flat[index] * weight).sum(1) - y
    return 0.5 * (d * d).sum()


def point_misfit(eta, obs_x, obs_y, y, dx, dy):
    """J = 0.5 * sum_k (eta(x_k, y_k) - y_k)^2 with bilinear interpolation."""
    torch = _torch()
    d = bilinear_eta(eta, obs_x, obs_y, dx, dy) - torch.as_tensor(y, dtype=eta.dtype, device=eta.device)
    return 0.5 * (d * d).sum()
