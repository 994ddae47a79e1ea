"""Host-side pieces of the reverse-mode module that need no GPU: the flat
bilinear observation stencil against the per-plane bilinear interpolation,
and the ghost-cell index maps against the boundary rules they encode."""

import numpy as np
import torch

from SWESimulators import adjoint as A


def test_stencil_matches_bilinear_eta():
    rng = np.random.default_rng(0)
    ny, nx, pitch, off = 20, 30, 36, 11
    e = torch.randn(ny + 4, pitch, dtype=torch.float64)
    ox = rng.uniform(1.0, nx - 1.0, 25) * 200.0
    oy = rng.uniform(1.0, ny - 1.0, 25) * 150.0
    ref = A.bilinear_eta(e, ox, oy, 200.0, 150.0)
    idx, w = A.bilinear_stencil(off, pitch, ox, oy, 200.0, 150.0)
    flat = torch.cat([torch.zeros(off, dtype=torch.float64), e.reshape(-1)])
    got = (flat[torch.from_numpy(idx)] * torch.from_numpy(w).double()).sum(1)
    # weights are rounded to float32
    assert torch.allclose(got, ref, atol=1e-6 * float(ref.abs().max()))
    y = torch.from_numpy(rng.standard_normal(25))
    J = A.stencil_misfit(flat, torch.from_numpy(idx), torch.from_numpy(w).double(), y)
    assert abs(float(J) - 0.5 * float(((got - y) ** 2).sum())) < 1e-12


def test_ghost_index_rules():
    nx, ny, pitch = 7, 5, 13
    for bc_n, bc_e in [(1, 1), (2, 2), (2, 1), (1, 2)]:
        src, dst, keep = A._ghost_index(nx, ny, pitch, bc_n, bc_e)
        R, C = np.divmod(src, pitch)
        # every rewritten ghost is a periodic image of the interior cell it copies
        Rd, Cd = np.divmod(dst, pitch)
        assert np.all((Rd >= 2) & (Rd <= ny + 1) & (Cd >= 2) & (Cd <= nx + 1))
        assert np.all((R == Rd) | (bc_n == 2)) and np.all((C == Cd) | (bc_e == 2))
        assert np.all((R - Rd) % ny == 0) and np.all((C - Cd) % nx == 0)
        # rewritten + untouched = every ghost cell of the padded plane
        ghost = sum(1 for r in range(ny + 4) for c in range(nx + 4)
                    if r < 2 or r > ny + 1 or c < 2 or c > nx + 1)
        assert len(src) + len(keep) == ghost
        assert len(set(src.tolist()) & set(keep.tolist())) == 0


def test_wrap_groups_have_distinct_destinations():
    """The periodic fold runs as index_add_ groups whose destinations are
    distinct (so the device's atomic adds cannot reorder a sum); together
    the groups hold every (ghost, image) pair once, and on a doubly periodic
    domain a cell next to a corner takes its three images in three groups."""
    nx, ny, pitch = 7, 5, 13
    for bc_n, bc_e in [(1, 1), (2, 2), (2, 1), (1, 2)]:
        src, dst, keep = A._ghost_index(nx, ny, pitch, bc_n, bc_e)
        groups = A._unique_dst_groups(src, dst)
        pairs = []
        for a, b in groups:
            assert len(set(b.tolist())) == len(b)
            pairs += list(zip(a.tolist(), b.tolist()))
        assert sorted(pairs) == sorted(zip(src.tolist(), dst.tolist()))
        if (bc_n, bc_e) == (2, 2):
            assert len(groups) == 3
            assert sum(1 for a, b in groups for d in b.tolist() if d == 2 * pitch + 2) == 3
        elif (bc_n, bc_e) == (1, 1):
            assert groups == []
        else:
            assert len(groups) == 1
