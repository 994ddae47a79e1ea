"""Host side of the device LETKF (SWESimulators/ETKFOcean.py): the weight
stencil and the groups equal the oracle's restatement of the reference
(ETKFOcean.py:396-418, 478-514), and the tile lists handed to the apply
kernel name exactly the buoys whose nonzero weights reach each cell."""

import numpy as np
import pytest

from oracle import etkf_np as E
from SWESimulators import ETKFOcean as P


@pytest.mark.parametrize("r,nx,ny,dx", [(1.0, 200, 200, 1.0), (2.0, 40, 36, 400.0), (3.3, 60, 50, 1.0),
                                         (15.0, 500, 300, 2000.0), (4.0, 13, 13, 1.0)])
def test_weight_stencil_equals_oracle(r, nx, ny, dx):
    np.testing.assert_allclose(P.local_weight_shape(r, dx, dx, nx, ny, 1.7),
                               E.local_weight_shape(r, dx, dx, nx, ny, 1.7), rtol=0, atol=1e-14)


def test_gaspari_cohn_equals_oracle():
    z = np.linspace(0, 2.5, 1001)
    np.testing.assert_allclose(P.gaspari_cohn(z), E.gaspari_cohn(z * 3.0, 3.0), rtol=0, atol=1e-14)


@pytest.mark.parametrize("seed", range(6))
def test_groups_equal_oracle(seed):
    rs = np.random.RandomState(seed)
    n = rs.randint(1, 70)
    pos = np.column_stack([rs.uniform(0, 120, n), rs.uniform(0, 90, n)])
    r = rs.uniform(0.5, 6.0)
    assert P.observation_groups(pos, r, 1.0, 1.0, 120, 90) == E.observation_groups(pos, r, 1.0, 1.0, 120, 90)


def test_lattice_groups_of_the_double_jet_buoys():
    """The reference's buoy layout (every 25 cells of the 500 x 300 double
    jet, r = 15): neighbours 25 cells apart are closer than 3 r = 45 cells."""
    ys, xs = np.mgrid[0:300:25, 0:500:25]
    pos = np.column_stack([(xs.ravel() + 0.5), (ys.ravel() + 0.5)])
    groups = P.observation_groups(pos, 15.0, 1.0, 1.0, 500, 300)
    assert groups == E.observation_groups(pos, 15.0, 1.0, 1.0, 500, 300)
    assert len(groups) > 1 and sorted(k for g in groups for k in g) == list(range(len(pos)))
    sparse = pos[(xs.ravel() % 50 == 0) & (ys.ravel() % 50 == 0)]
    assert len(P.observation_groups(sparse, 15.0, 1.0, 1.0, 500, 300)) == 1


@pytest.mark.parametrize("seed", range(4))
def test_tiles_list_every_covering_buoy(seed):
    rs = np.random.RandomState(seed)
    nx, ny, r = 70, 50, 2.5
    pos = np.column_stack([rs.uniform(0, nx, 25), rs.uniform(0, ny, 25)])
    W = P.local_weight_shape(r, 1.0, 1.0, nx, ny)
    n = W.shape[0]
    R = n // 2
    bx = (np.round(pos[:, 0]).astype(int) - R) % nx
    by = (np.round(pos[:, 1]).astype(int) - R) % ny
    jj, ii = np.nonzero(W)
    for g in P.observation_groups(pos, r, 1.0, 1.0, nx, ny):
        tiles, tile_obs, cover = P.letkf_tiles(g, bx, by, jj, ii, nx, ny, 64, 4)
        seen = {}
        for x0, y0, b, c in tiles:
            assert c > 0 and list(tile_obs[b:b + c]) == [o for o in g if o in set(tile_obs[b:b + c])]
            for y in range(y0, min(y0 + 4, ny)):
                for x in range(x0, min(x0 + 64, nx)):
                    seen[(x, y)] = set(tile_obs[b:b + c])
        worst = 0
        for y in range(ny):
            for x in range(nx):
                want = {o for o in g if (x - bx[o]) % nx < n and (y - by[o]) % ny < n
                        and W[(y - by[o]) % ny, (x - bx[o]) % nx] != 0}
                assert want <= seen.get((x, y), set())
                worst = max(worst, len(want))
        assert cover == worst
        # the combined weight the kernel forms equals the oracle's
        Wk = np.zeros((ny, nx))
        for o in g:
            Wk[(by[o] + jj) % ny, (bx[o] + ii) % nx] += W[jj, ii]
        np.testing.assert_allclose(Wk, E.combined_weights(pos[g], r, 1.0, 1.0, nx, ny), rtol=0, atol=1e-14)
