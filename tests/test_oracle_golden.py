"""Pin the CPU oracle against the reference test-suite's golden runs.

gpu_ocean/tests/schemes/CDKLM16_test.py:154-421 runs 13 cases (50x70 cells,
56 RK2 steps to t=50 s) and accepts a signed max difference of < 5e-4 of
max|ref| (assertAlmostEqual places=3).  The goldens were written by the CUDA
kernel with approximate division and square root, so they pin the oracle to
about 1e-4, not to the last bit; the measured agreement is asserted below.

The beta-plane golden was produced with an older Coriolis convention,
f + beta*(j+1/2)*dy; the current reference constructor builds
f - 2*beta*dy + beta*(j-3/2)*dy (CDKLM16.py:146-148, :295-302), which would
miss that golden by ~4e-2.  The legacy field is used to pin the per-cell
Coriolis path, and the current-code field is checked to reproduce that miss.
"""

import numpy as np
import pytest

from oracle import cdklm16_np as O
import helpers as H

GOLD_TOL = 1.0e-4       # measured worst case 5.5e-5 (normalised max |diff|)


def _run(bc, ic, f, bathy, f_field=None):
    eta, hu, hv, Hi = H.ref_initial(ic, bathy)
    p = O.Params(nx=H.REF_NX, ny=H.REF_NY, dx=H.REF_DX, dy=H.REF_DY, bc=bc,
                 f=np.float32(f) if f_field is None else f_field)
    sim = O.OracleSim(eta, hu, hv, Hi, p, dt=H.REF_DT)
    sim.step(H.REF_T)
    return sim


@pytest.mark.parametrize("case", H.ref_cases(), ids=[c[0] for c in H.ref_cases()])
def test_oracle_matches_reference_golden(case):
    name, bc, ic, stem, f, beta, bathy = case
    gold = H.load_goldens()
    f_field = H.legacy_beta_field(f, beta) if beta else None
    sim = _run(bc, ic, f, bathy, f_field)
    # the reference's step loop: 55 steps of 0.9 s and one of 0.5 s
    assert sim.num_iterations == 56
    assert sim.t == pytest.approx(50.0, abs=1e-9)
    err = H.golden_errors(sim.download(), gold, stem)
    for k, (signed, absmax, l2) in err.items():
        assert abs(signed) < 5e-4, (k, err)          # the reference's own bar
        assert absmax < GOLD_TOL, (k, err)


def test_beta_plane_current_convention_misses_legacy_golden():
    gold = H.load_goldens()
    f = 0.01 - 2 * 1e-6 * H.REF_DY
    sim = _run((1, 1, 1, 1), "central", f, False, H.current_beta_field(f, 1e-6))
    err = H.golden_errors(sim.download(), gold, "betamodel_central_")
    assert err["eta"][1] > 1e-2


def test_golden_fixture_shape():
    gold = H.load_goldens()
    assert len(gold.files) == 36
    for k in gold.files:
        assert gold[k].shape == (H.REF_NY, H.REF_NX)
        assert gold[k].dtype == np.float32
