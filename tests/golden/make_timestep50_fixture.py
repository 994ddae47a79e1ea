"""Write ``timestep50_cdklm16.npz`` from the reference test-suite's golden runs.

Source: gpu_ocean/tests/timestep50/CDKLM16_*.dat (text dumps of the padded
(ny+4, nx+4) = (74, 54) eta, hu, hv after ``sim.step(50.0)``, loaded by
gpu_ocean/tests/testUtils.py:175-180 and compared by
gpu_ocean/tests/schemes/CDKLM16_test.py:114-150).  The fixture keeps the
interior cells as float32 -- data only, no code.

Run from the repository root (needs the reference checkout):
    python tests/golden/make_timestep50_fixture.py
"""
import os
import sys

import numpy as np

REF = os.environ.get("GOCEAN_REFERENCE", "/root/reference")
SRC = os.path.join(REF, "gpu_ocean", "tests", "timestep50")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "timestep50_cdklm16.npz")

# golden file stems used by the 13 CDKLM16 regression cases
STEMS = [
    "wallBC_central_", "wallBC_corner_", "wallBC_upperCorner_",
    "periodic_corner_", "periodic_upperCorner_",
    "periodicNS_corner_", "periodicNS_upperCorner_",
    "periodicEW_corner_", "periodicEW_upperCorner_",
    "coriolis_central_", "betamodel_central_", "wallBC_central_bathymetry_",
]


def main():
    if not os.path.isdir(SRC):
        print("reference golden directory not found:", SRC, file=sys.stderr)
        return 1
    out = {}
    for stem in STEMS:
        for k in ("eta", "u", "v"):
            a = np.loadtxt(os.path.join(SRC, "CDKLM16_" + stem + k + ".dat"))
            assert a.shape == (74, 54), a.shape
            out[stem + k] = a[2:-2, 2:-2].astype(np.float32)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    return 0


if __name__ == "__main__":
    sys.exit(main())
