"""Host-side checks of bench.py's measurement contract (no GPU): the
committed counter summary it reads for ``roofline.traffic`` and
``extra.valu`` covers every marching configuration the autotuner can pick
at the 2048 x 2048 headline size, its byte counts are self-consistent, and
a multi-GPU request is refused with a message."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

CELLS = 2048 * 2048
MARCH = ("march60x8/256", "march60x16/256", "march60xauto/256")


def _counters():
    with open(os.path.join(ROOT, "profiles", "cdklm16_substep_2048_counters.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("tile", MARCH + ("32x16/256",))
def test_traffic_and_valu_cover_the_timed_tiles(tile):
    t = bench.measured_traffic(CELLS, tile)
    assert t is not None and t > 0
    # the measured traffic is at least the algorithmic 34 B per cell-substep
    assert t >= 34 * CELLS
    v = bench.measured_valu(CELLS, tile, 50e-6)
    assert v is not None and v["valu_wave_instructions_per_launch"] > 0
    assert v["valu_lane_ops_per_cell"] == pytest.approx(v["valu_wave_instructions_per_launch"] * 64 / CELLS, abs=0.1)


def test_counter_summary_is_consistent():
    d = _counters()
    assert d["cells"] == CELLS and d["algorithmic_bytes_per_launch"] == 34 * CELLS
    cal = d["calibration_copy_256MiB"]
    # the copy's read requests are 128-byte lines: 256 MiB / 128 B of them
    assert cal["TCC_EA0_RDREQ"] == pytest.approx(256 * 2**20 / 128, rel=1e-3)
    assert cal["WRITE_SIZE_KiB"] == 256 * 1024
    for name in MARCH:
        e = d["tiles"][name]
        assert e["bytes_per_launch"] == pytest.approx(e["fetch_bytes_per_launch"] + e["write_bytes_per_launch"], abs=2)
        # bytes read from the doubled FETCH_SIZE agree with the request count x 128 B
        assert e["fetch_bytes_per_launch"] == pytest.approx(e["fetch_bytes_from_rdreq"], rel=1e-3)
        assert e["traffic_over_algorithmic"] == pytest.approx(e["bytes_per_launch"] / (34 * CELLS), abs=1e-3)
    # only the marching configurations of a different size are absent
    assert bench.measured_traffic(1024 * 1024, "march60xauto/256") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
