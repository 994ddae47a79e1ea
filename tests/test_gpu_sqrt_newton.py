"""The square root that the substep kernels take for the wave speeds
sqrt(g h) when their fast arithmetic is on (hardware reciprocal square root
plus one Newton step) against the correctly rounded square root, over every
float bit pattern of the range the host allows it for: [2^-60, FLT_MAX].  The
host enables it only when RN(g * depth_cutoff) >= 2^-60, and the kernels take
it only for depths above the cutoff.  The lowest binade of the normal range
is where the two sequences differ: that part shows the check can fail."""

import ctypes

import pytest

from SWESimulators import _gocean as G

pytestmark = pytest.mark.gpu

TWO_M60 = 0x21800000          # bits of 2^-60
FLT_MAX = 0x7F7FFFFF


def _mismatches(lo, hi):
    lib = G.load()
    n = ctypes.c_ulonglong(123)
    G.check(lib.gocean_sqrt_newton_check(lo, hi, ctypes.byref(n)), "sqrt check")
    return n.value


def test_newton_sqrt_exact_on_allowed_range():
    print("sqrt check [2^-60, FLT_MAX]", flush=True)
    assert _mismatches(TWO_M60, FLT_MAX) == 0


def test_check_detects_differences_below_the_range():
    print("sqrt check lowest binade", flush=True)
    assert _mismatches(0x00800000, 0x00FFFFFF) > 0


def test_single_values():
    print("sqrt check single values", flush=True)
    for bits in (0x3F800000, 0x40800000, 0x461C4000, TWO_M60, FLT_MAX):   # 1, 4, 1e4, 2^-60, max
        assert _mismatches(bits, bits) == 0
