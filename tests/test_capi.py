"""The C-ABI library loads, exports every symbol ``include/gocean.h`` declares,
and rejects invalid arguments with a message -- all without a GPU (the
argument checks run before any HIP call)."""

import ctypes
import os
import re

import pytest

from SWESimulators import _gocean as G
from SWESimulators import Common

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(gocean_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_built():
    assert G.library_exists(), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported_and_bound():
    lib = G.load()
    decl = declared_symbols()
    assert len(decl) >= 9
    bound = {n for n, _, _ in G.PROTOTYPES}
    for name in decl:
        assert hasattr(lib, name), name
        assert name in bound, f"{name} has no ctypes prototype"
    assert bound <= decl


def test_version_and_configs():
    lib = G.load()
    assert b"gfx950" in lib.gocean_version()
    names = (ctypes.c_char_p * 32)()
    n = lib.gocean_tile_configs(names, 32)
    assert n >= 4
    assert all(b"/" in names[i] for i in range(n))


def _dom(**kw):
    args = dict(nx=32, ny=20, dx=100.0, dy=100.0, g=9.81, r=0.0, theta=1.3, rk_order=2,
                bc=(1, 1, 1, 1))
    args.update(kw)
    return Common.make_domain(**args)


@pytest.mark.parametrize("kw,msg", [
    (dict(nx=1), "too small"),
    (dict(pitch=10), "pitch"),
    (dict(bc=(3, 1, 1, 1)), "not supported"),
    (dict(bc=(2, 1, 1, 1)), "paired"),
    (dict(rk_order=4), "rk_order"),
    (dict(rk_order=3, r=0.1), "r == 0"),
    (dict(dx=0.0), "positive"),
])
def test_invalid_domain_rejected(kw, msg):
    lib = G.load()
    rc = lib.gocean_periodic_bc(_dom(**kw), 16, 16, 16, None)
    assert rc == -1
    assert msg in lib.gocean_last_error().decode()


def test_null_and_alias_rejected():
    lib = G.load()
    d = _dom()
    assert lib.gocean_periodic_bc(d, None, None, None, None) == -1
    assert lib.gocean_bathymetry_init(d, None, None, None) == -1
    assert lib.gocean_max_dt(d, None, None, None, None, 0.7, None, None) == -1
    rc = lib.gocean_cdklm_substep(d, 1.0, None, 0, 16, 16, 16, None, None, None,
                                  16, 32, 48, 64, None, None, -1, None)
    assert rc == -1 and "alias" in lib.gocean_last_error().decode()
    rc = lib.gocean_cdklm_substep(d, 1.0, None, 2, 16, 16, 16, None, None, None,
                                  32, 32, 32, 64, None, None, -1, None)
    assert rc == -1 and "rk_step" in lib.gocean_last_error().decode()
    f = G.Fields()
    f.f = 16
    rc = lib.gocean_cdklm_substep(d, 1.0, None, 0, 16, 16, 16, None, None, None,
                                  32, 32, 32, 64, f, None, -1, None)
    assert rc == -1 and "together" in lib.gocean_last_error().decode()
    rc = lib.gocean_cdklm_substep(d, 1.0, None, 0, 16, 16, 16, None, None, None,
                                  32, 32, 32, 64, None, None, 99, None)
    assert rc == -1 and "tile configuration" in lib.gocean_last_error().decode()


def test_check_raises_with_message():
    lib = G.load()
    with pytest.raises(RuntimeError, match="too small"):
        G.check(lib.gocean_periodic_bc(_dom(nx=1), 16, 16, 16, None), "periodic")


def test_no_gpu_means_zero_devices_or_more():
    assert G.load().gocean_device_count() >= 0


@pytest.mark.parametrize("kw,msg", [
    (dict(rk_order=1), "rk_order 2"),
    (dict(r=0.01), "r == 0"),
    (dict(north_x=0.1, north_y=0.99), "rotation"),
])
def test_adjoint_scope_rejected(kw, msg):
    lib = G.load()
    d = _dom(**kw)
    rc = lib.gocean_adjoint_rk2_step(d, None, 1.0, None, *([16] * 6), None, 16, 16, 16, 32, 32, 32,
                                     64, 128, None)
    assert rc == -1 and msg in lib.gocean_last_error().decode()
    rc = lib.gocean_adjoint_substep(d, None, 1.0, None, *([16] * 6), None, None, None, None, None,
                                    1.0, 0.0, 32, 32, 32, 64, 128, None)
    assert rc == -1 and msg in lib.gocean_last_error().decode()


def test_adjoint_arguments_rejected():
    lib = G.load()
    d = _dom()
    assert lib.gocean_adjoint_scratch_floats(d) == 19 * (d.ny + 4) * d.pitch
    rc = lib.gocean_adjoint_rk2_step(d, None, 1.0, None, *([16] * 6), None, 16, 16, 16, 16, 32, 32,
                                     64, 128, None)
    assert rc == -1 and "alias" in lib.gocean_last_error().decode()
    rc = lib.gocean_adjoint_substep(d, None, 1.0, None, *([16] * 6), None, 16, None, None, None,
                                    1.0, 0.0, 32, 32, 32, 64, 128, None)
    assert rc == -1 and "extra" in lib.gocean_last_error().decode()
    rc = lib.gocean_adjoint_substep(d, None, 1.0, None, *([16] * 6), None, None, None, None, None,
                                    1.0, 0.0, 32, 32, 32, 64, None, None)
    assert rc == -1 and "NULL" in lib.gocean_last_error().decode()


def test_adjoint_segment_arguments_rejected():
    lib = G.load()
    d = _dom()
    pl = (d.ny + 4) * d.pitch
    assert lib.gocean_adjoint_segment_work_floats(d, 5) == (16 * 5 + 9) * pl
    assert lib.gocean_adjoint_segment_work_floats(d, 0) == -1
    rc = lib.gocean_adjoint_segment(d, None, 5, 1.0, None, 16, 16, 16, 32, 32, 32, 48, 48, 48, 64, 128,
                                    (16 * 5 + 9) * pl - 1, None)
    assert rc == -1 and "too small" in lib.gocean_last_error().decode()
    rc = lib.gocean_adjoint_segment(d, None, 0, 1.0, None, 16, 16, 16, 32, 32, 32, 48, 48, 48, 64, 128,
                                    1 << 40, None)
    assert rc == -1 and "n_steps" in lib.gocean_last_error().decode()
    rc = lib.gocean_adjoint_segment(_dom(rk_order=3), None, 2, 1.0, None, 16, 16, 16, 32, 32, 32, 48, 48,
                                    48, 64, 128, 1 << 40, None)
    assert rc == -1 and "rk_order 2" in lib.gocean_last_error().decode()


# ---------------------------------------------------------------------------
# model error
# ---------------------------------------------------------------------------
def _noise(**kw):
    d = G.Noise()
    vals = dict(nx=30, ny=40, pitch=34, hi_pitch=35, dx=7.0, dy=7.0, interpolation_factor=1,
                periodic_ns=1, periodic_ew=1, soar_q0=7e-5, soar_L=5.25, g=9.81, f=0.02, beta=0.0,
                north_x=0.0, north_y=1.0, land_value=1e20)
    vals.update(kw)
    for k, v in vals.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw,want", [
    (dict(), (30, 40, 30, 40, 15, 40, 0, 0)),
    (dict(periodic_ns=0, periodic_ew=0), (30, 40, 38, 48, 19, 48, 4, 4)),
    (dict(periodic_ew=0), (30, 40, 38, 40, 19, 40, 4, 0)),
    (dict(nx=45, ny=30, interpolation_factor=5, periodic_ns=0), (9, 6, 9, 14, 5, 14, 0, 4)),
])
def test_noise_sizes(kw, want):
    """OceanStateNoise.py:97-135 (and the odd rand_nx of a 9-wide coarse grid)."""
    lib = G.load()
    out = (ctypes.c_int32 * 8)()
    assert lib.gocean_noise_sizes(_noise(**kw), out) == 0
    assert tuple(out) == want


@pytest.mark.parametrize("kw,msg", [
    (dict(interpolation_factor=2), "odd"),
    (dict(interpolation_factor=7), "divisible"),
    (dict(nx=0), "too small"),
])
def test_noise_invalid_descriptor(kw, msg):
    lib = G.load()
    out = (ctypes.c_int32 * 8)()
    assert lib.gocean_noise_sizes(_noise(**kw), out) == -1
    assert msg in lib.gocean_last_error().decode()


def test_noise_arguments_rejected():
    lib = G.load()
    d = _noise()
    assert lib.gocean_noise_generate(d, 7, 16, 16, 0, 16, None) == -1
    assert "kind" in lib.gocean_last_error().decode()
    assert lib.gocean_noise_generate(d, G.RNG_LCG_NORMAL, None, 16, 0, None, None) == -1
    assert "seed" in lib.gocean_last_error().decode()
    assert lib.gocean_noise_generate(d, G.RNG_PHILOX_NORMAL, None, 16, 0, None, None) == -1
    assert "counter" in lib.gocean_last_error().decode()
    assert lib.gocean_noise_soar(_noise(soar_L=0.0), 1.0, 1.0, 16, 16, 0, None) == -1
    assert "soar_L" in lib.gocean_last_error().decode()
    assert lib.gocean_noise_balance(_noise(f=0.0), 16, 16, 16, 16, 16, None, 0, 0, None) == -1
    assert "Coriolis field" in lib.gocean_last_error().decode()
    assert lib.gocean_noise_balance(_noise(pitch=20), 16, 16, 16, 16, 16, None, 0, 0, None) == -1
    assert "pitch" in lib.gocean_last_error().decode()
    d5 = _noise(nx=30, ny=40, interpolation_factor=5)
    assert lib.gocean_noise_balance(d5, 16, 16, 16, 16, 16, None, 3, 0, None) == -1
    assert "offsets" in lib.gocean_last_error().decode()
    f = G.Fields()
    f.north_x = 16
    assert lib.gocean_noise_balance(d, 16, 16, 16, 16, 16, f, 0, 0, None) == -1
    assert "together" in lib.gocean_last_error().decode()


def test_noise_batch_rejects_bad_member():
    lib = G.load()
    arr = (G.NoiseMember * 2)()
    arr[0].noise = _noise()
    arr[1].noise = _noise(interpolation_factor=4)
    h = ctypes.c_void_p()
    assert lib.gocean_noise_batch_create(arr, 2, G.RNG_LCG_NORMAL, ctypes.byref(h)) == -1
    assert "member 1" in lib.gocean_last_error().decode()
    arr[1].noise = _noise(f=0.0)
    arr[1].field_offset = -1
    assert lib.gocean_noise_batch_create(arr, 2, G.RNG_LCG_NORMAL, ctypes.byref(h)) == -1
    assert "Coriolis" in lib.gocean_last_error().decode()
    assert lib.gocean_noise_batch_create(arr, 2, 9, ctypes.byref(h)) == -1


def test_letkf_shape_and_argument_checks():
    lib = G.load()
    v = [ctypes.c_int32() for _ in range(4)]
    assert lib.gocean_letkf_tile_shape(*[ctypes.byref(x) for x in v]) == 0
    assert [x.value for x in v] == [64, 1, 4, 1024]
    lay = G.LetkfLayout(8, 30, 20, 34, 24 * 34, 3 * 24 * 34)
    cov = (ctypes.c_double * 4)(0.1, 0.0, 0.0, 0.2)
    ok = dict(state=16, group=16, cells=16, y=16, coef=16, oc=16)

    def coef_call(layout=lay, cov=cov, infl=1.0, **kw):
        a = dict(ok, **kw)
        return lib.gocean_letkf_coefficients(layout, a["state"], a["group"], 1, a["cells"], a["y"], 4, cov, infl,
                                             a["coef"], a["oc"], None)

    assert coef_call(layout=None) == -1 and "layout" in lib.gocean_last_error().decode()
    assert coef_call(layout=G.LetkfLayout(1, 30, 20, 34, 24 * 34, 3 * 24 * 34)) == -1
    assert "members" in lib.gocean_last_error().decode()
    assert coef_call(layout=G.LetkfLayout(8, 30, 20, 33, 24 * 34, 3 * 24 * 34)) == -1
    assert "pitch" in lib.gocean_last_error().decode()
    assert coef_call(layout=G.LetkfLayout(8, 30, 20, 34, 24 * 34, 2 * 24 * 34)) == -1
    assert "stride" in lib.gocean_last_error().decode()
    assert coef_call(state=None) == -1 and "NULL" in lib.gocean_last_error().decode()
    assert coef_call(cov=(ctypes.c_double * 4)(0.1, 0.3, 0.3, 0.2)) == -1
    assert "positive definite" in lib.gocean_last_error().decode()
    assert coef_call(cov=(ctypes.c_double * 4)(0.1, 0.01, 0.0, 0.2)) == -1
    assert coef_call(infl=-1.0) == -1 and "inflation" in lib.gocean_last_error().decode()
    lay2 = G.LetkfLayout(2, 30, 20, 34, 24 * 34, 3 * 24 * 34)
    assert coef_call(layout=lay2, infl=0.0) == -1 and "adaptive" in lib.gocean_last_error().decode()

    def apply_call(local_n=5, n_tiles=1, tiles=16):
        return lib.gocean_letkf_apply(lay, 16, 16, local_n, 16, 4, 16, 16, tiles, n_tiles, 16, 3, None)

    assert apply_call(local_n=4) == -1 and "odd" in lib.gocean_last_error().decode()
    assert apply_call(local_n=21) == -1 and "fit" in lib.gocean_last_error().decode()
    assert apply_call(tiles=None) == -1 and "NULL" in lib.gocean_last_error().decode()
    assert apply_call(n_tiles=-1) == -1


def test_iewpf_kalman_window_argument_checks():
    lib = G.load()
    plane = 24 * 34

    def call(n=4, nx=30, ny=20, stride=3 * plane, nd=2, wh=5, ww=7, n_tiles=3, state=16, lst=16, cs=32):
        return lib.gocean_iewpf_kalman_window(state, n, nx, ny, stride, 16, nd, wh, ww, 16, 16, cs, 16, n_tiles,
                                              16, lst, None)

    assert call(n=0) == -1 and "members" in lib.gocean_last_error().decode()
    assert call(stride=3 * plane - 1) == -1 and "stride" in lib.gocean_last_error().decode()
    assert call(wh=21) == -1 and "window" in lib.gocean_last_error().decode()
    assert call(ww=0) == -1 and "window" in lib.gocean_last_error().decode()
    assert call(nd=0) == -1 and "without drifters" in lib.gocean_last_error().decode()
    assert call(n_tiles=-1) == -1
    assert call(state=None) == -1 and "NULL" in lib.gocean_last_error().decode()
    assert call(lst=None) == -1 and "NULL" in lib.gocean_last_error().decode()
    assert call(n=33, cs=47) == -1 and "coefficient stride" in lib.gocean_last_error().decode()
    assert call(n_tiles=0) == 0                          # nothing to add, no launch


def test_sqrt_newton_check_arguments_rejected():
    lib = G.load()
    n = ctypes.c_ulonglong(7)
    assert lib.gocean_sqrt_newton_check(0x3F800000, 0x3F000000, ctypes.byref(n)) == -1
    assert b"not within" in lib.gocean_last_error()
    assert lib.gocean_sqrt_newton_check(0x3F800000, 0x7F800000, ctypes.byref(n)) == -1   # +inf excluded
    assert lib.gocean_sqrt_newton_check(0x3F800000, 0x3F800001, None) == -1
    assert n.value == 7


def test_marching_config_rejects_planes_of_4gib():
    """The column-marching kernels address planes with 32-bit byte offsets:
    an explicit marching configuration on a larger plane is an error (the
    default configuration falls back to the tile kernel instead)."""
    lib = G.load()
    d = _dom(pitch=1 << 26)                    # 24 rows x 256 MiB
    rc = lib.gocean_cdklm_substep(d, 1.0, None, 0, 16, 16, 16, None, None, None,
                                  32, 32, 32, 64, None, None, 9, None)
    assert rc == -1 and "4 GiB" in lib.gocean_last_error().decode()
    m = (G.Member * 1)()
    m[0].domain = d
    m[0].field_offset = -1
    h = ctypes.c_void_p()
    assert lib.gocean_ensemble_create(m, 1, 9, ctypes.byref(h)) == -1
    assert b"4 GiB" in lib.gocean_last_error()
