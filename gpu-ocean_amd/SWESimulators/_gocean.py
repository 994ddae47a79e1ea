"""ctypes binding of ``libgocean_hip.so`` (declared in ``include/gocean.h``).

The HIP library is the only compute path: if it is missing or does not load,
every simulator call raises ``RuntimeError`` -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.abspath(os.path.join(_HERE, "..", "lib"))
LIB_PATH = os.path.join(LIB_DIR, "libgocean_hip.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int = ctypes.c_int
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p


class Domain(ctypes.Structure):
    """gocean_domain_t"""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("pitch", ctypes.c_int32), ("hi_pitch", ctypes.c_int32),
        ("dx", c_float), ("dy", c_float), ("g", c_float), ("r", c_float),
        ("theta", c_float), ("desing_eps", c_float), ("flux_slope_eps", c_float),
        ("depth_cutoff", c_float), ("land_value", c_float),
        ("f", c_float), ("north_x", c_float), ("north_y", c_float),
        ("bc_north", ctypes.c_int32), ("bc_east", ctypes.c_int32),
        ("bc_south", ctypes.c_int32), ("bc_west", ctypes.c_int32),
        ("rk_order", ctypes.c_int32),
    ]


class Wind(ctypes.Structure):
    """gocean_wind_t"""
    _fields_ = [("t", c_float), ("x0", c_float), ("x1", c_float),
                ("y0", c_float), ("y1", c_float),
                ("grid", c_void_p), ("grid_nx", ctypes.c_int32), ("grid_ny", ctypes.c_int32)]


class Fields(ctypes.Structure):
    """gocean_fields_t"""
    _fields_ = [("f", c_void_p), ("north_x", c_void_p), ("north_y", c_void_p)]


# (name, restype, argtypes) of every exported symbol declared in gocean.h
PROTOTYPES = [
    ("gocean_version", ctypes.c_char_p, []),
    ("gocean_last_error", ctypes.c_char_p, []),
    ("gocean_device_count", c_int, []),
    ("gocean_tile_configs", c_int, [ctypes.POINTER(ctypes.c_char_p), c_int]),
    ("gocean_bathymetry_init", c_int, [ctypes.POINTER(Domain), c_void_p, c_void_p, c_void_p]),
    ("gocean_cdklm_substep", c_int,
     [ctypes.POINTER(Domain), c_float, c_void_p, c_int,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p,
      ctypes.POINTER(Fields), ctypes.POINTER(Wind), c_int, c_void_p]),
    ("gocean_rk_steps", c_int,
     [ctypes.POINTER(Domain), c_int, c_float, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p,
      ctypes.POINTER(Fields), ctypes.POINTER(Wind), c_int, c_void_p]),
    ("gocean_periodic_bc", c_int, [ctypes.POINTER(Domain), c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gocean_max_dt", c_int,
     [ctypes.POINTER(Domain), c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    ("gocean_max_dt_static", c_int,
     [ctypes.POINTER(Domain), c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    ("gocean_default_tile_config", c_int, [c_int, c_int]),
    ("gocean_sqrt_newton_check", c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_ulonglong)]),
    ("gocean_adjoint_set_record_tile_config", c_int, [c_int]),
    ("gocean_adjoint_set_reverse_kernel", c_int, [c_int]),
    ("gocean_adjoint_reverse_kernel", c_int, []),
    ("gocean_adjoint_record_tile_config", c_int, []),
]

class Member(ctypes.Structure):
    """gocean_member_t"""
    _fields_ = [("domain", Domain), ("state_offset", ctypes.c_int64),
                ("hi_offset", ctypes.c_int64), ("hm_offset", ctypes.c_int64),
                ("field_offset", ctypes.c_int64), ("wind", Wind)]


PROTOTYPES.extend([
    ("gocean_ensemble_create", ctypes.c_int,
     [ctypes.POINTER(Member), ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("gocean_ensemble_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("gocean_ensemble_set_tile_config", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("gocean_ensemble_info", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
      ctypes.POINTER(ctypes.c_int)]),
    ("gocean_ensemble_bathymetry_init", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("gocean_ensemble_periodic_bc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("gocean_ensemble_max_dt", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    ("gocean_ensemble_substep", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("gocean_ensemble_rk_steps", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("gocean_ensemble_capture", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("gocean_ensemble_replay", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("gocean_adjoint_scratch_floats", ctypes.c_longlong, [ctypes.POINTER(Domain)]),
    ("gocean_adjoint_substep", c_int,
     [ctypes.POINTER(Domain), ctypes.POINTER(Wind), c_float, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_float, c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gocean_adjoint_rk2_step", c_int,
     [ctypes.POINTER(Domain), ctypes.POINTER(Wind), c_float, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gocean_ensemble_adjoint_setup", ctypes.c_longlong, [c_void_p, c_int]),
    ("gocean_ensemble_adjoint_segment", c_int,
     [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      ctypes.c_longlong, c_void_p]),
    ("gocean_ensemble_adjoint_record", c_int,
     [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_longlong, c_void_p, c_void_p]),
    ("gocean_ensemble_adjoint_reverse", c_int,
     [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      ctypes.c_longlong, c_void_p]),
    ("gocean_adjoint_segment_work_floats", ctypes.c_longlong, [ctypes.POINTER(Domain), c_int]),
    ("gocean_adjoint_segment", c_int,
     [ctypes.POINTER(Domain), ctypes.POINTER(Wind), c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, ctypes.c_longlong, c_void_p]),
    ("gocean_adjoint_record", c_int,
     [ctypes.POINTER(Domain), ctypes.POINTER(Wind), c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, ctypes.c_longlong, c_void_p]),
    ("gocean_adjoint_reverse", c_int,
     [ctypes.POINTER(Domain), c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_longlong, c_void_p]),
])


class Noise(ctypes.Structure):
    """gocean_noise_t"""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("pitch", ctypes.c_int32), ("hi_pitch", ctypes.c_int32),
        ("dx", c_float), ("dy", c_float),
        ("interpolation_factor", ctypes.c_int32),
        ("periodic_ns", ctypes.c_int32), ("periodic_ew", ctypes.c_int32),
        ("soar_q0", c_float), ("soar_L", c_float),
        ("g", c_float), ("f", c_float), ("beta", c_float),
        ("north_x", c_float), ("north_y", c_float),
        ("land_value", c_float),
    ]


class NoiseMember(ctypes.Structure):
    """gocean_noise_member_t"""
    _fields_ = [("noise", Noise), ("state_offset", ctypes.c_int64),
                ("hi_offset", ctypes.c_int64), ("field_offset", ctypes.c_int64)]


RNG_LCG_NORMAL, RNG_LCG_UNIFORM, RNG_PHILOX_NORMAL, RNG_PHILOX_UNIFORM = 0, 1, 2, 3

PROTOTYPES.extend([
    ("gocean_noise_sizes", c_int, [ctypes.POINTER(Noise), ctypes.POINTER(ctypes.c_int32)]),
    ("gocean_noise_generate", c_int,
     [ctypes.POINTER(Noise), c_int, c_void_p, c_void_p, ctypes.c_uint64, c_void_p, c_void_p]),
    ("gocean_noise_soar", c_int,
     [ctypes.POINTER(Noise), c_float, c_float, c_void_p, c_void_p, c_int, c_void_p]),
    ("gocean_noise_balance", c_int,
     [ctypes.POINTER(Noise), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      ctypes.POINTER(Fields), c_int, c_int, c_void_p]),
    ("gocean_noise_soar_balance", c_int,
     [ctypes.POINTER(Noise), c_float, c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      ctypes.POINTER(Fields), c_void_p]),
    ("gocean_noise_norms", c_int, [ctypes.POINTER(Noise), c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gocean_noise_make_perpendicular", c_int,
     [ctypes.POINTER(Noise), c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gocean_noise_batch_create", c_int,
     [ctypes.POINTER(NoiseMember), c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("gocean_noise_batch_destroy", c_int, [c_void_p]),
    ("gocean_noise_batch_layout", c_int,
     [c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ("gocean_noise_batch_generate", c_int,
     [c_void_p, c_void_p, c_void_p, ctypes.c_uint64, c_void_p, c_void_p]),
    ("gocean_noise_batch_perturb", c_int,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_int,
      c_void_p, c_void_p, c_void_p, ctypes.c_uint64, c_void_p, c_void_p]),
])


class DriftMember(ctypes.Structure):
    """gocean_drift_member_t"""
    _fields_ = [("domain", Domain), ("state_offset", ctypes.c_int64), ("hm_offset", ctypes.c_int64),
                ("positions_offset", ctypes.c_int64), ("observations_offset", ctypes.c_int64),
                ("num_drifters", ctypes.c_int32), ("sensitivity", c_float),
                ("wind_drift_factor", c_float), ("wind", Wind)]


PROTOTYPES.extend([
    ("gocean_drift", c_int,
     [ctypes.POINTER(Domain), c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float,
      ctypes.POINTER(Wind), c_float, c_int, c_int, c_void_p]),
    ("gocean_drifters_enforce_bc", c_int, [c_float, c_float, c_int, c_int, c_int, c_void_p, c_void_p]),
    ("gocean_observe_flow", c_int,
     [ctypes.POINTER(Domain), c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    ("gocean_drift_batch_create", c_int, [ctypes.POINTER(DriftMember), c_int, ctypes.POINTER(c_void_p)]),
    ("gocean_drift_batch_destroy", c_int, [c_void_p]),
    ("gocean_drift_batch_step", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("gocean_observe_flow_batch", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
])


class LetkfLayout(ctypes.Structure):
    """gocean_letkf_layout_t"""
    _fields_ = [("n_members", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("pitch", ctypes.c_int32), ("plane", ctypes.c_int64), ("member_stride", ctypes.c_int64)]


PROTOTYPES.extend([
    ("gocean_iewpf_dots", c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p]),
    ("gocean_iewpf_make_perpendicular", c_int,
     [c_void_p, c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p]),
    ("gocean_iewpf_axpby", c_int,
     [c_void_p, c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, c_float, c_void_p]),
    ("gocean_iewpf_local_svd", c_int,
     [c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p, ctypes.c_int32,
      c_void_p]),
    ("gocean_iewpf_kalman_patch", c_int,
     [c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p, ctypes.c_int32,
      ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_void_p]),
    ("gocean_iewpf_kalman_window", c_int,
     [c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, c_void_p, ctypes.c_int32,
      ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p, ctypes.c_int32, c_void_p, ctypes.c_int32, c_void_p,
      c_void_p, c_void_p]),
])

# sizeof(gocean_letkf_coef_t): two floats, two doubles
LETKF_COEF_BYTES = 24

PROTOTYPES.extend([
    ("gocean_letkf_tile_shape", c_int, [ctypes.POINTER(ctypes.c_int32)] * 4),
    ("gocean_letkf_coefficients", c_int,
     [ctypes.POINTER(LetkfLayout), c_void_p, c_void_p, ctypes.c_int32, c_void_p, c_void_p, ctypes.c_int32,
      ctypes.POINTER(ctypes.c_double), ctypes.c_double, c_void_p, c_void_p, c_void_p]),
    ("gocean_letkf_apply", c_int,
     [ctypes.POINTER(LetkfLayout), c_void_p, c_void_p, ctypes.c_int32, c_void_p, ctypes.c_int32, c_void_p,
      c_void_p, c_void_p, ctypes.c_int32, c_void_p, ctypes.c_int32, c_void_p]),
])



_lib = None
_lock = threading.Lock()


def library_exists() -> bool:
    return os.path.isfile(LIB_PATH)


def load(path: str = None):
    """Load (once) and return the ctypes handle of the HIP library."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.isfile(p):
            raise RuntimeError(
                f"HIP library {p} is missing: build it with "
                f"`python gpu-ocean_amd/build.py` (or __graft_entry__.build()). "
                f"There is no CPU fallback.")
        lib = ctypes.CDLL(p)
        for name, res, args in PROTOTYPES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return _lib


def check(rc: int, what: str = "gocean call"):
    """Turn a non-zero return code into RuntimeError with the library's
    message (the reference raises Python exceptions on launch failure)."""
    if rc != 0:
        msg = load().gocean_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (code {rc}): {msg}")


def ptr(t) -> int:
    """Device address of a torch tensor (or an int)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def stream_handle(stream) -> int:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream

