"""ctypes binding of ``libsomaxhip.so`` (the C ABI declared in include/somaxhip.h).

This is the binding a JAX-less host uses; under JAX the same entry points are
registered as ``jax.ffi`` custom calls.  Loading fails loudly: there is no
Python or CPU fallback for any computation routed through this module.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libsomaxhip.so"

SOMAXHIP_OK = 0
ADV_ARAKAWA = 0
ADV_WENO = 1
PREC_FP32 = 0
PREC_MIXED = 1
PREC_BF16X6 = 2
PREC_FP64 = 3
ENS_CONCURRENT = 1
ENS_GRAPH = 2
ENS_MERGED = 4
SOLVER_SSPRK3 = 0
SOLVER_TSIT5 = 1
BC_PERIODIC = 0
BC_WALL = 1
BC_CHANNEL = 2
NOALLOC_SCRATCH_BYTES = 64 << 20  # SOMAXHIP_NOALLOC_SCRATCH_BYTES
STATS_MAX_FIELDS = 64


class SomaxHipError(RuntimeError):
    """Raised when a somaxhip entry point returns a non-zero status."""


c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)


class QGDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nl", ctypes.c_int),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("cl2m", c_double_p),
        ("cm2l", c_double_p),
        ("lambdas", c_double_p),
        ("beta_y_row", c_double_p),
        ("wind_row", c_double_p),
        ("wind_scale", c_double_p),
        ("zero_psi_ghost", ctypes.c_int),
        ("advection", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("mass_correction", ctypes.c_int),
        ("homsol", c_float_p),
        ("homsol_mean", c_double_p),
    ]


class QGParams(ctypes.Structure):
    _fields_ = [
        ("nu", ctypes.c_double),
        ("nu4", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("tau0", ctypes.c_double),
    ]


class SWMDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nl", ctypes.c_int),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("g_prime", c_double_p),
        ("H", c_double_p),
        ("f_row", c_double_p),
        ("wind_row", c_double_p),
        ("bc", ctypes.c_int),
        ("vorticity_flux", ctypes.c_int),
    ]


class SWMParams(ctypes.Structure):
    _fields_ = [
        ("nu", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("tau0", ctypes.c_double),
    ]


class KernelStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("launches", ctypes.c_int),
        ("total_ms", ctypes.c_double),
        ("flops", ctypes.c_double),
        ("bytes", ctypes.c_double),
    ]


# Every exported symbol and its signature: (restype, argtypes).
_vp = ctypes.c_void_p
_SIGNATURES = {
    "somaxhip_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "somaxhip_init": (ctypes.c_int, []),
    "somaxhip_version": (ctypes.c_int, []),
    "somaxhip_qg_plan_create": (ctypes.c_int, [ctypes.POINTER(QGDesc), ctypes.POINTER(_vp)]),
    "somaxhip_qg_plan_destroy": (ctypes.c_int, [_vp]),
    "somaxhip_qg_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "somaxhip_qg_invert": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "somaxhip_qg_rhs": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(QGParams), ctypes.c_int, _vp, _vp]),
    "somaxhip_qg_step": (
        ctypes.c_int,
        [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(QGParams), _vp, _vp],
    ),
    "somaxhip_qg_step_noalloc": (
        ctypes.c_int,
        [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(QGParams), _vp],
    ),
    "somaxhip_qg_tsit5_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "somaxhip_qg_step_tsit5": (
        ctypes.c_int,
        [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(QGParams), _vp, _vp, _vp],
    ),
    "somaxhip_zero_ghost": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    "somaxhip_qg_diagnose": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "somaxhip_qg_invert_f64": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "somaxhip_qg_rhs_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(QGParams), ctypes.c_int, _vp, _vp]),
    "somaxhip_qg_step_f64": (
        ctypes.c_int,
        [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(QGParams), _vp, _vp],
    ),
    "somaxhip_qg_step_tsit5_f64": (
        ctypes.c_int,
        [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(QGParams), _vp, _vp, _vp],
    ),
    "somaxhip_qg_diagnose_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "somaxhip_zero_ghost_f64": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    "somaxhip_qg_step_profile": (
        ctypes.c_int,
        [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(QGParams), _vp, _vp, ctypes.POINTER(KernelStat),
         ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    ),
    "somaxhip_swm_plan_create": (ctypes.c_int, [ctypes.POINTER(SWMDesc), ctypes.POINTER(_vp)]),
    "somaxhip_swm_plan_destroy": (ctypes.c_int, [_vp]),
    "somaxhip_swm_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "somaxhip_swm_rhs": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(SWMParams), ctypes.c_int, _vp]),
    "somaxhip_swm_apply_bc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "somaxhip_swm_diagnose": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "somaxhip_swm_step": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(SWMParams), _vp, _vp],
    ),
    "somaxhip_swm_tsit5_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "somaxhip_swm_step_tsit5": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(SWMParams), _vp, _vp],
    ),
    "somaxhip_qg_project": (
        ctypes.c_int,
        [_vp, ctypes.c_double, c_double_p, c_double_p, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp],
    ),
    "somaxhip_rqg_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "somaxhip_rqg_step": (
        ctypes.c_int,
        [_vp, _vp, ctypes.c_int, ctypes.c_double, c_double_p, c_double_p, _vp, _vp, _vp, ctypes.c_int,
         ctypes.c_double, ctypes.POINTER(SWMParams), _vp, _vp, _vp],
    ),
    "somaxhip_swm_step_profile": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(SWMParams), _vp, _vp,
         ctypes.POINTER(KernelStat), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    ),
    "somaxhip_cmm_create": (
        ctypes.c_int,
        [_vp, c_float_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), c_double_p, ctypes.POINTER(_vp)],
    ),
    "somaxhip_cmm_destroy": (ctypes.c_int, [_vp]),
    "somaxhip_cmm_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "somaxhip_cmm_solve": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "somaxhip_stats_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int]),
    "somaxhip_field_stats": (
        ctypes.c_int,
        [ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, _vp, _vp, _vp],
    ),
    "somaxhip_ens_create": (
        ctypes.c_int,
        [ctypes.POINTER(QGDesc), ctypes.POINTER(QGParams), ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
    ),
    "somaxhip_ens_destroy": (ctypes.c_int, [_vp]),
    "somaxhip_ens_state_floats": (ctypes.c_size_t, [_vp]),
    "somaxhip_ens_member_offset": (ctypes.c_long, [_vp, ctypes.c_int]),
    "somaxhip_ens_workspace_bytes": (ctypes.c_size_t, [_vp]),
    "somaxhip_ens_num_classes": (ctypes.c_int, [_vp]),
    "somaxhip_ens_class_info": (ctypes.c_int, [_vp, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 6),
    "somaxhip_ens_set_stage_rows": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "somaxhip_ens_set_params": (ctypes.c_int, [_vp, ctypes.POINTER(QGParams), _vp]),
    "somaxhip_ens_step": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp]),
    "somaxhip_ens_step_profile": (
        ctypes.c_int,
        [_vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp, ctypes.POINTER(KernelStat), ctypes.c_int,
         ctypes.POINTER(ctypes.c_int)],
    ),
}

_lib = None


def library_path() -> Path:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise SomaxHipError(
            f"{_LIB_PATH} is missing: build it with `python -m somax_amd.build` "
            "(or __graft_entry__.build()); there is no fallback implementation"
        )
    lib = ctypes.CDLL(str(_LIB_PATH), mode=os.RTLD_NOW | os.RTLD_LOCAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(_SIGNATURES)


def last_error() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(4096)
    lib.somaxhip_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != SOMAXHIP_OK:
        raise SomaxHipError(f"{what} failed (status {rc}): {last_error()}")


_initialised = False


def init_device() -> None:
    """Check the current device is gfx950 (once per process)."""
    global _initialised
    if _initialised:
        return
    check(load().somaxhip_init(), "somaxhip_init")
    _initialised = True
