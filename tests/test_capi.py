"""The C-ABI library loads and exports every entry point include/somaxhip.h
declares (no compute: this runs without a GPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ERR_ARG = 1  # SOMAXHIP_ERR_ARG


def _declared():
    text = (ROOT / "include" / "somaxhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(somaxhip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "somaxhip_qg_step" in names and "somaxhip_qg_plan_create" in names


def test_library_exports_every_declared_symbol():
    from somax_amd import _lib

    path = _lib.library_path()
    if not path.exists():
        pytest.fail(f"{path} not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(str(path))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    from somax_amd import _lib

    assert sorted(_lib.exported_symbols()) == _declared()


def test_version():
    from somax_amd import _lib

    assert _lib.load().somaxhip_version() >= 100


@pytest.mark.parametrize("kind", ["qg", "swm"])
def test_plan_create_rejects_planes_of_2_31_cells(kind):
    """Argument checks run before any device call, so this needs no GPU: a
    plane of (nx + 2) * (ny + 2) >= 2^31 cells is refused (the kernels index a
    plane with 32-bit offsets), one just below passes the size check and
    fails on the next one (null tables)."""
    from somax_amd import _lib

    lib = _lib.load()
    plan = ctypes.c_void_p()
    desc_t, create = ((_lib.QGDesc, lib.somaxhip_qg_plan_create) if kind == "qg"
                      else (_lib.SWMDesc, lib.somaxhip_swm_plan_create))
    d = desc_t()
    d.nx, d.ny, d.nl, d.dx, d.dy = 46340, 46340, 1, 1.0, 1.0  # 46342^2 > 2^31 - 1
    assert create(ctypes.byref(d), ctypes.byref(plan)) == ERR_ARG
    assert "2^31" in _lib.last_error()
    d.nx = d.ny = 46338  # 46340^2 < 2^31 - 1
    assert create(ctypes.byref(d), ctypes.byref(plan)) == ERR_ARG
    assert "2^31" not in _lib.last_error()


def test_noalloc_step_checks_arguments_first():
    """The no-allocation step refuses a null plan before touching a device."""
    from somax_amd import _lib

    lib = _lib.load()
    prm = _lib.QGParams(0.0, 0.0, 0.0, 0.0)
    assert lib.somaxhip_qg_step_noalloc(None, None, None, 1, 1200.0, ctypes.byref(prm), None) == ERR_ARG
    assert "somaxhip_qg_step_noalloc" in _lib.last_error()
    header = (ROOT / "include" / "somaxhip.h").read_text()
    assert f"#define SOMAXHIP_NOALLOC_SCRATCH_BYTES ({_lib.NOALLOC_SCRATCH_BYTES >> 20}u << 20)" in header


def _qg_desc(lib_mod, keep):
    """A valid 16 x 16, 2-layer FP32 description; ``keep`` holds the arrays."""
    d = lib_mod.QGDesc()
    d.nx, d.ny, d.nl, d.dx, d.dy = 16, 16, 2, 1e4, 1e4
    tabs = {name: (ctypes.c_double * n)(*([1.0] * n)) for name, n in
            (("cl2m", 4), ("cm2l", 4), ("lambdas", 2), ("beta_y_row", 18), ("wind_row", 18), ("wind_scale", 2))}
    keep.append(tabs)
    for name, arr in tabs.items():
        setattr(d, name, ctypes.cast(arr, lib_mod.c_double_p))
    d.advection, d.precision = lib_mod.ADV_ARAKAWA, lib_mod.PREC_FP32
    return d


@pytest.mark.parametrize("field,value,message", [
    ("nx", 1, "nx, ny >= 2"), ("ny", 0, "nx, ny >= 2"), ("nl", 0, "1 <= nl <= 8"), ("nl", 9, "1 <= nl <= 8"),
    ("dx", 0.0, "dx, dy > 0"), ("dy", float("nan"), "dx, dy > 0"), ("lambdas", None, "table pointer is null"),
    ("wind_scale", None, "table pointer is null"), ("advection", 7, "unknown advection"),
    ("precision", 9, "unknown precision"), ("mass_correction", 1, "needs homsol"),
])
def test_qg_plan_create_refuses_bad_descriptions(field, value, message):
    """Each check of a QG description runs before any device call (no GPU
    needed): one bad field at a time, the rest valid."""
    from somax_amd import _lib

    lib = _lib.load()
    keep = []
    d = _qg_desc(_lib, keep)
    setattr(d, field, value)
    plan = ctypes.c_void_p()
    assert lib.somaxhip_qg_plan_create(ctypes.byref(d), ctypes.byref(plan)) == ERR_ARG
    assert message in _lib.last_error()
    assert plan.value is None
    assert lib.somaxhip_qg_plan_create(None, ctypes.byref(plan)) == ERR_ARG
    assert lib.somaxhip_qg_plan_create(ctypes.byref(d), None) == ERR_ARG


def test_entry_points_refuse_a_null_plan():
    """Steps, inversions and diagnostics refuse a null plan before touching
    a device."""
    from somax_amd import _lib

    lib = _lib.load()
    prm = _lib.QGParams(0.0, 0.0, 0.0, 0.0)
    calls = [
        ("somaxhip_qg_step", lambda: lib.somaxhip_qg_step(None, None, None, 1, 1200.0, ctypes.byref(prm), None, None)),
        ("somaxhip_qg_step_f64",
         lambda: lib.somaxhip_qg_step_f64(None, None, None, 1, 1200.0, ctypes.byref(prm), None, None)),
        ("somaxhip_qg_invert", lambda: lib.somaxhip_qg_invert(None, None, None, 1, None, None)),
    ]
    for name, call in calls:
        assert call() == ERR_ARG, name
        assert _lib.last_error(), name


@pytest.mark.parametrize("field,value,message", [
    ("nl", 9, "1 <= nl <= 8"), ("nx", 1, "nx, ny >= 2"), ("dx", -1.0, "dx, dy > 0"), ("f_row", None, "null"),
    ("bc", 7, "unknown boundary condition"), ("H0", 0.0, "H[0] must be positive"),
    ("vorticity_flux", 5, "unknown vorticity flux"),
])
def test_swm_plan_create_refuses_bad_descriptions(field, value, message):
    """The shallow-water description's checks, before any device call."""
    from somax_amd import _lib

    lib = _lib.load()
    d = _lib.SWMDesc()
    d.nx, d.ny, d.nl, d.dx, d.dy = 16, 16, 2, 1e4, 1e4
    tabs = {"g_prime": (ctypes.c_double * 2)(9.81, 0.02), "H": (ctypes.c_double * 2)(500.0, 3500.0),
            "f_row": (ctypes.c_double * 18)(*([1e-4] * 18)), "wind_row": (ctypes.c_double * 18)(*([0.0] * 18))}
    for name, arr in tabs.items():
        setattr(d, name, ctypes.cast(arr, _lib.c_double_p))
    d.bc, d.vorticity_flux = _lib.BC_WALL, 0
    if field == "H0":
        tabs["H"][0] = value
    else:
        setattr(d, field, value)
    plan = ctypes.c_void_p()
    assert lib.somaxhip_swm_plan_create(ctypes.byref(d), ctypes.byref(plan)) == ERR_ARG
    assert message in _lib.last_error()
    assert plan.value is None


def test_ensemble_create_refuses_bad_members():
    """An ensemble refuses no members, unknown flags and (naming the member)
    an invalid or FP64 member description, before any device call."""
    from somax_amd import _lib

    lib = _lib.load()
    keep = []
    members = (_lib.QGDesc * 3)(*(_qg_desc(_lib, keep) for _ in range(3)))
    params = (_lib.QGParams * 3)(*(_lib.QGParams(0.0, 0.0, 0.0, 0.0) for _ in range(3)))
    ens = ctypes.c_void_p()
    assert lib.somaxhip_ens_create(members, params, 0, 0, ctypes.byref(ens)) == ERR_ARG
    assert lib.somaxhip_ens_create(members, params, 3, 1 << 10, ctypes.byref(ens)) == ERR_ARG
    assert "unknown flags" in _lib.last_error()
    members[1].nl = 9
    assert lib.somaxhip_ens_create(members, params, 3, 0, ctypes.byref(ens)) == ERR_ARG
    assert "member 1" in _lib.last_error() and "nl" in _lib.last_error()
    members[1].nl = 2
    members[2].precision = _lib.PREC_FP64
    assert lib.somaxhip_ens_create(members, params, 3, 0, ctypes.byref(ens)) == ERR_ARG
    assert "member 2" in _lib.last_error() and "FP64" in _lib.last_error()
    assert ens.value is None
