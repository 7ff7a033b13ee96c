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
