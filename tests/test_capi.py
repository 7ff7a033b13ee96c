"""The C-ABI library loads and exports every entry point include/somaxhip.h
declares (no compute: this runs without a GPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


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
