"""The C ABI: the built library loads and exports every entry point the header declares.

No compute call is made here (no GPU is needed); the GPU tests call each
entry point through ctypes.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gm1hip.h"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|int64_t)\s+(gm1hip_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_every_registered_signature():
    from gm1_amd import _native

    declared = set(declared_symbols())
    registered = set(_native.exported_symbols())
    assert registered <= declared, f"registered but not declared: {sorted(registered - declared)}"
    assert declared <= registered, f"declared but without a Python signature: {sorted(declared - registered)}"


def test_library_exports_every_declared_symbol():
    from gm1_amd import _native

    if not _native.LIBRARY_PATH.exists():
        pytest.fail(f"{_native.LIBRARY_PATH} is not built; run __graft_entry__.build()")
    handle = ctypes.CDLL(str(_native.LIBRARY_PATH))
    missing = [name for name in declared_symbols() if not hasattr(handle, name)]
    assert not missing, f"library lacks {missing}"


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """The product path has no fallback: a missing library raises before any launch."""
    from gm1_amd import _native

    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIBRARY_PATH", tmp_path / "absent.so")
    with pytest.raises(RuntimeError, match="missing"):
        _native.lib()
