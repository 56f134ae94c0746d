"""Loader for the gfx950 HIP library ``libgm1hip.so`` (C ABI, via ctypes).

The product path has no CPU fallback: if the library is missing or a symbol is
absent, every operator raises.  Entry points take raw device pointers, sizes
and a ``hipStream_t`` and return a ``hipError_t`` code; :func:`check` turns a
non-zero code into ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIBRARY_PATH = Path(__file__).resolve().parent / "libgm1hip.so"

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None

P = ctypes.c_void_p
I = ctypes.c_int
U = ctypes.c_uint
Fl = ctypes.c_float
D = ctypes.c_double

# name -> argument types; entry points return int (hipError_t) unless a
# query registers another result type.
SIGNATURES: dict[str, tuple] = {}
RESTYPES: dict[str, object] = {}


def signature(name: str, *argtypes, restype=ctypes.c_int) -> None:
    SIGNATURES[name] = argtypes
    RESTYPES[name] = restype


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIBRARY_PATH.exists():
                raise RuntimeError(
                    f"{LIBRARY_PATH} is missing: build it with "
                    "`python -c 'import __graft_entry__; __graft_entry__.build()'`"
                )
            handle = ctypes.CDLL(str(LIBRARY_PATH), mode=ctypes.RTLD_GLOBAL)
            for name, argtypes in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.argtypes = list(argtypes)
                fn.restype = RESTYPES.get(name, ctypes.c_int)
            _lib = handle
    return _lib


def call(name: str, *args) -> None:
    fn = getattr(lib(), name)
    check(fn(*args), name)


def query(name: str, *args):
    """Call an entry point that returns a value rather than a status."""
    return getattr(lib(), name)(*args)


def check(code: int, name: str) -> None:
    if code != 0:
        raise RuntimeError(f"{name} failed with hipError_t {code}")


def exported_symbols() -> list[str]:
    return sorted(SIGNATURES)
