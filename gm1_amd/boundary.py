"""Boundary fills of the padded carriers (host side of ``boundary.hip``).

Mirrors the reference's family fills (``cuda_boundary_families.py:887-962``
``cpu_family_we/sn_inplace``, ``cpu_pressure_we/sn_inplace``), the rigid
``bcw`` lid/floor, ``bcwsfc`` (``cuda_terrain_open_boundary.py:641``) and the
open ``bc2d`` fill of ``w(k=1)`` (``cuda_momentum_rk.py:274-290``).
"""

from __future__ import annotations

from typing import Any

from . import _native
from ._device import lateral_code, ptr, stream_handle
from ._native import I, P, U, signature

KIND_SCALAR, KIND_U, KIND_V, KIND_W = 0, 1, 2, 3

signature("gm1hip_fill_family", P, I, U, I, I, I, P)
signature("gm1hip_fill_pressure", P, U, I, I, I, P)
signature("gm1hip_w_rigid", P, I, I, I, P)
signature("gm1hip_bcwsfc", P, P, P, P, P, P, I, I, P)
signature("gm1hip_bc2d_open", P, I, I, P)


def fill_family(array, kind: int, lateral_sides, *, ni: int, nj: int, nk: int, stream: Any | None = None) -> None:
    """bcs (kind 0), bcu (1), bcv (2) or bcw (3) ghost planes for a CM1 family."""
    if kind not in (0, 1, 2, 3):
        raise ValueError("kind must be 0 scalar, 1 u, 2 v or 3 w")
    _native.call("gm1hip_fill_family", ptr(array), kind, lateral_code(lateral_sides), ni, nj, nk,
                 stream_handle(stream))


def fill_pressure(array, lateral_sides, *, ni: int, nj: int, nk: int, stream: Any | None = None) -> None:
    """bcp: one-cell halo of a pressure-family carrier (nk+2, nj+2, ni+2)."""
    _native.call("gm1hip_fill_pressure", ptr(array), lateral_code(lateral_sides), ni, nj, nk, stream_handle(stream))


def w_rigid(w3d, *, ni: int, nj: int, nk: int, stream: Any | None = None) -> None:
    _native.call("gm1hip_w_rigid", ptr(w3d), ni, nj, nk, stream_handle(stream))


def bcwsfc(arrays, metrics, *, ni: int, nj: int, stream: Any | None = None) -> None:
    _native.call("gm1hip_bcwsfc", ptr(arrays["u3d"]), ptr(arrays["v3d"]), ptr(metrics["dzdx"]),
                 ptr(metrics["dzdy"]), ptr(metrics["gz"]), ptr(arrays["w3d"]), ni, nj, stream_handle(stream))


def bc2d_open(w3d, *, ni: int, nj: int, stream: Any | None = None) -> None:
    _native.call("gm1hip_bc2d_open", ptr(w3d), ni, nj, stream_handle(stream))
