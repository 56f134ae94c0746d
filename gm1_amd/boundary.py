"""Boundary fills of the padded carriers (host side of ``boundary.hip``).

Mirrors the reference's family fills (``cuda_boundary_families.py:887-962``
``cpu_family_we/sn_inplace``, ``cpu_pressure_we/sn_inplace``), the rigid
``bcw`` lid/floor, ``bcwsfc`` (``cuda_terrain_open_boundary.py:641``) and the
open ``bc2d`` fill of ``w(k=1)`` (``cuda_momentum_rk.py:274-290``).

Every entry checks what the reference's launchers check before anything is
enqueued: the grid (``_validate_dimensions``, ``cuda_boundary_families.py:832-841``),
the family kind and its padded extent (``family_array_shape``, :844-853), the
lateral family codes, and each carrier's shape, dtype and device.
"""

from __future__ import annotations

from typing import Any

from . import _native
from ._device import NG, lateral_code, ptr, require, stream_handle
from ._native import I, P, U, signature

KIND_SCALAR, KIND_U, KIND_V, KIND_W = 0, 1, 2, 3

signature("gm1hip_fill_family", P, I, U, I, I, I, P)
signature("gm1hip_fill_pressure", P, U, I, I, I, P)
signature("gm1hip_w_rigid", P, I, I, I, P)
signature("gm1hip_bcwsfc", P, P, P, P, P, P, I, I, P)
signature("gm1hip_bc2d_open", P, I, I, P)


def validate_dimensions(ni: int, nj: int, nk: int, ng: int = NG) -> None:
    """The grid the fills accept: ni, nj >= 4, nk >= 3, and more interior planes than ghost planes."""
    for name, value, lower in (("ni", ni, 4), ("nj", nj, 4), ("nk", nk, 3), ("ng", ng, 1)):
        if isinstance(value, bool) or not isinstance(value, int) or value < lower:
            raise ValueError(f"{name} must be an integer >= {lower}")
    if ni <= ng or nj <= ng:
        raise ValueError("a periodic or rigid fill reads ngxy interior planes: ni and nj must exceed ngxy")


def family_array_shape(kind: int, *, ni: int, nj: int, nk: int, ng: int = NG) -> tuple[int, int, int]:
    """The padded (z, y, x) extent of the carrier family ``kind`` (0 scalar, 1 u, 2 v, 3 w)."""
    validate_dimensions(ni, nj, nk, ng)
    if isinstance(kind, bool) or kind not in (KIND_SCALAR, KIND_U, KIND_V, KIND_W):
        raise ValueError("kind must be 0 scalar, 1 u, 2 v or 3 w")
    nx = ni + 1 + 2 * ng if kind == KIND_U else ni + 2 * ng
    ny = nj + 1 + 2 * ng if kind == KIND_V else nj + 2 * ng
    nz = nk + 3 if kind == KIND_W else nk + 2
    return (nz, ny, nx)


def pressure_array_shape(*, ni: int, nj: int, nk: int) -> tuple[int, int, int]:
    """The pressure family's one-cell-halo extent (nk+2, nj+2, ni+2)."""
    validate_dimensions(ni, nj, nk)
    return (nk + 2, nj + 2, ni + 2)


def fill_family(array, kind: int, lateral_sides, *, ni: int, nj: int, nk: int, stream: Any | None = None) -> None:
    """bcs (kind 0), bcu (1), bcv (2) or bcw (3) ghost planes for a CM1 family."""
    shape = family_array_shape(kind, ni=ni, nj=nj, nk=nk)
    code = lateral_code(lateral_sides)
    if code == 0:
        raise ValueError("a family fill needs CM1's (wbc, ebc, sbc, nbc)")
    require(array, "array", shape)
    _native.call("gm1hip_fill_family", ptr(array), kind, code, ni, nj, nk, stream_handle(stream))


def fill_pressure(array, lateral_sides, *, ni: int, nj: int, nk: int, stream: Any | None = None) -> None:
    """bcp: one-cell halo of a pressure-family carrier (nk+2, nj+2, ni+2)."""
    shape = pressure_array_shape(ni=ni, nj=nj, nk=nk)
    code = lateral_code(lateral_sides)
    if code == 0:
        raise ValueError("a pressure fill needs CM1's (wbc, ebc, sbc, nbc)")
    require(array, "array", shape)
    _native.call("gm1hip_fill_pressure", ptr(array), code, ni, nj, nk, stream_handle(stream))


def w_rigid(w3d, *, ni: int, nj: int, nk: int, stream: Any | None = None) -> None:
    """bcw's rigid floor and lid: w = 0 at levels 1 and nk+1 over the padded plane."""
    require(w3d, "w3d", family_array_shape(KIND_W, ni=ni, nj=nj, nk=nk))
    _native.call("gm1hip_w_rigid", ptr(w3d), ni, nj, nk, stream_handle(stream))


def bcwsfc(arrays, metrics, *, ni: int, nj: int, nk: int | None = None, stream: Any | None = None) -> None:
    """w at the surface from the terrain slopes and the surface u, v (bcwsfc)."""
    if nk is None:
        nk = int(arrays["w3d"].shape[0]) - 3
    u = require(arrays["u3d"], "u3d", family_array_shape(KIND_U, ni=ni, nj=nj, nk=nk))
    v = require(arrays["v3d"], "v3d", family_array_shape(KIND_V, ni=ni, nj=nj, nk=nk))
    w = require(arrays["w3d"], "w3d", family_array_shape(KIND_W, ni=ni, nj=nj, nk=nk))
    m = {n: require(metrics[n], n, (nj, ni)) for n in ("dzdx", "dzdy", "gz")}
    if w.data_ptr() in (u.data_ptr(), v.data_ptr()):
        raise ValueError("w3d must not alias u3d or v3d")
    _native.call("gm1hip_bcwsfc", ptr(u), ptr(v), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(m["gz"]), ptr(w), ni, nj,
                 stream_handle(stream))


def bc2d_open(w3d, *, ni: int, nj: int, nk: int | None = None, stream: Any | None = None) -> None:
    """The open bc2d fill of w(k=1)'s ghost planes."""
    if nk is None:
        nk = int(w3d.shape[0]) - 3
    require(w3d, "w3d", family_array_shape(KIND_W, ni=ni, nj=nj, nk=nk))
    _native.call("gm1hip_bc2d_open", ptr(w3d), ni, nj, stream_handle(stream))
