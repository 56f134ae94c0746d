"""Scalar fifth-order / WENO5 advection and the terrain mass flux (gfx950).

Host mirror of the reference's ``cuda_advection5.py`` operator interface:
``scalar_advection_uses_weno`` (:129-147), ``launch_cuda_terrain_mass_flux5``
(:849-938) and ``launch_cuda_scalar_advection5`` (:948-1049) keep their names,
keywords, dict keys, shapes and errors; arrays are ROCm ``torch.Tensor``
carriers instead of ``cupy.ndarray``.  The work runs in
``libgm1hip.so`` (``gm1hip_terrain_mass_flux5``, ``gm1hip_scalar_advection5``).
"""

from __future__ import annotations

import math
from collections.abc import Mapping, MutableMapping
from dataclasses import dataclass
from typing import Any

from . import _native
from ._device import (
    NG, Launch, check_grid, check_spacing, lateral_code, no_alias, ptr, record,
    require, shapes, stream_handle, terrain_static_shapes,
)
from ._native import D, Fl, I, P, U, signature
from .errors import refuse

CM1_NRKMAX = 3
CM1_WENO_SIFORM = 2
SCALAR_TILE_LOCAL_OBC = 0

signature("gm1hip_terrain_mass_flux5", P, P, P, P, P, P, P, P, P, P, P, P, P, P, I, I, I, Fl, Fl, Fl, P)
signature("gm1hip_scalar_advection5", P, P, P, P, P, P, P, P, P, P, P, P, P, I, I, I, I, D, Fl, Fl, Fl, U, P)


@dataclass(frozen=True)
class Advection5Selection:
    """The one admitted branch (cuda_advection5.py:60-90)."""

    scalar_horizontal_order: int = 5
    scalar_vertical_order: int = 5
    scalar_weno: int = 2
    velocity_horizontal_order: int = 5
    velocity_vertical_order: int = 5
    velocity_weno: int = 0
    weno_order: int = 5
    weno_siform: int = CM1_WENO_SIFORM
    terrain: bool = True
    axisymmetric: int = 0
    mpi: bool = False
    positive_definite: bool = False


DEFAULT_ADVECTION5_SELECTION = Advection5Selection()
_EXPECTED = {
    "dynamics.scalar_horizontal_order": ("scalar_horizontal_order", 5),
    "dynamics.scalar_vertical_order": ("scalar_vertical_order", 5),
    "dynamics.scalar_weno": ("scalar_weno", 2),
    "dynamics.velocity_horizontal_order": ("velocity_horizontal_order", 5),
    "dynamics.velocity_vertical_order": ("velocity_vertical_order", 5),
    "dynamics.velocity_weno": ("velocity_weno", 0),
    "dynamics.weno_order": ("weno_order", 5),
    "dynamics.weno_siform": ("weno_siform", 2),
    "grid.terrain": ("terrain", True),
    "grid.axisymmetric": ("axisymmetric", 0),
    "runtime.mpi": ("mpi", False),
    "advection.positive_definite": ("positive_definite", False),
}


def validate_advection5_selection(selection: Advection5Selection) -> None:
    """Refuse any unadmitted branch before launch (cuda_advection5.py:93-126)."""
    if not isinstance(selection, Advection5Selection):
        raise TypeError("selection must be an Advection5Selection")
    for option, (attribute, expected) in _EXPECTED.items():
        value = getattr(selection, attribute)
        if type(value) is not type(expected) or value != expected:
            refuse(
                code="CM1-ADVECTION5-UNSUPPORTED",
                gate="dry-dynamics",
                option=option,
                value=value,
                reason="only the fifth-order terrain scalar branch with final-stage WENO-Z is built",
                remedy="use the fifth-order terrain scalar profile",
            )


def scalar_advection_uses_weno(
    nrk: int, *, nrkmax: int = CM1_NRKMAX, wflag: bool = True,
    selection: Advection5Selection = DEFAULT_ADVECTION5_SELECTION,
) -> bool:
    """WENO only on the final RK stage (cuda_advection5.py:129-147)."""
    validate_advection5_selection(selection)
    if isinstance(nrk, bool) or not isinstance(nrk, int):
        raise TypeError("nrk must be an integer")
    if isinstance(nrkmax, bool) or not isinstance(nrkmax, int):
        raise TypeError("nrkmax must be an integer")
    if nrkmax != CM1_NRKMAX or nrk < 1 or nrk > nrkmax:
        raise ValueError("the sealed low-storage schedule requires nrkmax=3 and nrk=1..3")
    if not isinstance(wflag, bool):
        raise TypeError("wflag must be boolean")
    return bool(wflag and selection.scalar_weno == 2 and nrk == nrkmax)


def scalar_open_boundary_code(lateral_sides: Any) -> int:
    """cuda_advection5.py:825-846."""
    return lateral_code(lateral_sides)


@dataclass(frozen=True)
class CompiledHipLibrary:
    """Stand-in for the reference's compiled-kernel handle: the loaded library."""

    path: str
    target_arch: str = "gfx950"


def compile_cuda_advection5() -> CompiledHipLibrary:
    _native.lib()
    return CompiledHipLibrary(str(_native.LIBRARY_PATH))


def launch_cuda_terrain_mass_flux5(
    arrays: MutableMapping[str, Any], metrics: Mapping[str, Any], divx: Any, *,
    ni: int, nj: int, nk: int, dx: float, dy: float, dz: float,
    selection: Advection5Selection = DEFAULT_ADVECTION5_SELECTION,
    stream: Any | None = None, compiled: Any | None = None,
) -> Launch:
    """Terrain mass flux and interior divergence (cuda_advection5.py:849-938)."""
    validate_advection5_selection(selection)
    ni, nj, nk = check_grid(ni, nj, nk)
    dx, dy, dz = (check_spacing(n, v) for n, v in (("dx", dx), ("dy", dy), ("dz", dz)))
    sh = shapes(ni, nj, nk)
    expected = {"u3d": sh["u"], "v3d": sh["v"], "w3d": sh["w"], "rru": sh["u"], "rrv": sh["v"], "rrw": sh["w"]}
    a = {n: require(arrays[n], n, s) for n, s in expected.items()}
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    names = ("rho0_3d", "rf0", "rgzu", "rgzv", "dzdx", "dzdy", "gz")
    m = {n: require(metrics[n], n, ms[n]) for n in names}
    divx = require(divx, "divx", (nk, nj, ni))
    no_alias({**a, **m, "divx": divx}, "terrain mass-flux inputs and outputs must not alias")
    compiled = compiled or compile_cuda_advection5()
    _native.call(
        "gm1hip_terrain_mass_flux5",
        ptr(a["u3d"]), ptr(a["v3d"]), ptr(a["w3d"]), ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]),
        ptr(m["rho0_3d"]), ptr(m["rf0"]), ptr(m["rgzu"]), ptr(m["rgzv"]), ptr(m["dzdx"]),
        ptr(m["dzdy"]), ptr(m["gz"]), ptr(divx), ni, nj, nk, dx, dy, dz, stream_handle(stream),
    )
    return Launch(record(stream), compiled, {
        "operation": "terrain-mass-flux-and-divergence",
        "domain": {"ni": ni, "nj": nj, "nk": nk},
        "spacing_m": {"dx": dx, "dy": dy, "dz": dz},
        "source_anchors": {"terrain_mass_flux": "src/solve2.F:490-571", "terrain_divergence": "src/solve2.F:613-627"},
        "target_arch": "gfx950",
    })


def launch_cuda_scalar_advection5(
    scalar: Any, arrays: Mapping[str, Any], metrics: Mapping[str, Any], divx: Any,
    tendency: Any, fluxes: Mapping[str, Any] | None, *,
    ni: int, nj: int, nk: int, nrk: int, dx: float, dy: float, dz: float, weps: float,
    wflag: bool = True, selection: Advection5Selection = DEFAULT_ADVECTION5_SELECTION,
    stream: Any | None = None, compiled: Any | None = None,
    lateral_sides: tuple[int, int, int, int] | None = None,
) -> Launch:
    """Scalar face fluxes and tendency of one RK stage (cuda_advection5.py:948-1049).

    ``fluxes`` may be ``None``: the fused kernel then keeps face fluxes in LDS
    and registers only.  With a mapping, its ``x``/``y``/``z`` arrays receive
    the face fluxes as the reference's scratch does.
    """
    use_weno = scalar_advection_uses_weno(nrk, wflag=wflag, selection=selection)
    ni, nj, nk = check_grid(ni, nj, nk)
    dx, dy, dz = (check_spacing(n, v) for n, v in (("dx", dx), ("dy", dy), ("dz", dz)))
    if isinstance(weps, bool) or not isinstance(weps, (int, float)):
        raise TypeError("weps must be numeric")
    weps = float(weps)
    if not math.isfinite(weps) or weps <= 0.0:
        raise ValueError("weps must be finite and > 0")
    sh = shapes(ni, nj, nk)
    scalar = require(scalar, "scalar", sh["s"])
    rru = require(arrays["rru"], "rru", sh["u"])
    rrv = require(arrays["rrv"], "rrv", sh["v"])
    rrw = require(arrays["rrw"], "rrw", sh["w"])
    fx = fy = fz = None
    if fluxes is not None:
        fx = require(fluxes["x"], "flux_x", sh["u"])
        fy = require(fluxes["y"], "flux_y", sh["v"])
        fz = require(fluxes["z"], "flux_z", sh["w"])
    divx = require(divx, "divx", (nk, nj, ni))
    tendency = require(tendency, "tendency", (nk, nj, ni))
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    c1 = require(metrics["c1"], "c1", ms["c1"])
    c2 = require(metrics["c2"], "c2", ms["c2"])
    rr0 = require(metrics["rr0"], "rr0", ms["rr0"])
    gz = require(metrics["gz"], "gz", ms["gz"])
    no_alias(
        {"scalar": scalar, "rru": rru, "rrv": rrv, "rrw": rrw, "fx": fx, "fy": fy, "fz": fz,
         "divx": divx, "tendency": tendency},
        "scalar advection inputs, flux scratch, and tendency must not alias",
    )
    no_alias({"tendency": tendency, "c1": c1, "c2": c2, "rr0": rr0, "gz": gz},
             "scalar advection tendency must not alias a metric")
    obc = scalar_open_boundary_code(lateral_sides)
    compiled = compiled or compile_cuda_advection5()
    _native.call(
        "gm1hip_scalar_advection5",
        ptr(scalar), ptr(rru), ptr(rrv), ptr(rrw), ptr(c1), ptr(c2), ptr(rr0), ptr(gz), ptr(divx),
        ptr(tendency), ptr(fx), ptr(fy), ptr(fz), ni, nj, nk, int(use_weno), weps, dx, dy, dz,
        obc, stream_handle(stream),
    )
    return Launch(record(stream), compiled, {
        "domain": {"ni": ni, "nj": nj, "nk": nk},
        "spacing_m": {"dx": dx, "dy": dy, "dz": dz},
        "stage": {"nrk": nrk, "nrkmax": CM1_NRKMAX, "wflag": wflag, "uses_weno": use_weno},
        "weno": {"order": 5, "siform": CM1_WENO_SIFORM, "weps_f64": weps},
        "lateral_sides": None if lateral_sides is None else list(lateral_sides),
        "tendency_cells": "tile-local interior 3<=i<ni-3, 3<=j<nj-3" if obc == 0 else "every cell with open-side inflow edits",
        "target_arch": "gfx950",
    })


__all__ = [
    "Advection5Selection", "CM1_NRKMAX", "CM1_WENO_SIFORM", "DEFAULT_ADVECTION5_SELECTION",
    "SCALAR_TILE_LOCAL_OBC", "compile_cuda_advection5", "launch_cuda_scalar_advection5",
    "launch_cuda_terrain_mass_flux5", "scalar_advection_uses_weno", "scalar_open_boundary_code",
    "validate_advection5_selection", "NG",
]
