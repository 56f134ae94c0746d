"""Fifth-order terrain/open momentum advection (gfx950).

Host mirror of the reference's ``cuda_momentum_advection5.py`` interface:
``momentum_lateral_sides`` (:103-122), ``momentum_open_boundary_code``
(:125-137), ``momentum_advection5_counts`` (:140-152) and
``launch_cuda_terrain_open_momentum_advection5`` (:1173-1330), with the same
keywords, hooks, dict keys and errors.  Kernels: ``gm1hip_momentum_advection5``
and ``gm1hip_momentum_transport5`` in ``libgm1hip.so``.
"""

from __future__ import annotations

from collections.abc import Callable, Mapping
from dataclasses import dataclass
from typing import Any

import torch

from . import _native
from ._device import (
    NG, check_grid, check_spacing, no_alias, ptr, record, require, shapes, stream_handle,
    terrain_static_shapes,
)
from ._native import Fl, I, P, U, signature

MOMENTUM_ALL_OPEN_SIDES = (2, 2, 2, 2)
CM1_SMEPS_F32 = 1.0e-20

signature("gm1hip_momentum_advection5", *([P] * 23), I, I, I, Fl, Fl, Fl, U, I, P)
signature("gm1hip_momentum_transport5", *([P] * 18), I, I, I, Fl, Fl, Fl, U, P)


@dataclass(frozen=True)
class MomentumAdvection5Selection:
    """The admitted branch (cuda_momentum_advection5.py:81-92)."""

    hadvordrv: int = 5
    vadvordrv: int = 5
    advwenov: int = 0
    axisymmetric: int = 0
    terrain: bool = True
    idiff: int = 0
    immersed_boundary: bool = False
    uniform_horizontal_grid: bool = True
    uniform_vertical_sigma: bool = True
    lateral_boundaries: tuple[int, int, int, int] = (2, 2, 2, 2)


DEFAULT_MOMENTUM_ADVECTION5_SELECTION = MomentumAdvection5Selection()


def validate_momentum_advection5_selection(selection: MomentumAdvection5Selection) -> None:
    if not isinstance(selection, MomentumAdvection5Selection):
        raise TypeError("selection must be MomentumAdvection5Selection")
    expected = DEFAULT_MOMENTUM_ADVECTION5_SELECTION
    for name in expected.__dataclass_fields__:
        if getattr(selection, name) != getattr(expected, name):
            raise ValueError(
                f"unsupported momentum-advection selection {name}={getattr(selection, name)!r}; "
                f"required {getattr(expected, name)!r}")


def momentum_lateral_sides(lateral_sides: Any) -> tuple[int, int, int, int]:
    """CM1's (wbc, ebc, sbc, nbc), periodic on both sides of a direction or neither."""
    sides = tuple(lateral_sides)
    if len(sides) != 4 or any(
        isinstance(v, bool) or not isinstance(v, int) or v not in (1, 2, 3, 4) for v in sides
    ):
        raise ValueError("lateral_sides must be CM1's (wbc, ebc, sbc, nbc), each 1, 2, 3 or 4")
    if (sides[0] == 1) != (sides[1] == 1) or (sides[2] == 1) != (sides[3] == 1):
        raise ValueError(f"lateral_sides {sides}: periodic b.c.'s must be on both sides of a direction")
    return sides  # type: ignore[return-value]


def momentum_open_boundary_code(lateral_sides: Any) -> int:
    w, e, s, n = momentum_lateral_sides(lateral_sides)
    return w | (e << 4) | (s << 8) | (n << 12)


def momentum_advection5_counts(ni: int, nj: int, nk: int, lateral_sides: Any = MOMENTUM_ALL_OPEN_SIDES):
    """Words advu, advv and advw visit (faces i1..i2, rows j1..j2, every w cell)."""
    w, e, s, n = momentum_lateral_sides(lateral_sides)
    faces = ni - 1 + int(w == 1) + int(e == 1)
    rows = nj - 1 + int(s == 1) + int(n == 1)
    return nk * nj * faces, nk * rows * ni, (nk - 1) * nj * ni


@dataclass(frozen=True)
class CudaMomentumAdvection5Launch:
    divx: Any
    ready_event: Any
    compiled: Any
    receipt: dict[str, Any]


def launch_cuda_terrain_open_momentum_advection5(
    arrays: Mapping[str, Any], metrics: Mapping[str, Any], *,
    ni: int, nj: int, nk: int, nrk: int, dx: float, dy: float, dz: float,
    divx: Any | None = None,
    selection: MomentumAdvection5Selection = DEFAULT_MOMENTUM_ADVECTION5_SELECTION,
    stream: Any | None = None, compiled: Any | None = None,
    rrw_boundary_hook: Any | None = None,
    post_divergence_hook: Callable[..., Mapping[str, object]] | None = None,
    lateral_sides: tuple[int, int, int, int] = MOMENTUM_ALL_OPEN_SIDES,
) -> CudaMomentumAdvection5Launch:
    """Mass fluxes, divx and advu/advv/advw increments of one RK stage.

    The hooks run at the reference's seams: ``rrw_boundary_hook(rrw, ...)``
    after the vertical mass flux, ``post_divergence_hook(arrays, ...)`` after
    ``divx`` and before the three advection increments.
    """
    validate_momentum_advection5_selection(selection)
    sides = momentum_lateral_sides(lateral_sides)
    obc = momentum_open_boundary_code(sides)
    if rrw_boundary_hook is not None and not callable(rrw_boundary_hook):
        raise TypeError("rrw_boundary_hook must be callable or None")
    if post_divergence_hook is not None and not callable(post_divergence_hook):
        raise TypeError("post_divergence_hook must be callable or None")
    ni, nj, nk = check_grid(ni, nj, nk, 7)
    if nk < 8:
        raise ValueError("nk must be an integer >= 8")
    if isinstance(nrk, bool) or not isinstance(nrk, int) or nrk not in (1, 2, 3):
        raise ValueError("nrk must be an integer in 1..3")
    dx, dy, dz = (check_spacing(n, v) for n, v in (("dx", dx), ("dy", dy), ("dz", dz)))
    sh = shapes(ni, nj, nk)
    expected = {
        "u3d": sh["u"], "uten": sh["u"], "rru": sh["u"], "v3d": sh["v"], "vten": sh["v"], "rrv": sh["v"],
        "w3d": sh["w"], "wten": sh["w"], "rrw": sh["w"], "uh": (ni + 2 * NG,), "vh": (nj + 2 * NG,),
    }
    a = {n: require(arrays[n], n, s) for n, s in expected.items()}
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    names = ("rho0_3d", "rf0", "rr0", "rrf0", "c1", "c2", "gz", "gzu", "gzv", "rgzu", "rgzv", "dzdx", "dzdy")
    m = {n: require(metrics[n], n, ms[n]) for n in names}
    if divx is None:
        divx = torch.empty((nk, nj, ni), dtype=torch.float32, device=a["u3d"].device)
    else:
        divx = require(divx, "divx", (nk, nj, ni))
    no_alias({**{f"state.{k}": v for k, v in a.items()}, "divx": divx},
             "momentum-advection mutable arrays and unproved metrics must not alias")
    no_alias({**{f"out.{k}": a[k] for k in ("uten", "vten", "wten", "rru", "rrv", "rrw")}, "divx": divx,
              **{f"metric.{k}": v for k, v in m.items()}},
             "momentum-advection mutable arrays and unproved metrics must not alias")
    h = stream_handle(stream)
    split = rrw_boundary_hook is not None or post_divergence_hook is not None
    _native.call(
        "gm1hip_momentum_advection5",
        ptr(a["u3d"]), ptr(a["v3d"]), ptr(a["w3d"]), ptr(a["uten"]), ptr(a["vten"]), ptr(a["wten"]),
        ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]), ptr(a["uh"]), ptr(a["vh"]), ptr(m["rho0_3d"]),
        ptr(m["rf0"]), ptr(m["rrf0"]), ptr(m["c2"]), ptr(m["gz"]), ptr(m["gzu"]), ptr(m["gzv"]),
        ptr(m["rgzu"]), ptr(m["rgzv"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(divx),
        ni, nj, nk, dx, dy, dz, obc, int(split), h,
    )
    hook_receipts: dict[str, Any] = {}
    if split:
        if rrw_boundary_hook is not None:
            hook_receipts["rrw_boundary"] = rrw_boundary_hook(a["rrw"], stream=stream, ni=ni, nj=nj, nk=nk)
        if post_divergence_hook is not None:
            r = post_divergence_hook(a, stream=stream, ni=ni, nj=nj, nk=nk, nrk=nrk, ng=NG)
            if not isinstance(r, Mapping):
                raise TypeError("post_divergence_hook must return a receipt mapping")
            hook_receipts["post_divergence"] = dict(r)
        _native.call(
            "gm1hip_momentum_transport5",
            ptr(a["u3d"]), ptr(a["v3d"]), ptr(a["w3d"]), ptr(a["uten"]), ptr(a["vten"]), ptr(a["wten"]),
            ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]), ptr(a["uh"]), ptr(a["vh"]), ptr(m["rho0_3d"]),
            ptr(m["rrf0"]), ptr(m["c2"]), ptr(m["gz"]), ptr(m["gzu"]), ptr(m["gzv"]), ptr(divx),
            ni, nj, nk, dx, dy, dz, obc, h,
        )
    return CudaMomentumAdvection5Launch(divx, record(stream), compiled, {
        "selection": {"hadvordrv": 5, "vadvordrv": 5, "advwenov": 0, "lateral_boundaries": list(sides), "nrk": nrk},
        "domain": {"ni": ni, "nj": nj, "nk": nk},
        "hooks": hook_receipts,
        "target_arch": "gfx950",
    })


__all__ = [
    "CudaMomentumAdvection5Launch", "DEFAULT_MOMENTUM_ADVECTION5_SELECTION", "MOMENTUM_ALL_OPEN_SIDES",
    "MomentumAdvection5Selection", "launch_cuda_terrain_open_momentum_advection5",
    "momentum_advection5_counts", "momentum_lateral_sides", "momentum_open_boundary_code",
]
