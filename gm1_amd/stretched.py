"""The flat stretched column (CM1 stretch_z on a flat grid) on gfx950.

Host mirror of the stretched Cartesian lane's operators
(``cuda_cartesian_dynamics.py:1070-1262``): CM1's terrain_flag=.false.
statements on the column's own ``mh``/``mf`` profiles and ``rdz = 1/dz`` of
the metric reference, with the base state read at CM1's (1,1,k):

* ``launch_cuda_gm1_child_stretched_momentum5`` (cuda_gm1_child_stretched_momentum.py:430-757);
* ``launch_cuda_gm1_flat_stretched_buoyancy`` and
  ``launch_cuda_gm1_child_stretched_post_acoustic`` (cuda_gm1_child_stretched_moist.py:240-372);
* ``launch_cuda_gm1_flat_acoustic_stage`` (cuda_gm1_flat_lateral_acoustic.py:117,
  the all-open Durran-Klemp family);
* the scalar RK compose and the final-stage pdef chain on the flat texts
  (cuda_gm1_child_stretched_scalar.py:262-448, :666, :851).

The column rides on the device as one buffer ``mh (nk+2) | mf (nk+3)``.
"""

from __future__ import annotations

from collections.abc import Mapping
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import _native
from ._device import NG, lateral_code, ptr, record, require, shapes, stream_handle, terrain_static_shapes
from ._native import D, Fl, I, P, U, signature
from .errors import refuse
from .grid import VerticalLevels, rk_stage_dt_f32, sound3_stage_schedule_f32
from .psolver3 import (CM1_ALPH_F32, CM1_KDIV_F32, FLAT_COLUMN_PASSES, FUSED_PASS_FLAG, CudaPsolver3TerrainOpenLaunch,
                       TerrainOpenSoundAblations, device_profiles, terrain_open_sound_shapes)
from .scalar_moist_rk import CM1_NRKMAX, moisture_weps_f64, scalar_moist_rk_stage_timestep_f32

signature("gm1hip_momentum_advection5_flat", *([P] * 18), I, I, I, Fl, Fl, Fl, U, P)
signature("gm1hip_buoyancy_flat", P, P, P, P, I, I, I, P)
signature("gm1hip_post_acoustic_flat", *([P] * 10), I, I, I, Fl, Fl, Fl, P)
signature("gm1hip_scalar_rk_stage_flat", *([P] * 13), I, I, I, I, D, Fl, Fl, Fl, U, Fl, P)
signature("gm1hip_pdef_moisture_stage_flat", *([P] * 16), I, I, I, I, D, Fl, Fl, Fl, U, Fl, P)
signature("gm1hip_psolver3_stage_flat", *([P] * 42), Fl, P, I, I, I, I, I, Fl, Fl, Fl, Fl, Fl, Fl, U, I, P)
signature("gm1hip_psolver3_stage_flat_fused", *([P] * 42), Fl, P, P, P, P, I, I, I, I, I, Fl, Fl, Fl, Fl, Fl, Fl, U, I,
          P)


@dataclass(frozen=True)
class FlatStretchedColumn:
    """A generated flat column on the device: the levels, ``mh | mf`` and ``rdz``."""

    levels: VerticalLevels
    column: torch.Tensor
    rdz: float

    @property
    def nk(self) -> int:
        return self.levels.nk


def flat_stretched_column(levels: VerticalLevels, device: Any = "cuda") -> FlatStretchedColumn:
    if not isinstance(levels, VerticalLevels):
        raise TypeError("levels must be a VerticalLevels")
    col = torch.from_numpy(levels.flat_column()).to(device)
    return FlatStretchedColumn(levels, col, levels.rdz)


def _require_column(vertical: Any, nk: int) -> FlatStretchedColumn:
    if not isinstance(vertical, FlatStretchedColumn):
        raise TypeError("vertical must be a FlatStretchedColumn (flat_stretched_column(levels))")
    if vertical.nk != nk:
        raise ValueError(f"the vertical column has nk={vertical.nk}, the launch nk={nk}")
    require(vertical.column, "vertical.column", (2 * nk + 5,))
    return vertical


def launch_cuda_gm1_child_stretched_momentum5(
    arrays: Mapping[str, Any], metrics: Mapping[str, Any], vertical: FlatStretchedColumn, *, ni: int, nj: int,
    nk: int, nrk: int, dx: float, dy: float, divx: Any, stream: Any | None = None,
    rrw_boundary_hook: Any | None = None, post_divergence_hook: Any | None = None,
    lateral_sides: tuple[int, int, int, int] = (2, 2, 2, 2), **_: Any,
) -> dict[str, Any]:
    """Mass fluxes, divx and the advu/advv/advw increments on a flat stretched column.

    The flat lane binds no rrw hook (solve2.F:575's bcw(rrw) is a
    terrain_flag statement); a post-divergence hook is refused here.
    """
    if rrw_boundary_hook is not None or post_divergence_hook is not None:
        refuse(code="GM1-STRETCHED-MOMENTUM-HOOK", gate="stretched", option="hooks", value="callable",
               reason="the flat stretched column binds no rrw or post-divergence hook in this build",
               remedy="pass rrw_boundary_hook=None and post_divergence_hook=None")
    if nrk not in (1, 2, 3):
        raise ValueError("nrk must be an integer in 1..3")
    v = _require_column(vertical, nk)
    sh = shapes(ni, nj, nk)
    expected = {"u3d": sh["u"], "uten": sh["u"], "rru": sh["u"], "v3d": sh["v"], "vten": sh["v"], "rrv": sh["v"],
                "w3d": sh["w"], "wten": sh["w"], "rrw": sh["w"], "uh": (ni + 2 * NG,), "vh": (nj + 2 * NG,)}
    a = {n: require(arrays[n], n, s) for n, s in expected.items()}
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("rho0_3d", "rf0", "rr0", "rrf0", "c2")}
    divx = require(divx, "divx", (nk, nj, ni))
    _native.call("gm1hip_momentum_advection5_flat", ptr(a["u3d"]), ptr(a["v3d"]), ptr(a["w3d"]), ptr(a["uten"]),
                 ptr(a["vten"]), ptr(a["wten"]), ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]), ptr(a["uh"]),
                 ptr(a["vh"]), ptr(m["rho0_3d"]), ptr(m["rf0"]), ptr(m["rr0"]), ptr(m["rrf0"]), ptr(m["c2"]),
                 ptr(v.column), ptr(divx), ni, nj, nk, float(dx), float(dy), float(v.rdz),
                 lateral_code(lateral_sides), stream_handle(stream))
    return {"ready_event": record(stream), "receipt": {"family": "flat-stretched", "nrk": nrk, "rdz_f32": v.rdz}}


def launch_cuda_gm1_flat_stretched_buoyancy(arrays: Mapping[str, Any], metrics: Mapping[str, Any], *, ni: int,
                                            nj: int, nk: int, stream: Any | None = None) -> None:
    """w's buoyancy c1*dum8(k-1) + c2*dum8(k), and nothing else (solve2.F:897-906 on a flat column)."""
    sh = shapes(ni, nj, nk)
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    _native.call("gm1hip_buoyancy_flat", ptr(require(arrays["dum8"], "dum8", sh["s"])),
                 ptr(require(metrics["c1"], "c1", ms["c1"])), ptr(require(metrics["c2"], "c2", ms["c2"])),
                 ptr(require(arrays["wten"], "wten", sh["w"])), ni, nj, nk, stream_handle(stream))


def launch_cuda_gm1_child_stretched_post_acoustic(
    arrays: Mapping[str, Any], metrics: Mapping[str, Any], vertical: FlatStretchedColumn, *, ni: int, nj: int,
    nk: int, dx: float, dy: float, thterm: Any, divergence: Any, theta_divergence_term: Any, divx: Any,
    stream: Any | None = None,
) -> None:
    """Flat getdiv (div, div*thterm) then getdivx (rho0/rf0-weighted mass fluxes, divx)."""
    v = _require_column(vertical, nk)
    sh = shapes(ni, nj, nk)
    cell = (nk, nj, ni)
    a = {n: require(arrays[n], n, sh[k]) for n, k in (("rru", "u"), ("rrv", "v"), ("rrw", "w"))}
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("rho0_3d", "rf0")}
    for name, t in (("thterm", thterm), ("divergence", divergence), ("theta_divergence_term", theta_divergence_term),
                    ("divx", divx)):
        require(t, name, cell)
    _native.call("gm1hip_post_acoustic_flat", ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]), ptr(m["rho0_3d"]),
                 ptr(m["rf0"]), ptr(v.column), ptr(thterm), ptr(divergence), ptr(theta_divergence_term), ptr(divx),
                 ni, nj, nk, float(dx), float(dy), float(v.rdz), stream_handle(stream))


def launch_flat_scalar_rk_compose(
    padded: Any, mass_flux: Mapping[str, Any], metrics: Mapping[str, Any], vertical: FlatStretchedColumn,
    divx: Any, a0: Any, held: Any, divterm: Any | None, out: Any, *, ni: int, nj: int, nk: int, nrk: int,
    dt: float, dx: float, dy: float, weps: float, lateral_sides, stream: Any | None = None,
) -> None:
    """One scalar's flat-column advection (5th order, WENO5 at nrk=3) fused with the A-anchored compose."""
    v = _require_column(vertical, nk)
    sh = shapes(ni, nj, nk)
    cell = (nk, nj, ni)
    padded = require(padded, "padded", sh["s"])
    rru = require(mass_flux["rru"], "rru", sh["u"])
    rrv = require(mass_flux["rrv"], "rrv", sh["v"])
    rrw = require(mass_flux["rrw"], "rrw", sh["w"])
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("c1", "c2", "rr0")}
    for name, t in (("divx", divx), ("a0", a0), ("held", held), ("out", out)):
        require(t, name, cell)
    if divterm is not None:
        require(divterm, "divterm", cell)
    _native.call("gm1hip_scalar_rk_stage_flat", ptr(padded), ptr(rru), ptr(rrv), ptr(rrw), ptr(m["c1"]),
                 ptr(m["c2"]), ptr(m["rr0"]), ptr(v.column), ptr(divx), ptr(a0), ptr(held), ptr(divterm), ptr(out),
                 ni, nj, nk, int(nrk == CM1_NRKMAX), float(weps), float(dx), float(dy), float(v.rdz),
                 lateral_code(lateral_sides), scalar_moist_rk_stage_timestep_f32(dt, nrk), stream_handle(stream))


def launch_cuda_gm1_child_stretched_pdef_moisture(
    padded: Any, mass_flux: Mapping[str, Any], metrics: Mapping[str, Any], vertical: FlatStretchedColumn,
    divx: Any, qa: Any, held: Any, out: Any, scratch: Mapping[str, Any], *, species: int, ni: int, nj: int,
    nk: int, dt: float, dx: float, dy: float, lateral_sides, stream: Any | None = None,
) -> None:
    """Final-stage moisture row on a flat column: the pdef chain fused with the compose."""
    v = _require_column(vertical, nk)
    sh = shapes(ni, nj, nk)
    cell = (nk, nj, ni)
    padded = require(padded, "padded", sh["s"])
    rru = require(mass_flux["rru"], "rru", sh["u"])
    rrv = require(mass_flux["rrv"], "rrv", sh["v"])
    rrw = require(mass_flux["rrw"], "rrw", sh["w"])
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("c1", "c2", "rho0_3d", "rr0")}
    for name, t in (("divx", divx), ("qa", qa), ("held", held), ("out", out), ("advx", scratch["advx"]),
                    ("advy", scratch["advy"]), ("advz", scratch["advz"])):
        require(t, name, cell)
    _native.call("gm1hip_pdef_moisture_stage_flat", ptr(padded), ptr(rru), ptr(rrv), ptr(rrw), ptr(m["c1"]),
                 ptr(m["c2"]), ptr(m["rho0_3d"]), ptr(m["rr0"]), ptr(v.column), ptr(divx), ptr(qa), ptr(held),
                 ptr(out), ptr(scratch["advx"]), ptr(scratch["advy"]), ptr(scratch["advz"]), ni, nj, nk,
                 int(species != 0), moisture_weps_f64(species), float(dx), float(dy), float(v.rdz),
                 lateral_code(lateral_sides), scalar_moist_rk_stage_timestep_f32(dt, CM1_NRKMAX),
                 stream_handle(stream))


def launch_cuda_gm1_flat_acoustic_stage(
    arrays: Mapping[str, Any], metrics: Mapping[str, Any], *, ni: int, nj: int, nk: int, nrk: int, dt: float,
    nsound: int, dx: float, dy: float, vertical: FlatStretchedColumn, ablations: TerrainOpenSoundAblations | None = None,
    stream: Any | None = None, lateral_sides: tuple[int, int, int, int] = (2, 2, 2, 2), alph: float | None = None,
    kdiv: float | None = None, lateral: Any | None = None, column_pass: str = "pipelined",
    alternates: Mapping[str, Any] | None = None, **_: Any,
) -> CudaPsolver3TerrainOpenLaunch:
    """One acoustic RK stage on a flat stretched column (sound.F with terrain_flag=.false.).

    Only the all-open Durran-Klemp lateral family is built (``lateral=None``,
    the reference's sealed path); irbc=2 phase speeds and roflux are refused.
    The fused column passes need ``alternates``: workspace arrays ``u3d``,
    ``v3d`` and ``ppd`` of those arrays' shapes, distinct from them, which the
    small steps alternate with the caller's arrays (no state is kept there).
    """
    if lateral is not None:
        refuse(code="GM1-FLAT-ACOUSTIC-LATERAL", gate="stretched", option="lateral", value=type(lateral).__name__,
               reason="the gfx950 flat stage builds the all-open irbc=4 family only",
               remedy="pass lateral=None")
    if nrk not in (1, 2, 3):
        raise ValueError("nrk must select explicit CM1 RK stage 1, 2, or 3")
    v = _require_column(vertical, nk)
    abl = ablations or TerrainOpenSoundAblations()
    for name in ("open_consumption", "terrain_pressure_gradient", "transformed_divergence", "implicit_vertical_solve",
                 "between_step_pressure_fill"):
        if not getattr(abl, name):
            refuse(code="GM1-ACOUSTIC-ABLATION", gate="acoustic", option=f"ablations.{name}", value=False,
                   reason="the gfx950 stage always runs every small-step operator of sound.F",
                   remedy="keep the small-step operators enabled")
    flags = (1 if abl.radiation_history_production else 0) | (2 if abl.stage_family_fills else 0)
    if column_pass not in FLAT_COLUMN_PASSES:
        raise ValueError(f"the flat column runs the LDS-staged passes {tuple(FLAT_COLUMN_PASSES)}, not {column_pass!r}")
    flags |= FLAT_COLUMN_PASSES[column_pass]
    shp = terrain_open_sound_shapes(ni, nj, nk)
    names = ("ppi", "pp3d", "ppx", "ppd", "piadv", "ppterm", "ppten", "thv", "qk", "pk1", "pk2", "mm", "fpk",
             "elim_c", "elim_d", "ua", "u3d", "uten", "rru", "va", "v3d", "vten", "rrv", "wa", "w3d", "wten", "rrw",
             "radbcw", "radbce", "radbcs", "radbcn", "uh", "vh", "ruf", "rvf")
    a = {n: require(arrays[n], n, shp[n]) for n in names}
    scalar = arrays.get("scalar")
    if scalar is not None:
        scalar = require(scalar, "scalar", shp["scalar"])
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("rho0_3d", "rr0", "rf0", "c1", "c2")}
    prof = device_profiles(v.levels.profiles(), a["u3d"].device)
    nloop, dts = sound3_stage_schedule_f32(dt, nsound, nrk)
    alph32 = float(np.float32(CM1_ALPH_F32 if alph is None else alph))
    kdiv32 = float(np.float32(CM1_KDIV_F32 if kdiv is None else kdiv))
    common = (*(ptr(a[n]) for n in names), ptr(scalar), *(ptr(m[n]) for n in ("rho0_3d", "rr0", "rf0", "c1", "c2")),
              ptr(v.column), float(v.rdz), ptr(prof))
    tail = (ni, nj, nk, nrk, nloop, dts, rk_stage_dt_f32(dt, nrk), float(dx), float(dy), alph32, kdiv32,
            lateral_code(lateral_sides), flags, stream_handle(stream))
    if flags & FUSED_PASS_FLAG:
        if alternates is None:
            raise ValueError(f"the {column_pass!r} column pass needs alternates u3d, v3d and ppd")
        alt = {n: require(alternates[n], f"alternates[{n!r}]", shp[n]) for n in ("u3d", "v3d", "ppd")}
        for n in alt:
            if alt[n].data_ptr() == a[n].data_ptr():
                raise ValueError(f"alternates[{n!r}] must be a different array from {n}")
        _native.call("gm1hip_psolver3_stage_flat_fused", *common, ptr(alt["u3d"]), ptr(alt["v3d"]), ptr(alt["ppd"]),
                     *tail)
    else:
        _native.call("gm1hip_psolver3_stage_flat", *common, *tail)
    cadence = [{"nrk": nrk, "nloop": nloop, "dts_f32": dts}]
    return CudaPsolver3TerrainOpenLaunch(record(stream), None, tuple(cadence), {
        "domain": {"ni": ni, "nj": nj, "nk": nk, "ngxy": NG}, "cadence": cadence, "family": "flat-stretched",
        "acoustic_reals": {"alph": alph32, "kdiv": kdiv32}, "rdz_f32": v.rdz,
        "lateral_boundaries": list(lateral_sides), "target_arch": "gfx950",
    })


__all__ = [
    "FlatStretchedColumn", "flat_stretched_column", "launch_cuda_gm1_child_stretched_momentum5",
    "launch_cuda_gm1_child_stretched_pdef_moisture", "launch_cuda_gm1_child_stretched_post_acoustic",
    "launch_cuda_gm1_flat_acoustic_stage", "launch_cuda_gm1_flat_stretched_buoyancy",
    "launch_flat_scalar_rk_compose",
]
