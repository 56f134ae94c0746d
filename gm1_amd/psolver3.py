"""psolver=3 acoustic RK stage (CM1 ``sound``) on gfx950.

Host mirror of ``cuda_psolver3_terrain_open.py``: ``terrain_open_sound_shapes``
(:162-187), ``TerrainOpenSoundAblations`` (:275-284),
``launch_cuda_psolver3_terrain_open_three_stage`` (:1217-1433) and
``launch_cuda_psolver3_terrain_open_stage`` (:1436-1507).  The whole stage
(prepare, every small step, radiation history, stage fills) is one C-ABI call,
``gm1hip_psolver3_stage``, that enqueues its kernels on the caller's stream.

``lateral_sides`` is CM1's ``(wbc, ebc, sbc, nbc)``; the reference reaches the
same family through a bound ``lateral`` object.  Open sides take the
``ssopenbc`` step, periodic sides take the boundary-face pressure gradient,
and every fill is the family fill of that side.
"""

from __future__ import annotations

from collections.abc import Mapping
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import _native
from ._device import NG, check_spacing, lateral_code, no_alias, ptr, record, require, stream_handle, terrain_static_shapes
from ._native import Fl, I, P, U, signature
from .grid import VerticalProfiles, rk_stage_dt_f32, sound3_stage_schedule_f32, terrain_coordinate_profiles
from .errors import refuse

CM1_CP_F32 = 1005.7
CM1_ALPH_F32 = 0.60
CM1_KDIV_F32 = 0.1

PRESSURE_NAMES = ("ppi", "pp3d", "ppx", "ppd", "piadv", "ppterm", "ppten", "thv", "qk", "pk1", "pk2", "mm",
                  "fpk", "elim_c", "elim_d")
U_NAMES = ("ua", "u3d", "uten", "rru")
V_NAMES = ("va", "v3d", "vten", "rrv")
W_NAMES = ("wa", "w3d", "wten", "rrw")
BOUNDARY_NAMES = ("radbcw", "radbce", "radbcs", "radbcn", "uh", "vh", "ruf", "rvf")
METRIC_NAMES = ("rho0_3d", "rr0", "rf0", "rrf0", "c1", "c2", "gz", "rgzu", "rgzv", "dzdx", "dzdy",
                "gx", "gy", "gxu", "gyv")

signature("gm1hip_psolver3_stage", *([P] * 52), Fl, I, I, I, I, I, Fl, Fl, Fl, Fl, Fl, Fl, U, I, P)


def terrain_open_sound_shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    ng = NG
    pressure = (nk + 2, nj + 2, ni + 2)
    shapes = {name: pressure for name in PRESSURE_NAMES}
    shapes.update({name: (nk + 2, nj + 2 * ng, ni + 1 + 2 * ng) for name in U_NAMES})
    shapes.update({name: (nk + 2, nj + 1 + 2 * ng, ni + 2 * ng) for name in V_NAMES})
    shapes.update({name: (nk + 3, nj + 2 * ng, ni + 2 * ng) for name in W_NAMES})
    shapes.update({
        "radbcw": (nk + 2, nj + 2 * ng), "radbce": (nk + 2, nj + 2 * ng),
        "radbcs": (nk + 2, ni + 2 * ng), "radbcn": (nk + 2, ni + 2 * ng),
        "uh": (ni + 2 * ng,), "vh": (nj + 2 * ng,), "ruf": (ni + 1 + 2 * ng,), "rvf": (nj + 1 + 2 * ng,),
        "scalar": (nk + 2, nj + 2 * ng, ni + 2 * ng),
    })
    return shapes


@dataclass(frozen=True)
class TerrainOpenSoundAblations:
    """Which parts of the stage run (cuda_psolver3_terrain_open.py:275-284).

    Only the two stage-end parts may be switched off; every small-step part is
    the physics itself and is refused when disabled.
    """

    open_consumption: bool = True
    terrain_pressure_gradient: bool = True
    transformed_divergence: bool = True
    pressure_advection: bool = True
    implicit_vertical_solve: bool = True
    between_step_pressure_fill: bool = True
    radiation_history_production: bool = True
    stage_family_fills: bool = True


_profile_cache: dict[tuple, torch.Tensor] = {}


def device_profiles(profiles: VerticalProfiles, device: torch.device) -> torch.Tensor:
    key = (profiles.nk, profiles.packed().tobytes(), str(device))
    t = _profile_cache.get(key)
    if t is None:
        t = torch.from_numpy(profiles.packed()).to(device)
        _profile_cache[key] = t
    return t


@dataclass(frozen=True)
class CudaPsolver3TerrainOpenLaunch:
    ready_event: Any
    compiled: Any
    cadence: tuple[dict[str, Any], ...]
    receipt: dict[str, Any]


# (data_ptr, version) of gxu and gyv tensors whose every word has been checked to be +0
_slopes_checked: set[tuple] = set()


def _slopes_key(metrics: Mapping[str, Any]) -> tuple:
    return tuple((int(metrics[n].data_ptr()), int(metrics[n]._version)) for n in ("gxu", "gyv"))


def verify_terrain_slopes_zero(metrics: Mapping[str, Any]) -> bool:
    """Check once on the device that gxu and gyv are +0 in every word (reads one flag back).

    The result is remembered for these tensors until they are written again, so the check
    runs outside any step (a step under graph capture may not read back to the host).
    """
    key = _slopes_key(metrics)
    if key in _slopes_checked:
        return True
    bad = False
    for name in ("gxu", "gyv"):
        g = metrics[name]
        if g.numel():
            bad = bad or bool(torch.any((g != 0) | torch.signbit(g)).item())
    if not bad:
        _slopes_checked.add(key)
    return not bad


# column pass name -> the stage call's flag bits: "staged" is 32 columns x 512
# threads per workgroup, "staged16" 16 x 256; the "pipelined" forms issue each
# level's loads a level ahead and run the u/v time averages beside the
# recurrence; "register" walks each column in one thread
COLUMN_PASSES = {"staged": 0, "pipelined": 16, "staged16": 8, "pipelined16": 24, "register": 4}
# the flat column runs the LDS-staged passes (the register pass is the terrain text); its
# "fused" forms also step u and v inside the pipelined pass (one pass per small step instead
# of two), alternating the winds and ppd with a workspace of the same shapes; its "resident"
# forms keep pp and fpk of a thread's own levels in registers (four LDS planes instead of
# seven: three workgroups a CU where the pipelined pass fits two; nk <= 128)
FUSED_PASS_FLAG = 64
RESIDENT_PASS_FLAG = 128
FLAT_COLUMN_PASSES = {**{k: v for k, v in COLUMN_PASSES.items() if k != "register"},
                      "fused": 16 | FUSED_PASS_FLAG, "fused16": 24 | FUSED_PASS_FLAG,
                      "resident": RESIDENT_PASS_FLAG, "resident16": 8 | RESIDENT_PASS_FLAG}


def _check_ablations(ablations: TerrainOpenSoundAblations) -> int:
    if not isinstance(ablations, TerrainOpenSoundAblations):
        raise TypeError("ablations must be TerrainOpenSoundAblations")
    for name in ("open_consumption", "terrain_pressure_gradient", "transformed_divergence", "pressure_advection",
                 "implicit_vertical_solve", "between_step_pressure_fill"):
        if not getattr(ablations, name):
            refuse(code="GM1-ACOUSTIC-ABLATION", gate="acoustic", option=f"ablations.{name}", value=False,
                   reason="the gfx950 stage always runs every small-step operator of sound.F",
                   remedy="keep the small-step operators enabled")
    return (1 if ablations.radiation_history_production else 0) | (2 if ablations.stage_family_fills else 0)


def launch_cuda_psolver3_terrain_open_three_stage(
    arrays: Mapping[str, Any], metrics: Mapping[str, Any], *, ni: int, nj: int, nk: int, dt: float,
    nsound: int, dx: float, dy: float, dz: float, selection: Any | None = None,
    ablations: TerrainOpenSoundAblations | None = None, stream: Any | None = None, compiled: Any | None = None,
    lateral_sides: tuple[int, int, int, int] = (2, 2, 2, 2), alph: float | None = None,
    kdiv: float | None = None, profiles: VerticalProfiles | None = None, _nrk_sequence: tuple[int, ...] = (1, 2, 3),
    column_pass: str = "pipelined", terrain_slopes_zero: bool = False, **_: Any,
) -> CudaPsolver3TerrainOpenLaunch:
    """Enqueue the acoustic RK stages of ``_nrk_sequence`` on ``stream``.

    ``column_pass`` selects the vertical pass: "staged" (LDS-staged, k-parallel
    outside the Thomas recurrence, 32 columns x 512 threads per workgroup),
    "staged16" (16 x 256), "pipelined"/"pipelined16" (the same with loads
    issued a level ahead) or "register" (one thread walks each column); all
    produce the same words.  ``terrain_slopes_zero`` states that the metrics'
    gxu and gyv are +0 everywhere (checked once by the caller, e.g.
    ``large_step.slopes_are_zero``); the stage then gives the same words with
    less work.
    """
    if column_pass not in COLUMN_PASSES:
        raise ValueError(f"column_pass must be one of {COLUMN_PASSES}")
    for name, value in (("alph", alph), ("kdiv", kdiv)):
        if value is not None and not (0.0 <= float(value) <= 1.0):
            raise ValueError(f"{name}={value} is outside CM1's admitted [0, 1]")
    dt, dx, dy, dz = (check_spacing(n, v) for n, v in (("dt", dt), ("dx", dx), ("dy", dy), ("dz", dz)))
    flags = _check_ablations(ablations or TerrainOpenSoundAblations()) | COLUMN_PASSES[column_pass]
    sides = lateral_code(lateral_sides)
    shapes = terrain_open_sound_shapes(ni, nj, nk)
    a = {}
    for name in (*PRESSURE_NAMES, *U_NAMES, *V_NAMES, *W_NAMES, *BOUNDARY_NAMES):
        a[name] = require(arrays[name], name, shapes[name])
    scalar = arrays.get("scalar")
    if scalar is not None:
        scalar = require(scalar, "scalar", shapes["scalar"])
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {name: require(metrics[name], name, ms[name]) for name in METRIC_NAMES}
    if terrain_slopes_zero:
        # gxu and gyv are +0 everywhere (flat terrain): the pressure gradient then forms its terrain
        # term only where it can change a word.  The claim is checked on the device (once per
        # tensor version; a check inside a graph capture is refused rather than trusted).
        if _slopes_key(m) not in _slopes_checked and torch.cuda.is_current_stream_capturing():
            raise ValueError("terrain_slopes_zero=True on metrics not yet checked; call "
                             "verify_terrain_slopes_zero(metrics) before capturing")
        if not verify_terrain_slopes_zero(m):
            raise ValueError("terrain_slopes_zero=True but gxu/gyv hold a word other than +0")
        flags |= 32
    no_alias({**a, "scalar": scalar}, "terrain/open sound arrays must not alias")
    prof = profiles or terrain_coordinate_profiles(nk, dz)
    dprof = device_profiles(prof, a["u3d"].device)
    alph32 = float(np.float32(CM1_ALPH_F32 if alph is None else alph))
    kdiv32 = float(np.float32(CM1_KDIV_F32 if kdiv is None else kdiv))
    h = stream_handle(stream)
    cadence = []
    for nrk in _nrk_sequence:
        nloop, dts = sound3_stage_schedule_f32(dt, nsound, nrk)
        _native.call(
            "gm1hip_psolver3_stage",
            *(ptr(a[n]) for n in ("ppi", "pp3d", "ppx", "ppd", "piadv", "ppterm", "ppten", "thv", "qk", "pk1",
                                  "pk2", "mm", "fpk", "elim_c", "elim_d", "ua", "u3d", "uten", "rru", "va", "v3d",
                                  "vten", "rrv", "wa", "w3d", "wten", "rrw", "radbcw", "radbce", "radbcs",
                                  "radbcn", "uh", "vh", "ruf", "rvf")),
            ptr(scalar),
            *(ptr(m[n]) for n in METRIC_NAMES),
            ptr(dprof), prof.ztop,
            ni, nj, nk, nrk, nloop, dts, rk_stage_dt_f32(dt, nrk), dx, dy, alph32, kdiv32, sides, flags, h,
        )
        cadence.append({"nrk": nrk, "nloop": nloop, "dts_f32": dts})
    return CudaPsolver3TerrainOpenLaunch(record(stream), compiled, tuple(cadence), {
        "domain": {"ni": ni, "nj": nj, "nk": nk, "ngxy": NG},
        "cadence": cadence,
        "acoustic_reals": {"alph": alph32, "kdiv": kdiv32},
        "lateral_boundaries": list(lateral_sides),
        "target_arch": "gfx950",
    })


def launch_cuda_psolver3_terrain_open_stage(
    arrays: Mapping[str, Any], metrics: Mapping[str, Any], *, ni: int, nj: int, nk: int, nrk: int, dt: float,
    nsound: int, dx: float, dy: float, dz: float, **kwargs: Any,
) -> CudaPsolver3TerrainOpenLaunch:
    """Exactly one acoustic RK stage (cuda_psolver3_terrain_open.py:1436-1507)."""
    if isinstance(nrk, bool) or not isinstance(nrk, int) or nrk not in (1, 2, 3):
        raise ValueError("nrk must select explicit CM1 RK stage 1, 2, or 3")
    return launch_cuda_psolver3_terrain_open_three_stage(
        arrays, metrics, ni=ni, nj=nj, nk=nk, dt=dt, nsound=nsound, dx=dx, dy=dy, dz=dz,
        _nrk_sequence=(nrk,), **kwargs)


__all__ = [
    "CM1_ALPH_F32", "CM1_CP_F32", "CM1_KDIV_F32", "CudaPsolver3TerrainOpenLaunch", "TerrainOpenSoundAblations",
    "launch_cuda_psolver3_terrain_open_stage", "launch_cuda_psolver3_terrain_open_three_stage",
    "terrain_open_sound_shapes", "verify_terrain_slopes_zero",
]
