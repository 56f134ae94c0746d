"""The momentum RK stage's stateful entry points on gfx950 (host mirror of cuda_momentum_rk.py).

``create_cuda_momentum_rk_state`` (cuda_momentum_rk.py:587-654) binds the
acoustic arrays, the static metrics and the held tendencies uten1/vten1/wten1;
``launch_cuda_momentum_rk_stage`` (:683-850) runs one stage in CM1's source
order (``cpu_momentum_rk_stage_inplace``, :303-372):

1. the held tendencies seed uten/vten/wten (``gm1hip_seed_held_momentum``);
2. from nrk 2 on, the terrain surface w (bcwsfc) and its W, E, S, N bc2d copy;
3. 5th-order momentum advection with the mass fluxes and divx;
4. the psolver=3 acoustic stage;

and ``begin_cuda_momentum_rk_step`` (:657) resets the stage cadence once the
caller has refreshed the A-level and held carriers.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Mapping

import torch

from . import _native
from ._device import no_alias, ptr, record, require, stream_handle, terrain_static_shapes
from ._native import I, P, signature
from .momentum_advection5 import launch_cuda_terrain_open_momentum_advection5
from .psolver3 import TerrainOpenSoundAblations, launch_cuda_psolver3_terrain_open_stage, terrain_open_sound_shapes

signature("gm1hip_seed_held_momentum", P, P, P, P, P, P, I, I, I, P)
signature("gm1hip_bcwsfc", P, P, P, P, P, P, I, I, P)
signature("gm1hip_bc2d_open", P, I, I, P)

HELD_NAMES = ("uten1", "vten1", "wten1")


@dataclass
class CudaMomentumRKState:
    ni: int
    nj: int
    nk: int
    arrays: dict[str, torch.Tensor]
    metrics: dict[str, torch.Tensor]
    held: dict[str, torch.Tensor]
    divx: torch.Tensor
    stream: Any
    device_id: int
    next_nrk: int = 1
    completed: bool = False
    lateral_sides: tuple[int, int, int, int] = (2, 2, 2, 2)


@dataclass(frozen=True)
class CudaMomentumRKStageLaunch:
    state: CudaMomentumRKState
    ready_event: Any
    advection_launch: Any
    acoustic_launch: Any
    receipt: dict[str, Any] = field(default_factory=dict)


def _positive_int(name: str, value: Any, minimum: int) -> int:
    if isinstance(value, bool) or not isinstance(value, int) or value < minimum:
        raise ValueError(f"{name} must be an integer >= {minimum}")
    return value


def _positive(name: str, value: Any) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise TypeError(f"{name} must be numeric")
    v = float(value)
    if not math.isfinite(v) or v <= 0.0:
        raise ValueError(f"{name} must be finite and > 0")
    return v


def create_cuda_momentum_rk_state(arrays: Mapping[str, Any], metrics: Mapping[str, Any], held: Mapping[str, Any], *,
                                  ni: int, nj: int, nk: int, stream: Any | None = None,
                                  lateral_sides: Any = (2, 2, 2, 2)) -> CudaMomentumRKState:
    """Bind the acoustic arrays, static metrics and held tendencies (cuda_momentum_rk.py:587-654)."""
    ni, nj, nk = _positive_int("ni", ni, 7), _positive_int("nj", nj, 7), _positive_int("nk", nk, 8)
    array_shapes = terrain_open_sound_shapes(ni, nj, nk)
    missing = [n for n in array_shapes if n not in arrays]
    if missing:
        raise ValueError(f"missing momentum-RK arrays: {', '.join(missing)}")
    a = {n: require(arrays[n], n, s) for n, s in array_shapes.items()}
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    missing = [n for n in ms if n not in metrics]
    if missing:
        raise ValueError(f"missing momentum-RK metrics: {', '.join(missing)}")
    m = {n: require(metrics[n], n, s) for n, s in ms.items()}
    h = {n: require(held[n], n, tuple(a[n.replace("1", "")].shape)) for n in HELD_NAMES}
    dev = a["u3d"].device
    divx = torch.empty((nk, nj, ni), dtype=torch.float32, device=dev)
    no_alias({**{f"array.{k}": v for k, v in a.items()}, **{f"held.{k}": v for k, v in h.items()}, "divx": divx},
             "momentum-RK arrays, held carriers and scratch must not alias")
    return CudaMomentumRKState(ni, nj, nk, a, m, h, divx, stream,
                               dev.index if dev.index is not None else torch.cuda.current_device(),
                               lateral_sides=tuple(int(v) for v in lateral_sides))


def begin_cuda_momentum_rk_step(state: CudaMomentumRKState) -> None:
    """Reset the cadence after the caller refreshes the A and held carriers (cuda_momentum_rk.py:657-666)."""
    if not isinstance(state, CudaMomentumRKState):
        raise TypeError("state must be a CudaMomentumRKState")
    if not state.completed or state.next_nrk != 4:
        raise ValueError("cannot begin a new large step before stage 3 completes")
    state.next_nrk = 1
    state.completed = False


def launch_cuda_momentum_rk_stage(
    state: CudaMomentumRKState, *, nrk: int, dt: float, nsound: int, dx: float, dy: float, dz: float,
    acoustic_ablations: TerrainOpenSoundAblations | None = None, stream: Any | None = None, **_: Any,
) -> CudaMomentumRKStageLaunch:
    """Seed, surface w, momentum advection and the acoustic stage of one RK stage (cuda_momentum_rk.py:683-850)."""
    if not isinstance(state, CudaMomentumRKState):
        raise TypeError("state must be a CudaMomentumRKState")
    if isinstance(nrk, bool) or not isinstance(nrk, int) or nrk not in (1, 2, 3):
        raise ValueError("nrk must be 1, 2 or 3")
    if nrk != state.next_nrk or state.completed:
        raise ValueError(f"momentum RK cadence requires stage {state.next_nrk}, received {nrk}")
    dt, dx, dy, dz = (_positive(n, v) for n, v in (("dt", dt), ("dx", dx), ("dy", dy), ("dz", dz)))
    if isinstance(nsound, bool) or not isinstance(nsound, int) or nsound < 1:
        raise ValueError("nsound must be a positive integer")
    if torch.cuda.current_device() != state.device_id:
        raise ValueError("momentum RK state is bound to another HIP device")
    stream = stream if stream is not None else state.stream
    if stream_handle(stream) != stream_handle(state.stream):
        raise ValueError("momentum RK state must remain on its bound stream")
    a, m, hd = state.arrays, state.metrics, state.held
    ni, nj, nk = state.ni, state.nj, state.nk
    h = stream_handle(stream)
    _native.call("gm1hip_seed_held_momentum", ptr(hd["uten1"]), ptr(hd["vten1"]), ptr(hd["wten1"]), ptr(a["uten"]),
                 ptr(a["vten"]), ptr(a["wten"]), ni, nj, nk, h)
    if nrk >= 2:
        _native.call("gm1hip_bcwsfc", ptr(a["u3d"]), ptr(a["v3d"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(m["gz"]),
                     ptr(a["w3d"]), ni, nj, h)
        _native.call("gm1hip_bc2d_open", ptr(a["w3d"]), ni, nj, h)
    advection = launch_cuda_terrain_open_momentum_advection5(
        a, m, ni=ni, nj=nj, nk=nk, nrk=nrk, dx=dx, dy=dy, dz=dz, divx=state.divx, stream=stream,
        lateral_sides=state.lateral_sides)
    acoustic = launch_cuda_psolver3_terrain_open_stage(
        a, m, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=dt, nsound=nsound, dx=dx, dy=dy, dz=dz, stream=stream,
        lateral_sides=state.lateral_sides, ablations=acoustic_ablations or TerrainOpenSoundAblations())
    state.next_nrk = nrk + 1
    state.completed = nrk == 3
    receipt = {"nrk": nrk, "order": ["seed-held-tendencies"] + (["bcwsfc", "bc2d-w-surface"] if nrk >= 2 else [])
               + ["momentum-advection5", "psolver3-acoustic-stage"], "lateral_sides": list(state.lateral_sides),
               "target_arch": "gfx950"}
    return CudaMomentumRKStageLaunch(state, record(stream), advection, acoustic, receipt)


__all__ = [
    "CudaMomentumRKStageLaunch", "CudaMomentumRKState", "HELD_NAMES", "begin_cuda_momentum_rk_step",
    "create_cuda_momentum_rk_state", "launch_cuda_momentum_rk_stage",
]
