"""The moist stage bridge's stateful entry points on gfx950 (host mirror of cuda_moist_stage_bridge.py).

Per RK stage the reference runs three launchers against one bound state
(``CudaMoistStageBridgeState``, cuda_moist_stage_bridge.py:303-352):

* ``launch_cuda_eqtset2_stage_terms`` (:378): eqtset=2 thv, ppterm, piadv,
  the buoyancy carrier dum8 and thterm, with the open halos the later
  operators read;
* ``launch_cuda_moist_buoyancy`` (:433): the w and terrain u/v buoyancy
  added into the live tendencies;
* ``launch_cuda_post_acoustic_flux_bridge`` (:456): getdiv of the stage's
  time-averaged winds, then rru/rrv/rrw rebuilt in place as mass fluxes and
  divx;

with ``begin_cuda_moist_stage`` (:496) advancing the cadence guard between
stages.  The state owns dum8, thterm, divergence, theta_divergence_term and
divx; launches are asynchronous on the state's stream and return a ready
event.  The device work is ``gm1hip_stage_terms``, ``gm1hip_buoyancy`` and
``gm1hip_post_acoustic``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Mapping

import torch

from . import _native
from ._device import NG, lateral_code, no_alias, ptr, record, require, stream_handle, terrain_static_shapes
from ._native import Fl, I, P, U, signature
from .errors import refuse

signature("gm1hip_stage_terms", *([P] * 7), I, *([P] * 6), I, I, I, U, P)
signature("gm1hip_buoyancy", *([P] * 8), I, I, I, P)
signature("gm1hip_post_acoustic", *([P] * 14), I, I, I, Fl, Fl, Fl, P)

CUDA_MOIST_STAGE_BRIDGE_SCHEMA = "gm1_amd.moist-stage-bridge/v1"
MAX_SPECIES = 6


@dataclass
class CudaMoistStageBridgeState:
    ni: int
    nj: int
    nk: int
    base3d: bool
    dum8: torch.Tensor
    thterm: torch.Tensor
    divergence: torch.Tensor
    theta_divergence_term: torch.Tensor
    divx: torch.Tensor
    stream: Any
    device_id: int
    lateral_sides: tuple[int, int, int, int] = (2, 2, 2, 2)
    base_condensate: str = "exact-zero"
    terms_ready: bool = False
    buoyancy_applied: bool = False
    post_acoustic_ready: bool = False


@dataclass(frozen=True)
class CudaPostAcousticFluxLaunch:
    state: CudaMoistStageBridgeState
    ready_event: Any
    receipt: dict[str, Any] = field(default_factory=dict)


def _positive_int(name: str, value: Any, minimum: int) -> int:
    if isinstance(value, bool) or not isinstance(value, int) or value < minimum:
        raise ValueError(f"{name} must be an integer >= {minimum}")
    return value


def _spacing(name: str, value: Any) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise TypeError(f"{name} must be numeric")
    result = float(value)
    if not math.isfinite(result) or result <= 0.0:
        raise ValueError(f"{name} must be finite and > 0")
    return result


def create_cuda_moist_stage_bridge_state(
    *, ni: int, nj: int, nk: int, base3d: bool, stream: Any | None = None, device: Any = "cuda",
    lateral_sides: Any = (2, 2, 2, 2), compiled: Any = None, open_compiled: Any = None,
) -> CudaMoistStageBridgeState:
    """Allocate the bridge's own carriers (cuda_moist_stage_bridge.py:327-352).

    ``lateral_sides`` is the family the halo fills apply (the reference binds
    it through its compiled open-boundary image); ``compiled`` and
    ``open_compiled`` are accepted for the reference's call shape.
    """
    ni, nj, nk = _positive_int("ni", ni, 7), _positive_int("nj", nj, 7), _positive_int("nk", nk, 8)
    if type(base3d) is not bool:
        raise TypeError("base3d must be boolean")
    if base3d:
        refuse(code="GM1AMD-MSB-BASE3D", gate="moist-stage-bridge", option="base3d", value=True,
               reason="the gfx950 stage terms read one base column (th0, qv0, pi0 of shape (nk,))",
               remedy="pass the horizontally uniform base state as 1-D profiles")
    dev = torch.device(device)
    physical = (nk, nj, ni)
    z = lambda shape: torch.zeros(shape, dtype=torch.float32, device=dev)  # noqa: E731
    return CudaMoistStageBridgeState(
        ni, nj, nk, base3d, z((nk + 2, nj + 2 * NG, ni + 2 * NG)), z(physical), z(physical), z(physical),
        z(physical), stream, dev.index if dev.index is not None else torch.cuda.current_device(),
        tuple(int(s) for s in lateral_sides))


def _validate_common(state: CudaMoistStageBridgeState, arrays: Mapping[str, Any], metrics: Mapping[str, Any]):
    if not isinstance(state, CudaMoistStageBridgeState):
        raise TypeError("state must be CudaMoistStageBridgeState")
    if torch.cuda.current_device() != state.device_id:
        raise ValueError("moist-stage bridge is bound to another HIP device")
    ni, nj, nk, ng = state.ni, state.nj, state.nk, NG
    shapes = {
        "thv": (nk + 2, nj + 2, ni + 2), "ppterm": (nk + 2, nj + 2, ni + 2),
        "piadv": (nk + 2, nj + 2, ni + 2), "pp3d": (nk + 2, nj + 2, ni + 2),
        "uten": (nk + 2, nj + 2 * ng, ni + 1 + 2 * ng), "vten": (nk + 2, nj + 1 + 2 * ng, ni + 2 * ng),
        "wten": (nk + 3, nj + 2 * ng, ni + 2 * ng), "rru": (nk + 2, nj + 2 * ng, ni + 1 + 2 * ng),
        "rrv": (nk + 2, nj + 1 + 2 * ng, ni + 2 * ng), "rrw": (nk + 3, nj + 2 * ng, ni + 2 * ng),
    }
    a = {name: require(arrays[name], name, shape) for name, shape in shapes.items()}
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    names = ("c1", "c2", "rho0_3d", "rf0", "gz", "rgzu", "rgzv", "dzdx", "dzdy", "gxu", "gyv")
    m = {name: require(metrics[name], name, ms[name]) for name in names}
    return a, m


def _bound_stream(state: CudaMoistStageBridgeState, stream: Any | None) -> Any:
    stream = stream if stream is not None else state.stream
    if stream_handle(stream) != stream_handle(state.stream):
        raise ValueError("moist-stage bridge must remain on its bound stream")
    return stream


def launch_cuda_eqtset2_stage_terms(
    state: CudaMoistStageBridgeState, arrays: Mapping[str, Any], metrics: Mapping[str, Any], *, th0: Any, qv0: Any,
    pi0: Any, th3d: Any, q3d: Any, stream: Any | None = None, qc0: Any | None = None, qi0: Any | None = None,
) -> Any:
    """Current-stage moist terms and the halos the later operators read (cuda_moist_stage_bridge.py:378-430).

    ``q3d`` is (6, nk, nj, ni) in the order qv, qc, qr, qi, qs, qg (or fewer
    rows: the absent species are exact zeros); ``qc0``/``qi0`` are base.F's
    base-state condensate, ``None`` for CM1's zero.
    """
    a, _ = _validate_common(state, arrays, metrics)
    if state.terms_ready or state.buoyancy_applied or state.post_acoustic_ready:
        raise ValueError("begin the next moist stage before recomputing terms")
    stream = _bound_stream(state, stream)
    nk, nj, ni = state.nk, state.nj, state.ni
    physical = (nk, nj, ni)
    th0, qv0, pi0 = (require(v, n, (nk,)) for v, n in ((th0, "th0"), (qv0, "qv0"), (pi0, "pi0")))
    qc0 = None if qc0 is None else require(qc0, "qc0", (nk,))
    qi0 = None if qi0 is None else require(qi0, "qi0", (nk,))
    state.base_condensate = "exact-zero" if (qc0 is None and qi0 is None) else "base.F qc0/qi0"
    th3d = require(th3d, "th3d", physical)
    if not isinstance(q3d, torch.Tensor) or q3d.dim() != 4 or not 1 <= q3d.shape[0] <= MAX_SPECIES:
        raise ValueError(f"q3d must be (numq <= {MAX_SPECIES}, nk, nj, ni)")
    q3d = require(q3d, "q3d", (q3d.shape[0], *physical))
    no_alias({"th3d": th3d, "q3d": q3d, "thv": a["thv"], "ppterm": a["ppterm"], "piadv": a["piadv"],
              "dum8": state.dum8, "thterm": state.thterm}, "stage-term inputs and outputs must not alias")
    _native.call("gm1hip_stage_terms", ptr(th0), ptr(qv0), ptr(pi0), ptr(qc0), ptr(qi0), ptr(th3d), ptr(q3d),
                 int(q3d.shape[0]), ptr(a["pp3d"]), ptr(a["thv"]), ptr(a["ppterm"]), ptr(a["piadv"]),
                 ptr(state.dum8), ptr(state.thterm), ni, nj, nk, lateral_code(state.lateral_sides),
                 stream_handle(stream))
    state.terms_ready = True
    return record(stream)


def launch_cuda_moist_buoyancy(state: CudaMoistStageBridgeState, arrays: Mapping[str, Any],
                               metrics: Mapping[str, Any], *, stream: Any | None = None) -> Any:
    """Vertical and terrain-horizontal buoyancy into the live tendencies (cuda_moist_stage_bridge.py:433-447)."""
    a, m = _validate_common(state, arrays, metrics)
    if not state.terms_ready or state.buoyancy_applied or state.post_acoustic_ready:
        raise ValueError("buoyancy requires one fresh current-stage term production")
    stream = _bound_stream(state, stream)
    _native.call("gm1hip_buoyancy", ptr(state.dum8), ptr(m["c1"]), ptr(m["c2"]), ptr(m["gxu"]), ptr(m["gyv"]),
                 ptr(a["uten"]), ptr(a["vten"]), ptr(a["wten"]), state.ni, state.nj, state.nk,
                 stream_handle(stream))
    state.buoyancy_applied = True
    return record(stream)


def launch_cuda_post_acoustic_flux_bridge(
    state: CudaMoistStageBridgeState, arrays: Mapping[str, Any], metrics: Mapping[str, Any], *, dx: float,
    dy: float, dz: float, stream: Any | None = None,
) -> CudaPostAcousticFluxLaunch:
    """getdiv of the acoustic averages, then getdivx in place (cuda_moist_stage_bridge.py:456-493)."""
    a, m = _validate_common(state, arrays, metrics)
    if not state.buoyancy_applied or state.post_acoustic_ready:
        raise ValueError("post-acoustic bridge requires applied current-stage buoyancy")
    dx, dy, dz = (_spacing(n, v) for n, v in (("dx", dx), ("dy", dy), ("dz", dz)))
    stream = _bound_stream(state, stream)
    _native.call("gm1hip_post_acoustic", ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]), ptr(m["rho0_3d"]),
                 ptr(m["rf0"]), ptr(m["rgzu"]), ptr(m["rgzv"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(m["gz"]),
                 ptr(state.thterm), ptr(state.divergence), ptr(state.theta_divergence_term), ptr(state.divx),
                 state.ni, state.nj, state.nk, dx, dy, dz, stream_handle(stream))
    state.post_acoustic_ready = True
    receipt = {
        "schema": CUDA_MOIST_STAGE_BRIDGE_SCHEMA,
        "selection": {"eqtset": 2, "base_condensate": state.base_condensate, "terrain": True, "base3d": False,
                      "lateral_boundaries": list(state.lateral_sides)},
        "domain": {"ni": state.ni, "nj": state.nj, "nk": state.nk},
        "ordering": ["getdiv-on-acoustic-average-velocity", "getdivx-horizontal-mass-flux-in-place",
                     "getdivx-terrain-vertical-mass-flux-in-place", "getdivx-divergence"],
        "target_arch": "gfx950",
    }
    return CudaPostAcousticFluxLaunch(state, record(stream), receipt)


def begin_cuda_moist_stage(state: CudaMoistStageBridgeState) -> None:
    """Advance the cadence guard after a completed post-acoustic bridge (cuda_moist_stage_bridge.py:496-505)."""
    if not isinstance(state, CudaMoistStageBridgeState):
        raise TypeError("state must be CudaMoistStageBridgeState")
    if not state.post_acoustic_ready:
        raise ValueError("cannot begin another stage before post-acoustic getdiv/getdivx")
    state.terms_ready = False
    state.buoyancy_applied = False
    state.post_acoustic_ready = False


__all__ = [
    "CUDA_MOIST_STAGE_BRIDGE_SCHEMA", "CudaMoistStageBridgeState", "CudaPostAcousticFluxLaunch",
    "begin_cuda_moist_stage", "create_cuda_moist_stage_bridge_state", "launch_cuda_eqtset2_stage_terms",
    "launch_cuda_moist_buoyancy", "launch_cuda_post_acoustic_flux_bridge",
]
