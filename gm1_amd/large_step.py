"""The CM1 dynamical-core large step on gfx950 (composition).

Host mirror of the reference's large-step composer
``cuda_gm1_forecast.launch_cuda_gm1_large_step`` (:1569-1627) and its stage
driver ``launch_cuda_gm1_rk_stage`` (:1241-1345), in the same source order
(``gm1_runtime.large_step_schedule``, ``gm1_runtime.py:1969-2007``):

  solve1   held carriers zeroed (no SGS/TKE closure on this path), held
           compose of ppten and the microphysics held rates
  nrk=1..3 eqtset=2 stage terms -> held momentum seed (+ bcwsfc, bc2d for
           nrk>=2) -> momentum advection -> buoyancy -> acoustic stage ->
           post-acoustic getdiv/getdivx -> theta and moisture RK stage
  finish   microphysics (Kessler, moist runs) -> solve3 (ppx, ppi, halos,
           A <- current copies, density)

Every operator is one C-ABI call into ``libgm1hip.so``; the whole step is
enqueued on one HIP stream with no host synchronisation, so it can be
captured into a HIP graph and replayed.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from . import _native
from ._device import NG, lateral_code, ptr, stream_handle
from ._native import Fl, I, P, U, signature
from .advection5 import launch_cuda_scalar_advection5
from .fields import CaseConfig, base_state, initial_fields, metrics_for, profiles_for
from .grid import VerticalProfiles
from .momentum_advection5 import launch_cuda_terrain_open_momentum_advection5
from .psolver3 import TerrainOpenSoundAblations, launch_cuda_psolver3_terrain_open_stage

signature("gm1hip_stage_terms", P, P, P, P, P, P, P, I, P, P, P, P, P, P, I, I, I, U, P)
signature("gm1hip_buoyancy", P, P, P, P, P, P, P, P, I, I, I, P)
signature("gm1hip_post_acoustic", *([P] * 14), I, I, I, Fl, Fl, Fl, P)
signature("gm1hip_pack_scalar", P, P, P, I, I, I, I, U, P)
signature("gm1hip_rk_compose", P, P, P, P, P, ctypes.c_int64, Fl, P)
signature("gm1hip_seed_held_momentum", P, P, P, P, P, P, I, I, I, P)
signature("gm1hip_held_compose", *([P] * 13), I, I, I, Fl, P)
signature("gm1hip_solve3_pressure", P, P, P, I, I, I, U, P)
signature("gm1hip_solve3_density", P, P, P, P, P, I, I, I, P)

CM1_EPSILON_F32 = 1.0e-18
THETA_WEPS = float(np.float32(10.0) * np.float32(CM1_EPSILON_F32))
MOISTURE_WEPS = float(np.float32(1.0e-2) * np.float32(CM1_EPSILON_F32))


def padded_shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    return {
        "u": (nk + 2, nj + 2 * NG, ni + 1 + 2 * NG), "v": (nk + 2, nj + 1 + 2 * NG, ni + 2 * NG),
        "w": (nk + 3, nj + 2 * NG, ni + 2 * NG), "s": (nk + 2, nj + 2 * NG, ni + 2 * NG),
        "p": (nk + 2, nj + 2, ni + 2), "cell": (nk, nj, ni),
    }


def host_initial_state(cfg: CaseConfig) -> dict[str, Any]:
    """Every carrier of the step as NumPy arrays (the state both paths start from)."""
    prof = profiles_for(cfg)
    base = base_state(cfg, prof)
    met = metrics_for(cfg, base)
    init = initial_fields(cfg, prof, base)
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    sh = padded_shapes(ni, nj, nk)
    a: dict[str, np.ndarray] = {}
    for name in ("u3d", "ua", "uten", "uten1", "rru"):
        a[name] = np.zeros(sh["u"], np.float32)
    for name in ("v3d", "va", "vten", "vten1", "rrv"):
        a[name] = np.zeros(sh["v"], np.float32)
    for name in ("w3d", "wa", "wten", "wten1", "rrw"):
        a[name] = np.zeros(sh["w"], np.float32)
    for name in ("dum8", "theta_padded", "moisture_padded"):
        a[name] = np.zeros(sh["s"], np.float32)
    for name in ("pp3d", "ppi", "ppx", "ppd", "piadv", "ppterm", "ppten", "thv", "qk", "pk1", "pk2", "mm", "fpk",
                 "elim_c", "elim_d"):
        a[name] = np.zeros(sh["p"], np.float32)
    for name in ("th3d", "tha", "thten1", "thterm", "divergence", "theta_div", "divx", "advection"):
        a[name] = np.zeros(sh["cell"], np.float32)
    a["radbcw"] = np.zeros((nk + 2, nj + 2 * NG), np.float32)
    a["radbce"] = np.zeros((nk + 2, nj + 2 * NG), np.float32)
    a["radbcs"] = np.zeros((nk + 2, ni + 2 * NG), np.float32)
    a["radbcn"] = np.zeros((nk + 2, ni + 2 * NG), np.float32)
    a["uh"] = np.ones(ni + 2 * NG, np.float32)
    a["vh"] = np.ones(nj + 2 * NG, np.float32)
    a["ruf"] = np.ones(ni + 1 + 2 * NG, np.float32)
    a["rvf"] = np.ones(nj + 1 + 2 * NG, np.float32)
    # horizontally uniform winds fill halos consistently for any lateral family
    a["u3d"][1 : nk + 1] = init["u_profile"][:, None, None]
    a["v3d"][1 : nk + 1] = init["v_profile"][:, None, None]
    a["ua"][...] = a["u3d"]
    a["va"][...] = a["v3d"]
    a["th3d"][...] = init["th3d"]
    a["tha"][...] = init["th3d"]
    numq = 3 if cfg.moist else 0
    if numq:
        a["q3d"] = init["q3d"].copy()
        a["qa"] = init["q3d"].copy()
        a["qten"] = np.zeros_like(a["q3d"])
    base_d = {"th0": base["th0"], "pi0": base["pi0"], "qv0": base["qv0"]}
    return {"cfg": cfg, "prof": prof, "base": base_d, "metrics": met, "arrays": a, "numq": numq}


@dataclass
class Gm1LargeStepState:
    cfg: CaseConfig
    prof: VerticalProfiles
    base: dict[str, torch.Tensor]
    metrics: dict[str, torch.Tensor]
    a: dict[str, torch.Tensor]
    numq: int
    stream: Any | None = None
    step_index: int = 0
    receipts: list = field(default_factory=list)


def create_gm1_large_step_state(host: dict[str, Any], device: torch.device | str = "cuda",
                                stream: Any | None = None) -> Gm1LargeStepState:
    dev = torch.device(device)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    return Gm1LargeStepState(
        host["cfg"], host["prof"], {k: t(v) for k, v in host["base"].items()},
        {k: t(v) for k, v in host["metrics"].items()}, {k: t(v) for k, v in host["arrays"].items()},
        host["numq"], stream)


def state_to_host(st: Gm1LargeStepState) -> dict[str, np.ndarray]:
    return {k: v.detach().cpu().numpy() for k, v in st.a.items()}


def _rk_stage(st: Gm1LargeStepState, nrk: int) -> None:
    cfg, a, m, b = st.cfg, st.a, st.metrics, st.base
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    sides = cfg.lateral_sides
    code = lateral_code(sides)
    h = stream_handle(st.stream)
    q3d = a.get("q3d")
    _native.call("gm1hip_stage_terms", ptr(b["th0"]), ptr(b["qv0"]), ptr(b["pi0"]), None, None, ptr(a["th3d"]),
                 ptr(q3d), st.numq, ptr(a["pp3d"]), ptr(a["thv"]), ptr(a["ppterm"]), ptr(a["piadv"]),
                 ptr(a["dum8"]), ptr(a["thterm"]), ni, nj, nk, code, h)
    _native.call("gm1hip_seed_held_momentum", ptr(a["uten1"]), ptr(a["vten1"]), ptr(a["wten1"]), ptr(a["uten"]),
                 ptr(a["vten"]), ptr(a["wten"]), ni, nj, nk, h)
    if nrk >= 2:
        _native.call("gm1hip_bcwsfc", ptr(a["u3d"]), ptr(a["v3d"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(m["gz"]),
                     ptr(a["w3d"]), ni, nj, h)
        _native.call("gm1hip_bc2d_open", ptr(a["w3d"]), ni, nj, h)
    launch_cuda_terrain_open_momentum_advection5(
        a, m, ni=ni, nj=nj, nk=nk, nrk=nrk, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, divx=a["divx"], stream=st.stream,
        lateral_sides=sides)
    _native.call("gm1hip_buoyancy", ptr(a["dum8"]), ptr(m["c1"]), ptr(m["c2"]), ptr(m["gxu"]), ptr(m["gyv"]),
                 ptr(a["uten"]), ptr(a["vten"]), ptr(a["wten"]), ni, nj, nk, h)
    launch_cuda_psolver3_terrain_open_stage(
        a, m, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=cfg.dt, nsound=cfg.nsound, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz,
        stream=st.stream, lateral_sides=sides, profiles=st.prof, ablations=TerrainOpenSoundAblations())
    _native.call("gm1hip_post_acoustic", ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]), ptr(m["rho0_3d"]),
                 ptr(m["rf0"]), ptr(m["rgzu"]), ptr(m["rgzv"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(m["gz"]),
                 ptr(a["thterm"]), ptr(a["divergence"]), ptr(a["theta_div"]), ptr(a["divx"]), ni, nj, nk,
                 cfg.dx, cfg.dy, cfg.dz, h)
    dttmp = float(np.float32(np.float32(cfg.dt) / np.float32(4 - nrk)))
    n = ni * nj * nk
    # theta: (th0 - 300) + th3d, family halos, advection, A-anchored compose
    _native.call("gm1hip_pack_scalar", ptr(b["th0"]), ptr(a["th3d"]), ptr(a["theta_padded"]), ni, nj, nk, 1, code, h)
    mass = {"rru": a["rru"], "rrv": a["rrv"], "rrw": a["rrw"]}
    launch_cuda_scalar_advection5(
        a["theta_padded"], mass, m, a["divx"], a["advection"], None, ni=ni, nj=nj, nk=nk, nrk=nrk, dx=cfg.dx,
        dy=cfg.dy, dz=cfg.dz, weps=THETA_WEPS, stream=st.stream, lateral_sides=sides)
    _native.call("gm1hip_rk_compose", ptr(a["tha"]), ptr(a["thten1"]), ptr(a["theta_div"]), ptr(a["advection"]),
                 ptr(a["th3d"]), n, dttmp, h)
    if st.numq:
        raise NotImplementedError("moisture transport is composed by the moist large step")


def launch_cuda_gm1_large_step(st: Gm1LargeStepState) -> dict[str, Any]:
    """Enqueue one complete large step on ``st.stream`` (no host synchronisation)."""
    cfg, a = st.cfg, st.a
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    h = stream_handle(st.stream)
    code = lateral_code(cfg.lateral_sides)
    with torch.cuda.stream(st.stream) if st.stream is not None else _nullcontext():
        for name in ("uten1", "vten1", "wten1", "thten1", "ppten"):
            a[name].zero_()
    for nrk in (1, 2, 3):
        _rk_stage(st, nrk)
    _native.call("gm1hip_solve3_pressure", ptr(a["ppx"]), ptr(a["pp3d"]), ptr(a["ppi"]), ni, nj, nk, code, h)
    with torch.cuda.stream(st.stream) if st.stream is not None else _nullcontext():
        a["ua"].copy_(a["u3d"])
        a["va"].copy_(a["v3d"])
        a["wa"].copy_(a["w3d"])
        a["tha"].copy_(a["th3d"])
    st.step_index += 1
    return {"step": st.step_index, "cells": cfg.cells, "lateral_sides": list(cfg.lateral_sides)}


class _nullcontext:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False
