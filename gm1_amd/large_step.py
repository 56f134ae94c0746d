"""The CM1 dynamical-core large step on gfx950 (composition).

Host mirror of the reference's large-step composer
``cuda_gm1_forecast.launch_cuda_gm1_large_step`` (:1569-1627) and its stage
driver ``launch_cuda_gm1_rk_stage`` (:1241-1345), in the same source order
(``gm1_runtime.large_step_schedule``, ``gm1_runtime.py:1969-2007``):

  solve1   held carriers zeroed (no SGS/TKE closure on this path), held
           compose of ppten and the microphysics held rates (moist)
  nrk=1..3 eqtset=2 stage terms -> held momentum seed (+ bcwsfc, bc2d for
           nrk>=2) -> momentum advection -> buoyancy -> acoustic stage ->
           post-acoustic getdiv/getdivx -> theta and moisture RK stage
           (moisture at nrk=3 through the positive-definite chain)
  finish   Kessler transition (held remove, EOS, dry-mass correction,
           driver, held recapture; moist) -> solve3 (ppx, ppi, halos,
           A <- current copies, density)

Every operator is one C-ABI call into ``libgm1hip.so``; the whole step is
enqueued on one HIP stream with no host synchronisation, so it can be
captured into a HIP graph and replayed.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np
import torch

from . import _native
from ._device import NG, lateral_code, ptr, stream_handle
from ._native import Fl, I, P, U, signature
from .fields import CaseConfig, base_state, initial_fields, initial_thermodynamics, metrics_for, profiles_for
from .stretched import (flat_stretched_column, launch_cuda_gm1_child_stretched_post_acoustic,
                        launch_cuda_gm1_flat_acoustic_stage)
from .grid import VerticalProfiles
from .kessler import (
    Gm1KesslerDriverSelection,
    create_cuda_gm1_kessler_driver_workspace,
    create_dry_mass_workspace,
    launch_cuda_apmasscon,
    launch_cuda_gm1_kessler_transition,
)
from .errors import refuse
from .psolver3 import TerrainOpenSoundAblations, device_profiles, launch_cuda_psolver3_terrain_open_stage
from .scalar_moist_rk import (BATCH_MAX, launch_scalar_stage_from_carrier, launch_scalar_stages_from_carriers,
                              moisture_weps_f64, theta_weps_f64, verify_horizontally_uniform)

signature("gm1hip_stage_terms", P, P, P, P, P, P, P, I, P, P, P, P, P, P, I, I, I, U, P)
signature("gm1hip_buoyancy", P, P, P, P, P, P, P, P, I, I, I, P)
signature("gm1hip_post_acoustic", *([P] * 14), I, I, I, Fl, Fl, Fl, P)
signature("gm1hip_rk_compose", P, P, P, P, P, ctypes.c_int64, Fl, P)
signature("gm1hip_seed_held_momentum", P, P, P, P, P, P, I, I, I, P)
signature("gm1hip_momentum_stage", *([P] * 26), I, I, I, Fl, Fl, Fl, U, P)
signature("gm1hip_momentum_stage_flat", *([P] * 23), I, I, I, Fl, Fl, Fl, U, P)
signature("gm1hip_held_compose", *([P] * 13), I, I, I, Fl, I, P)
signature("gm1hip_solve3_pressure", P, P, P, P, I, I, I, U, P)
signature("gm1hip_solve3_density", P, P, P, P, P, I, I, I, P)

THETA_WEPS = theta_weps_f64()
MOISTURE_WEPS = moisture_weps_f64(0)
KESSLER_HELD = ("qpten", "qtten", "qvten", "qcten")


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
    a["dum8"] = np.zeros(sh["s"], np.float32)
    for name in ("pp3d", "ppi", "ppx", "ppd", "piadv", "ppterm", "ppten", "thv", "qk", "pk1", "pk2", "mm", "fpk",
                 "elim_c", "elim_d"):
        a[name] = np.zeros(sh["p"], np.float32)
    for name in ("th3d", "tha", "thten1", "thterm", "divergence", "theta_div", "divx", "scalar_next"):
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
        for name in ("advx", "advy", "advz", "pp_mass", "fall_velocity") + KESSLER_HELD:
            a[name] = np.zeros(sh["cell"], np.float32)
        thermo = initial_thermodynamics(base, init["th3d"], init["q3d"][0])
        a["prs"], a["rho"] = thermo["prs"], thermo["rho"]
        a["rr"] = (np.float32(1.0) / thermo["rho"]).astype(np.float32)
        a["rf"] = np.zeros((nk + 1, nj, ni), np.float32)
        a["pi0_3d"], a["th0_3d"] = thermo["pi0_3d"], thermo["th0_3d"]
        a["ruh"] = np.ones(ni, np.float32)
        a["rvh"] = np.ones(nj, np.float32)
        a["rmh"] = met["rmh"].copy()
        a["mh"] = met["mh"].copy()
        a["rain"] = np.zeros((1, nj, ni), np.float32)
        a["prate"] = np.zeros((nj, ni), np.float32)
        a["asq"] = np.zeros(3, np.float64)
        a["qbudget"] = np.zeros(10, np.float64)
    base_d = {"th0": base["th0"], "pi0": base["pi0"], "qv0": base["qv0"]}
    return {"cfg": cfg, "prof": prof, "base": base_d, "metrics": met, "arrays": a, "numq": numq,
            "levels": cfg.levels() if cfg.stretched is not None else None}


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
    selection: Gm1KesslerDriverSelection | None = None
    kessler_workspace: Any = None
    dry_mass: Any = None
    mass1: float | None = None
    column_pass: str = "pipelined"
    # gxu and gyv are +0 everywhere (flat terrain), measured once from the host metrics
    slopes_zero: bool = False
    # c1 and c2 hold one word per level at every column (flat terrain): the scalar advection reads one column
    c12_columns: bool = False
    # the scalars' second and third carriers: stage 1 writes the second (th3d, q3d), stage 2 the third
    # (th3d2, q3d2), stage 3 reads those and writes th3d/q3d; also the fused column passes' workspace
    alt: dict = field(default_factory=dict)
    receipts: list = field(default_factory=list)
    # a stretched (flat) column runs the terrain_flag=.false. family
    flat: Any = None
    # uten1, vten1, wten1 and qten's species beyond qc hold +0 since their last fill and no solve1
    # hook has run: a step without a hook then skips the fills and its passes form 0 + x themselves
    held_clean: bool = False
    zero_held: bool = False
    # scalars composed together in one pass at the stages before the last (1: one pass each)
    scalar_batch: int = 2


def slopes_are_zero(metrics: dict[str, Any]) -> bool:
    """True when the terrain slopes gxu and gyv are +0 in every word (flat terrain)."""
    for name in ("gxu", "gyv"):
        g = np.asarray(metrics[name])
        if g.size and (np.any(g != 0) or np.any(np.signbit(g))):
            return False
    return True


def create_gm1_large_step_state(host: dict[str, Any], device: torch.device | str = "cuda",
                                stream: Any | None = None, *, dry_mass_correction: bool = True,
                                equation_set: int = 2) -> Gm1LargeStepState:
    """Upload the state; for a moist run also the Kessler workspaces and the dry-mass ledger.

    The ledger ``mass1`` is measured once here (one host read, outside any
    step), as the reference persists it at initialisation.
    """
    dev = torch.device(device)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    st = Gm1LargeStepState(
        host["cfg"], host["prof"], {k: t(v) for k, v in host["base"].items()},
        {k: t(v) for k, v in host["metrics"].items()}, {k: t(v) for k, v in host["arrays"].items()},
        host["numq"], stream)
    st.slopes_zero = slopes_are_zero(host["metrics"])
    if st.slopes_zero:  # the device copies are checked too, before any step can be captured
        from .psolver3 import verify_terrain_slopes_zero

        st.slopes_zero = verify_terrain_slopes_zero(st.metrics)
    st.c12_columns = verify_horizontally_uniform(st.metrics)
    if host.get("levels") is not None:
        st.flat = flat_stretched_column(host["levels"], dev)
        if st.cfg.nk <= 128:  # the flat pass with its own levels' pressure words in registers
            st.column_pass = "resident"
    # device copies of the vertical profiles are made here, never inside a step
    # (a step may be captured into a graph, where host-to-device copies are not allowed)
    device_profiles(st.prof, dev)
    if st.flat is not None:
        device_profiles(st.flat.levels.profiles(), dev)
    st.alt = {"th3d": torch.empty_like(st.a["th3d"]), "th3d2": torch.empty_like(st.a["th3d"])}
    if st.flat is not None:  # the fused column passes' wind and ppd workspace
        for name in ("u3d", "v3d", "ppd"):
            st.alt[name] = torch.empty_like(st.a[name])
    if st.numq:
        st.alt["q3d"] = torch.empty_like(st.a["q3d"])
        st.alt["q3d2"] = torch.empty_like(st.a["q3d"])
        cfg = st.cfg
        st.selection = Gm1KesslerDriverSelection(equation_set=equation_set)
        st.kessler_workspace = create_cuda_gm1_kessler_driver_workspace(cfg.ni, cfg.nj, cfg.nk, dev)
        if dry_mass_correction:
            st.dry_mass = create_dry_mass_workspace(cfg.ni, cfg.nj, cfg.nk, dev)
            launch_cuda_apmasscon({"rho": st.a["rho"], "ruh": st.a["ruh"], "rvh": st.a["rvh"], "rmh": st.a["rmh"]},
                                  ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, mass1=0.0, workspace=st.dry_mass, stream=stream)
            torch.cuda.synchronize(dev)
            st.mass1 = float(st.dry_mass.ledger[0].item())
    return st


def state_to_host(st: Gm1LargeStepState) -> dict[str, np.ndarray]:
    return {k: v.detach().cpu().numpy() for k, v in st.a.items()}


def _kessler_arrays(st: Gm1LargeStepState) -> dict[str, torch.Tensor]:
    a = st.a
    return {"ruh": a["ruh"], "rvh": a["rvh"], "rmh": a["rmh"], "mh": a["mh"], "pi0": a["pi0_3d"],
            "th0": a["th0_3d"], "rho": a["rho"], "rr": a["rr"], "pp3d": a["pp_mass"], "th3d": a["th3d"],
            "prs": a["prs"], "qv": a["q3d"][0], "qc": a["q3d"][1], "qr": a["q3d"][2],
            "fall_velocity": a["fall_velocity"], "rain": a["rain"], "prate": a["prate"], "asq": a["asq"],
            "qbudget": a["qbudget"]}


def _scalar_carriers(st: Gm1LargeStepState, nrk: int, alternate: bool) -> tuple[Any, Any]:
    """Where this stage's theta' and moisture live: when alternating, the second carriers during stage 2 and
    the third during stage 3."""
    a = st.a
    if alternate and nrk == 2:
        return st.alt["th3d"], st.alt.get("q3d")
    if alternate and nrk == 3:
        return st.alt["th3d2"], st.alt.get("q3d2")
    return a["th3d"], a.get("q3d")


def _rk_stage(st: Gm1LargeStepState, nrk: int, alternate: bool = False) -> None:
    cfg, a, m, b = st.cfg, st.a, st.metrics, st.base
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    sides = cfg.lateral_sides
    code = lateral_code(sides)
    h = stream_handle(st.stream)
    th3d, q3d = _scalar_carriers(st, nrk, alternate)
    _native.call("gm1hip_stage_terms", ptr(b["th0"]), ptr(b["qv0"]), ptr(b["pi0"]), None, None, ptr(th3d),
                 ptr(q3d), st.numq, ptr(a["pp3d"]), ptr(a["thv"]), ptr(a["ppterm"]), ptr(a["piadv"]),
                 ptr(a["dum8"]), ptr(a["thterm"]), ni, nj, nk, code, h)
    if st.flat is not None:
        _flat_rk_stage(st, nrk, alternate)
        return
    if nrk >= 2:
        _native.call("gm1hip_bcwsfc", ptr(a["u3d"]), ptr(a["v3d"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(m["gz"]),
                     ptr(a["w3d"]), ni, nj, h)
        _native.call("gm1hip_bc2d_open", ptr(a["w3d"]), ni, nj, h)
    # the held seed (uten = uten1, ...) folded into the momentum advection updates: the seed
    # touches no array bcwsfc/bc2d read or write, so it moves past them unchanged
    held = (None, None, None) if st.zero_held else (a["uten1"], a["vten1"], a["wten1"])
    _native.call("gm1hip_momentum_stage", ptr(a["u3d"]), ptr(a["v3d"]), ptr(a["w3d"]), *(ptr(x) for x in held),
                 ptr(a["uten"]), ptr(a["vten"]), ptr(a["wten"]), ptr(a["rru"]),
                 ptr(a["rrv"]), ptr(a["rrw"]), ptr(a["uh"]), ptr(a["vh"]), ptr(m["rho0_3d"]), ptr(m["rf0"]),
                 ptr(m["rrf0"]), ptr(m["c2"]), ptr(m["gz"]), ptr(m["gzu"]), ptr(m["gzv"]), ptr(m["rgzu"]),
                 ptr(m["rgzv"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(a["divx"]), ni, nj, nk, cfg.dx, cfg.dy, cfg.dz,
                 code, h)
    _native.call("gm1hip_buoyancy", ptr(a["dum8"]), ptr(m["c1"]), ptr(m["c2"]), ptr(m["gxu"]), ptr(m["gyv"]),
                 ptr(a["uten"]), ptr(a["vten"]), ptr(a["wten"]), ni, nj, nk, h)
    launch_cuda_psolver3_terrain_open_stage(
        a, m, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=cfg.dt, nsound=cfg.nsound, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz,
        stream=st.stream, lateral_sides=sides, profiles=st.prof, ablations=TerrainOpenSoundAblations(),
        column_pass=st.column_pass, terrain_slopes_zero=st.slopes_zero)
    _native.call("gm1hip_post_acoustic", ptr(a["rru"]), ptr(a["rrv"]), ptr(a["rrw"]), ptr(m["rho0_3d"]),
                 ptr(m["rf0"]), ptr(m["rgzu"]), ptr(m["rgzv"]), ptr(m["dzdx"]), ptr(m["dzdy"]), ptr(m["gz"]),
                 ptr(a["thterm"]), ptr(a["divergence"]), ptr(a["theta_div"]), ptr(a["divx"]), ni, nj, nk,
                 cfg.dx, cfg.dy, cfg.dz, h)
    mass = {"rru": a["rru"], "rrv": a["rrv"], "rrw": a["rrw"]}
    # theta' and each moisture species advected straight from their carriers (the packed
    # (th0 - 300) + theta' and the lateral-family halos formed in the tiles' loads) and
    # composed A-anchored in one pass
    _scalar_stages(st, nrk, mass, flat_column=None, rdz=float(np.float32(1.0) / np.float32(cfg.dz)),
                   alternate=alternate)


def _flat_rk_stage(st: Gm1LargeStepState, nrk: int, alternate: bool = False) -> None:
    """The rest of one RK stage on a flat stretched column (cuda_cartesian_dynamics.py:1070-1262).

    No bcwsfc/bc2d (terrain_flag statements), then the flat momentum,
    buoyancy, acoustic stage, getdiv/getdivx, theta and moisture texts.
    """
    cfg, a, m = st.cfg, st.a, st.metrics
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    sides = cfg.lateral_sides
    # the held seed, the flat momentum advection and the flat buoyancy in three passes
    # (the seed folded into the advection updates, the buoyancy into the w update)
    held = (None, None, None) if st.zero_held else (a["uten1"], a["vten1"], a["wten1"])
    _native.call("gm1hip_momentum_stage_flat", ptr(a["u3d"]), ptr(a["v3d"]), ptr(a["w3d"]), *(ptr(x) for x in held),
                 ptr(a["uten"]), ptr(a["vten"]), ptr(a["wten"]), ptr(a["rru"]),
                 ptr(a["rrv"]), ptr(a["rrw"]), ptr(a["uh"]), ptr(a["vh"]), ptr(m["rho0_3d"]), ptr(m["rf0"]),
                 ptr(m["rr0"]), ptr(m["rrf0"]), ptr(m["c1"]), ptr(m["c2"]), ptr(a["dum8"]), ptr(st.flat.column),
                 ptr(a["divx"]), ni, nj, nk, cfg.dx, cfg.dy, float(st.flat.rdz), lateral_code(sides),
                 stream_handle(st.stream))
    launch_cuda_gm1_flat_acoustic_stage(a, m, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=cfg.dt, nsound=cfg.nsound, dx=cfg.dx,
                                        dy=cfg.dy, vertical=st.flat, stream=st.stream, lateral_sides=sides,
                                        column_pass=st.column_pass, alternates=st.alt)
    launch_cuda_gm1_child_stretched_post_acoustic(a, m, st.flat, ni=ni, nj=nj, nk=nk, dx=cfg.dx, dy=cfg.dy,
                                                  thterm=a["thterm"], divergence=a["divergence"],
                                                  theta_divergence_term=a["theta_div"], divx=a["divx"],
                                                  stream=st.stream)
    mass = {"rru": a["rru"], "rrv": a["rrv"], "rrw": a["rrw"]}
    _scalar_stages(st, nrk, mass, flat_column=st.flat.column, rdz=float(st.flat.rdz), alternate=alternate)


def _scalar_stages(st: Gm1LargeStepState, nrk: int, mass: dict[str, Any], *, flat_column: Any | None,
                   rdz: float, alternate: bool = False) -> None:
    """Theta' then each moisture species (the final stage's rows through the pdef chain).

    ``alternate``: stage 1 writes the second carriers (st.alt), stage 2 reads them and writes the third, stage 3
    reads those and writes th3d/q3d -- the same words as a copy-back after every scalar, with no copy.
    """
    cfg, a, m, b = st.cfg, st.a, st.metrics, st.base
    grid = dict(ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, nrk=nrk, dt=cfg.dt, dx=cfg.dx, dy=cfg.dy, rdz=rdz,
                lateral_sides=cfg.lateral_sides, flat_column=flat_column, stream=st.stream,
                c12_columns=st.c12_columns)
    if alternate:
        src_th, src_q = _scalar_carriers(st, nrk, True)
        if nrk == 1:
            out_th, out_q = st.alt["th3d"], st.alt.get("q3d")
        elif nrk == 2:
            out_th, out_q = st.alt["th3d2"], st.alt.get("q3d2")
        else:
            out_th, out_q = a["th3d"], a.get("q3d")
        grid["copy_back"] = False
    else:
        src_th, src_q = a["th3d"], a.get("q3d")
        out_th, out_q = a["scalar_next"], None
    if alternate and nrk < 3 and st.scalar_batch > 1:
        # theta' and the species in batches sharing the mass fluxes: the same words, fewer passes
        rows = [(src_th, b["th0"], a["tha"], a["thten1"], a["theta_div"], out_th)]
        rows += [(src_q[sp], None, a["qa"][sp], a["qten"][sp], None, out_q[sp]) for sp in range(st.numq)]
        grid.pop("copy_back")
        n = min(st.scalar_batch, BATCH_MAX)
        for lo in range(0, len(rows), n):
            cols = list(zip(*rows[lo:lo + n]))
            launch_scalar_stages_from_carriers(list(cols[0]), list(cols[1]), mass, m, a["divx"], list(cols[2]),
                                               list(cols[3]), list(cols[4]), list(cols[5]), **grid)
        return
    launch_scalar_stage_from_carrier(src_th, b["th0"], mass, m, a["divx"], a["tha"], a["thten1"],
                                     a["theta_div"], out_th, weps=THETA_WEPS, **grid)
    for sp in range(st.numq):
        scratch = out_q[sp] if out_q is not None else a["scalar_next"]
        if nrk < 3:
            launch_scalar_stage_from_carrier(src_q[sp], None, mass, m, a["divx"], a["qa"][sp], a["qten"][sp], None,
                                             scratch, weps=moisture_weps_f64(sp), **grid)
        else:
            launch_scalar_stage_from_carrier(src_q[sp], None, mass, m, a["divx"], a["qa"][sp], a["qten"][sp], None,
                                             scratch, pdef_species=sp,
                                             pdef_scratch={"advx": a["advx"], "advy": a["advy"],
                                                           "advz": a["advz"]}, **grid)


@dataclass(frozen=True)
class Gm1Solve1Carriers:
    """The held (solve1) tendencies a solve1 hook may add into (cuda_gm1_forecast.py Gm1Solve1Carriers)."""

    uten1: torch.Tensor
    vten1: torch.Tensor
    wten1: torch.Tensor
    thten1: torch.Tensor
    qten: torch.Tensor | None
    ppten: torch.Tensor


# the reference's name for the composer's state
CudaGm1ForecastState = Gm1LargeStepState


def launch_cuda_gm1_large_step(
    st: Gm1LargeStepState, *, solve1_hook: Callable[[Gm1LargeStepState, Gm1Solve1Carriers], object] | None = None,
    post_stage_hook: Callable[[Gm1LargeStepState, int], object] | None = None, lateral_forcing: Any | None = None,
    reflectivity_due: bool = False, allow_incomplete_validation: bool = False,
) -> dict[str, Any]:
    """Enqueue one complete large step on ``st.stream`` (no host synchronisation).

    The reference composer's keywords (cuda_gm1_forecast.py:1569-1627):
    ``solve1_hook(st, carriers)`` runs after the held carriers are cleared
    (this path has no built-in closure) and before the held compose, and may
    add tendencies into them; ``post_stage_hook(st, nrk)`` runs after each RK
    stage.  Both enqueue on ``st.stream``.  ``lateral_forcing`` (a parent
    model's boundary nudging) and ``reflectivity_due`` (no reflectivity
    consumer for ptype=1) are refused before any launch.
    """
    if solve1_hook is not None and not callable(solve1_hook):
        raise TypeError("solve1_hook must be callable or None")
    if post_stage_hook is not None and not callable(post_stage_hook):
        raise TypeError("post_stage_hook must be callable or None")
    if lateral_forcing is not None:
        refuse(code="GM1AMD-LATERAL-FORCING", gate="large-step", option="lateral_forcing", value="given",
               reason="parent-model lateral forcing is not part of the gfx950 large step",
               remedy="run with the open-radiative or periodic lateral families and lateral_forcing=None")
    if type(reflectivity_due) is not bool:
        raise TypeError("reflectivity_due must be boolean")
    if reflectivity_due:
        refuse(code="GM1-KESSLER-REFLECTIVITY-UNPORTED", gate="physics-seam", option="reflectivity_due", value=True,
               reason="CM1 does not admit the selected reflectivity consumer for ptype=1",
               remedy="disable reflectivity or select an admitted microphysics consumer")
    cfg, a = st.cfg, st.a
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    h = stream_handle(st.stream)
    code = lateral_code(cfg.lateral_sides)
    # the held tendencies start each step at +0; with no solve1 hook and the fields still +0 from
    # their last fill, the passes that read them form 0 + x without the fills (the same words)
    st.zero_held = solve1_hook is None and st.held_clean
    if not st.zero_held:
        with torch.cuda.stream(st.stream) if st.stream is not None else _nullcontext():
            for name in ("uten1", "vten1", "wten1", "thten1", "ppten"):
                a[name].zero_()
            if st.numq:
                a["qten"].zero_()
    st.held_clean = solve1_hook is None and post_stage_hook is None  # (a hook may write the held fields)
    solve1_result = None
    if solve1_hook is not None:
        carriers = Gm1Solve1Carriers(a["uten1"], a["vten1"], a["wten1"], a["thten1"], a.get("qten"), a["ppten"])
        solve1_result = solve1_hook(st, carriers)
    if st.numq:
        b = st.base
        _native.call("gm1hip_held_compose", ptr(b["th0"]), ptr(b["pi0"]), ptr(a["rho"]), ptr(a["ppi"]),
                     ptr(a["tha"]), ptr(a["qa"]), ptr(a["thten1"]), ptr(a["qten"]), ptr(a["ppten"]),
                     ptr(a["qpten"]), ptr(a["qtten"]), ptr(a["qvten"]), ptr(a["qcten"]), ni, nj, nk, cfg.dt,
                     int(st.zero_held), h)
    stage_hooks = []
    # with no hook between stages, the scalars rotate through three carriers (no copy-back)
    alternate = post_stage_hook is None
    for nrk in (1, 2, 3):
        _rk_stage(st, nrk, alternate)
        if post_stage_hook is not None:
            stage_hooks.append(post_stage_hook(st, nrk))
    if st.numq:
        launch_cuda_gm1_kessler_transition(
            _kessler_arrays(st), {n: a[n] for n in KESSLER_HELD}, ni=ni, nj=nj, nk=nk, dx=cfg.dx, dy=cfg.dy,
            dz=cfg.dz, dt=cfg.dt, selection=st.selection, workspace=st.kessler_workspace,
            pressure_family=a["pp3d"], mass1=st.mass1, dry_mass=st.dry_mass, stream=st.stream)
    _native.call("gm1hip_solve3_pressure", ptr(a["ppx"]), ptr(a["pp3d"]), ptr(a["ppi"]),
                 ptr(a["pp_mass"]) if st.numq else None, ni, nj, nk, code, h)
    with torch.cuda.stream(st.stream) if st.stream is not None else _nullcontext():
        a["ua"].copy_(a["u3d"])
        a["va"].copy_(a["v3d"])
        a["wa"].copy_(a["w3d"])
        a["tha"].copy_(a["th3d"])
        if st.numq:
            a["qa"].copy_(a["q3d"])
    if st.numq:
        m = st.metrics
        _native.call("gm1hip_solve3_density", ptr(a["rho"]), ptr(m["c1"]), ptr(m["c2"]), ptr(a["rr"]), ptr(a["rf"]),
                     ni, nj, nk, h)
    st.step_index += 1
    return {"step": st.step_index, "cells": cfg.cells, "lateral_sides": list(cfg.lateral_sides),
            "moist": bool(st.numq), "solve1_hook": solve1_hook is not None,
            "post_stage_hook": post_stage_hook is not None, "stages": [1, 2, 3],
            "hook_results": {"solve1": solve1_result, "post_stage": stage_hooks},
            "allow_incomplete_validation": bool(allow_incomplete_validation)}


def kessler_status(st: Gm1LargeStepState) -> dict[str, int]:
    """Max satadj iterations and non-converged cells so far (reads the device; synchronises)."""
    s = st.kessler_workspace.status.detach().cpu().tolist()
    return {"max_satadj_iterations": int(s[0]), "non_converged_cells": int(s[1])}


class _nullcontext:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False
