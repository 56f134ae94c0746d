"""Dry large step and its adjoint on halo-free periodic fields (gfx950).

GM1 has no adjoint.  This module runs the dry large step of the adjoint
configuration (flat terrain, horizontally uniform base state, CM1's
(1, 1, 1, 1) lateral sides) on fields without halos -- u, v and cell fields
(nk, nj, ni), w (nk+1, nj, ni) for levels 1..nk+1 -- and its adjoint:

* ``DryPeriodicStep.forward_stage`` / ``forward`` restate one RK stage / one
  large step (cuda_gm1_forecast.py:1241-1345, 1414-1457) in the reference's
  association order, keeping every acoustic substep's state;
* ``DryPeriodicStep.adjoint`` is the transpose of one large step at its
  trajectory: upwind choices held at their forward values, the smeps
  clamps taken as the identity, and the WENO5 weights and positive-definite
  limiters of the final stage's theta fluxes either differentiated
  (``weno="exact"``, the default: the gradient of the step itself) or frozen
  at their forward values (``weno="frozen"``);
* ``storm_intensity`` is the adjoint configuration's loss (base-density
  weighted mean of w^2 over the w levels between 2 and 8 km) and
  ``gradient`` its gradient with respect to the initial state over a window
  of large steps, recomputing forward segments from binomially placed
  checkpoints (``revolve_schedule``).
"""

from __future__ import annotations

import ctypes
from typing import Any, Mapping

import numpy as np
import torch

from . import _native
from ._device import stream_handle
from ._native import D as Dbl
from ._native import I, P, signature
from .grid import sound3_stage_schedule_f32, terrain_coordinate_profiles

signature("gm1hip_dry_stage_forward", P, P, I, I, I, I, I, Dbl, P)
signature("gm1hip_dry_stage_adjoint", P, P, I, I, I, I, I, Dbl, P)

F32 = np.float32
STATE = ("u", "v", "w", "th", "pp", "ppx")

# pointer-table slots (the order of adjoint_step.hip's Slot enum)
SLOTS = (
    "COL", "XU", "XV", "XW", "XTH", "XPP", "AU", "AV", "AW", "ATH", "API", "PPX",
    "THV", "PPTERM", "PIADV", "DUM8", "THTERM", "RRU", "RRV", "RRW", "DIVX", "UTEN", "VTEN", "WTEN",
    "QK", "PK1", "PK2", "MM", "FPK", "EC", "ED", "SU", "SV", "SW", "SPP", "SPPD", "ARU", "ARV", "ARW",
    "PPXO", "DIV", "TDIV", "DIVX2", "THO",
    "B_OU", "B_OV", "B_OW", "B_OTH", "B_OPP", "B_PPX", "B_XU", "B_XV", "B_XW", "B_XTH", "B_XPP",
    "B_AU", "B_AV", "B_AW", "B_ATH", "B_API",
    "B_THV", "B_PPTERM", "B_PIADV", "B_DUM8", "B_THTERM", "B_RRU", "B_RRV", "B_RRW", "B_DIVX", "B_UTEN",
    "B_VTEN", "B_WTEN", "B_QK", "B_PK1", "B_PK2", "B_MM", "B_U", "B_V", "B_W", "B_PP", "B_PPD", "B_ARU", "B_ARV",
    "B_ARW", "B_DIV", "B_XA", "B_TOT", "B_MU", "B_MV", "B_MW", "B_DIVX2", "B_SC", "B_EC", "B_ED", "B_FPKF",
    "RA", "RB", "RC", "RD", "G0", "G1", "G2", "G3", "G4", "G5",
)
W_SLOTS = {"XW", "AW", "RRW", "WTEN", "ARW", "B_OW", "B_XW", "B_AW", "B_RRW", "B_WTEN", "B_W", "B_ARW", "B_MW",
           "G0", "G1", "G2", "G3", "G4", "G5"}
_RD, _RV, _CP, _CPV, _G = F32(287.04), F32(461.5), F32(1005.7), F32(1870.0), F32(9.81)
THETA_WEPS = float(F32(10.0) * F32(1.0e-18))


def stage_consts(dt: float, nsound: int, nrk: int, dx: float, dy: float, dz: float,
                 weno: str = "exact") -> tuple[np.ndarray, int]:
    """The binary32 scalars of one stage (adjoint_step.hip's Consts) and its substep count.

    ``weno`` "exact" differentiates the WENO5 weights and limiters in the
    adjoint; "frozen" holds them at their forward values."""
    if weno not in ("exact", "frozen"):
        raise ValueError("weno must be 'exact' or 'frozen'")
    nloop, dts = sound3_stage_schedule_f32(dt, nsound, nrk)
    dts = F32(dts)
    rdx, rdy, rdz = F32(1.0) / F32(dx), F32(1.0) / F32(dy), F32(1.0) / F32(dz)
    alph, kdiv = F32(0.60), F32(0.10)
    rovcp = _RD / _CP
    cpm = _CP + _CPV * F32(0.0)
    vals = [rdx, rdy, rdz, dts, F32(F32(dt) / F32(4 - nrk)), alph, F32(1.0) - alph, kdiv, _CP * F32(0.5), _CP,
            F32(1.0) / F32(nloop), (dts * F32(0.5)) * rdx, (dts * F32(0.5)) * rdy,
            _G, rovcp, _RV / _RD, _RV / _RD - F32(1.0), cpm, F32(1.0) / ((_CP - _RD) + F32(0.0)),
            (_RD + _RV * F32(0.0)) - rovcp * cpm, alph * alph, F32(THETA_WEPS), F32(1.0 if weno == "exact" else 0.0)]
    return np.array(vals, F32), nloop


def pack_columns(metrics: dict[str, Any], base: dict[str, Any], nk: int, dz: float) -> np.ndarray:
    """rho rr0 rf0 rrf0 c1 c2 th0 pi0 qv0 rth rds rdsf of one uniform column, packed for the kernels."""
    col = lambda n: np.asarray(metrics[n], F32)[:, 0, 0]  # noqa: E731
    prof = terrain_coordinate_profiles(nk, dz)
    th0 = np.asarray(base["th0"], F32)
    parts = [col("rho0_3d"), col("rr0"), col("rf0"), col("rrf0"), col("c1"), col("c2"), th0,
             np.asarray(base["pi0"], F32), np.asarray(base["qv0"], F32), F32(1.0) / th0, prof.rds, prof.rdsf]
    out = np.concatenate(parts).astype(F32)
    assert out.size == 12 * nk + 8
    return out


def _require_flat_periodic(cfg, metrics) -> None:
    if tuple(cfg.lateral_sides) != (1, 1, 1, 1) or cfg.moist or cfg.stretched is not None:
        raise ValueError("the halo-free step is the dry, uniform-column, doubly periodic case")
    for name, val in (("gz", 1.0), ("rgzu", 1.0), ("rgzv", 1.0), ("dzdx", 0.0), ("dzdy", 0.0)):
        if not np.all(np.asarray(metrics[name]) == val):
            raise ValueError(f"the halo-free step needs flat metrics ({name} != {val})")
    rho = np.asarray(metrics["rho0_3d"])
    if not np.all(rho == rho[:, :1, :1]):
        raise ValueError("the halo-free step needs a horizontally uniform base state")


class DryPeriodicStep:
    """Workspaces and launches of the halo-free dry periodic step (one device, one stream).

    Each RK stage has a workspace of its own, chained without copies: stage
    2 reads stage 1's final substep planes as its input state, stage 3 reads
    stage 2's, the three share the step-start state and ppx; on the way
    back, stage 2 reads stage 3's input adjoints as its output adjoints.  A
    large step's adjoint therefore reruns its three forward stages once and
    keeps every stage's intermediates for the transposed pass.
    """

    SHARED = ("AU", "AV", "AW", "ATH", "API", "PPX", "COL", "B_AU", "B_AV", "B_AW", "B_ATH", "B_API", "B_PPX")

    def __init__(self, cfg, metrics, base, device="cuda", stream: Any | None = None, weno: str = "exact"):
        _require_flat_periodic(cfg, metrics)
        self.weno = weno
        self.cfg = cfg
        self.ni, self.nj, self.nk = cfg.ni, cfg.nj, cfg.nk
        if min(self.ni, self.nj, self.nk) < 6:
            raise ValueError("ni, nj and nk must be at least 6")
        self.dev = torch.device(device)
        self.stream = stream
        self.consts = {}
        self.nloops = {}
        for nrk in (1, 2, 3):
            self.consts[nrk], self.nloops[nrk] = stage_consts(cfg.dt, cfg.nsound, nrk, cfg.dx, cfg.dy, cfg.dz,
                                                              weno)
        ni, nj, nk = self.ni, self.nj, self.nk
        self.shape = (nk, nj, ni)
        self.wshape = (nk + 1, nj, ni)
        z = lambda shape: torch.zeros(shape, dtype=torch.float32, device=self.dev)  # noqa: E731
        ws = {}
        with self._on_stream():
            shared = {"COL": torch.from_numpy(pack_columns(metrics, base, nk, cfg.dz)).to(self.dev)}
            for sl in self.SHARED:
                if sl != "COL":
                    shared[sl] = z(self.wshape if sl in W_SLOTS else self.shape)
            for nrk in (1, 2, 3):
                nl = self.nloops[nrk]
                t = dict(shared)
                for sl in SLOTS:
                    if sl in t:
                        continue
                    if nrk == 1 and sl in ("XU", "XV", "XW", "XTH", "XPP"):
                        t[sl] = shared["A" + sl[1:]] if sl != "XPP" else shared["API"]
                    elif nrk > 1 and sl in ("XU", "XV", "XW", "XTH", "XPP"):
                        prev, pn = ws[nrk - 1], self.nloops[nrk - 1]
                        t[sl] = {"XU": prev["SU"][pn], "XV": prev["SV"][pn], "XW": prev["SW"][pn],
                                 "XTH": prev["THO"], "XPP": prev["SPP"][pn]}[sl]
                    elif sl in ("SU", "SV", "SPP", "SPPD"):
                        t[sl] = z((nl + 1,) + self.shape)
                    elif sl == "SW":
                        t[sl] = z((nl + 1,) + self.wshape)
                    else:
                        t[sl] = z(self.wshape if sl in W_SLOTS else self.shape)
                ws[nrk] = t
            for nrk in (2, 1):  # output adjoints of stage nrk are the input adjoints of stage nrk + 1
                for o, x in (("B_OU", "B_XU"), ("B_OV", "B_XV"), ("B_OW", "B_XW"), ("B_OTH", "B_XTH"),
                             ("B_OPP", "B_XPP")):
                    ws[nrk][o] = ws[nrk + 1][x]
        self.ws = ws
        self.tables = {nrk: (ctypes.c_void_p * len(SLOTS))(*(ws[nrk][sl].data_ptr() for sl in SLOTS))
                       for nrk in (1, 2, 3)}

    def _on_stream(self):
        return torch.cuda.stream(self.stream) if self.stream is not None else torch.cuda.stream(
            torch.cuda.current_stream(self.dev))

    def _h(self):
        return stream_handle(self.stream)

    def _load(self, nrk, pairs):
        with self._on_stream():
            for slot, x in pairs:
                dst = self.ws[nrk][slot]
                if dst.data_ptr() != x.data_ptr():
                    dst.copy_(x)

    def _launch(self, name: str, nrk: int) -> None:
        c = self.consts[nrk]
        _native.call(name, ctypes.cast(self.tables[nrk], ctypes.c_void_p), c.ctypes.data_as(ctypes.c_void_p),
                     self.ni, self.nj, self.nk, nrk, self.nloops[nrk], THETA_WEPS, self._h())

    def _stage_out(self, nrk: int) -> tuple[dict, torch.Tensor]:
        t, n = self.ws[nrk], self.nloops[nrk]
        return ({"u": t["SU"][n], "v": t["SV"][n], "w": t["SW"][n], "th": t["THO"], "pp": t["SPP"][n]},
                t["PPXO"] if nrk == 3 else t["PPX"])

    # ------------------------------------------------------------ forward
    def load_start(self, S: dict) -> None:
        """The step-start state (u v w th pp ppx), shared by the three stages."""
        self._load(1, [("AU", S["u"]), ("AV", S["v"]), ("AW", S["w"]), ("ATH", S["th"]), ("API", S["pp"]),
                       ("PPX", S["ppx"])])

    def forward_stage(self, nrk: int, X: dict, A: dict, ppx: torch.Tensor) -> tuple[dict, torch.Tensor]:
        """One RK stage from X (u v w th pp), the step-start state A and ppx; returns copies of the new X and
        ppx (at nrk = 3 solve3's pp + ppx).  The stage's workspace keeps its intermediates."""
        self._load(nrk, [("AU", A["u"]), ("AV", A["v"]), ("AW", A["w"]), ("ATH", A["th"]), ("API", A["pp"]),
                         ("PPX", ppx)])
        self._load(nrk, [("XU", X["u"]), ("XV", X["v"]), ("XW", X["w"]), ("XTH", X["th"]), ("XPP", X["pp"])])
        self._launch("gm1hip_dry_stage_forward", nrk)
        Xn, p = self._stage_out(nrk)
        with self._on_stream():
            return {k: v.clone() for k, v in Xn.items()}, p.clone()

    def _forward_chain(self, S: dict) -> None:
        self.load_start(S)
        for nrk in (1, 2, 3):
            self._launch("gm1hip_dry_stage_forward", nrk)

    def forward(self, S: dict) -> dict:
        """One large step of the state S (u v w th pp ppx); returns the new state (fresh tensors)."""
        self._forward_chain(S)
        Xn, p = self._stage_out(3)
        with self._on_stream():
            return {**{k: v.clone() for k, v in Xn.items()}, "ppx": p.clone()}

    # ------------------------------------------------------------ adjoint
    def adjoint_stage(self, nrk: int, X: dict, A: dict, ppx: torch.Tensor, Xb: dict, ppxb: torch.Tensor) -> tuple[
            dict, torch.Tensor]:
        """Adjoints of one stage's inputs (X, ppx) from those of its outputs; the adjoint of A accumulates
        (``zero_start_adjoint`` / ``start_adjoint``)."""
        self.forward_stage(nrk, X, A, ppx)
        self._load(nrk, [("B_OU", Xb["u"]), ("B_OV", Xb["v"]), ("B_OW", Xb["w"]), ("B_OTH", Xb["th"]),
                         ("B_OPP", Xb["pp"]), ("B_PPX", ppxb)])
        self._launch("gm1hip_dry_stage_adjoint", nrk)
        t = self.ws[nrk]
        with self._on_stream():
            out = {"u": t["B_XU"].clone(), "v": t["B_XV"].clone(), "w": t["B_XW"].clone(), "th": t["B_XTH"].clone(),
                   "pp": t["B_XPP"].clone()}
            return out, t["B_PPX"].clone()

    def zero_start_adjoint(self) -> None:
        with self._on_stream():
            for sl in ("B_AU", "B_AV", "B_AW", "B_ATH", "B_API"):
                self.ws[1][sl].zero_()

    def start_adjoint(self) -> dict:
        t = self.ws[1]
        with self._on_stream():
            return {"u": t["B_AU"].clone(), "v": t["B_AV"].clone(), "w": t["B_AW"].clone(), "th": t["B_ATH"].clone(),
                    "pp": t["B_API"].clone()}

    def adjoint(self, S: dict, Sb: dict) -> dict:
        """Adjoint of the state before one large step (S) from the adjoint Sb of the state after it."""
        self._forward_chain(S)
        self.zero_start_adjoint()
        self._load(3, [("B_OU", Sb["u"]), ("B_OV", Sb["v"]), ("B_OW", Sb["w"]), ("B_OTH", Sb["th"]),
                       ("B_OPP", Sb["pp"]), ("B_PPX", Sb["ppx"])])
        for nrk in (3, 2, 1):
            self._launch("gm1hip_dry_stage_adjoint", nrk)
        t = self.ws[1]
        with self._on_stream():
            out = {n: t["B_X" + k] + t["B_A" + a] for n, k, a in (("u", "U", "U"), ("v", "V", "V"), ("w", "W", "W"),
                                                                  ("th", "TH", "TH"), ("pp", "PP", "PI"))}
            out["ppx"] = t["B_PPX"].clone()
        return out

    def empty_state(self) -> dict:
        with self._on_stream():
            return {n: torch.zeros(self.wshape if n == "w" else self.shape, dtype=torch.float32, device=self.dev)
                    for n in STATE}


def revolve_schedule(nsteps: int, snapshots: int) -> tuple[dict, int]:
    """Binomial checkpointing: the split table of the cheapest recursive schedule and its forward-step count.

    A window of n steps starting from a held state with s free checkpoint
    slots is reversed by advancing m steps, storing that state, reversing the
    last n - m steps with s - 1 slots and then the first m with s slots; the
    table gives the m of least recomputation (Griewank's binomial bound is
    reached), with cost(1, s) = 0 and, without free slots, m = n - 1.
    """
    from functools import lru_cache

    @lru_cache(maxsize=None)
    def cost(n: int, s: int) -> tuple[int, int]:
        if n <= 1:
            return 0, 0
        if s <= 0:
            return n * (n - 1) // 2, n - 1
        best, arg = None, 1
        for m in range(1, n):
            c = m + cost(n - m, s - 1)[0] + cost(m, s)[0]
            if best is None or c < best:
                best, arg = c, m
        return best, arg

    table = {}
    for n in range(2, nsteps + 1):
        for s in range(0, snapshots + 1):
            table[(n, s)] = cost(n, s)[1]
    return table, cost(nsteps, snapshots)[0] if nsteps > 1 else 0


def smooth_direction(rng: np.random.Generator, shape: tuple, modes: int = 3) -> np.ndarray:
    """A random periodic perturbation of unit max-norm built from the lowest horizontal and vertical modes
    (wavenumbers 1..modes), the kind of direction a finite-difference check of the gradient resolves."""
    nk, nj, ni = shape
    k = np.arange(nk)[:, None, None] + 0.5
    j = np.arange(nj)[None, :, None]
    i = np.arange(ni)[None, None, :]
    out = np.zeros(shape)
    for _ in range(2 * modes):
        kx, ky, kz = rng.integers(1, modes + 1, size=3)
        ph = rng.uniform(0.0, 2.0 * np.pi, size=3)
        out += rng.standard_normal() * (np.cos(2 * np.pi * kx * i / ni + ph[0]) * np.cos(2 * np.pi * ky * j / nj + ph[1])
                                        * np.sin(np.pi * kz * k / nk + ph[2]))
    return out / np.abs(out).max()


def fd_directions(rng: np.random.Generator, fields: Mapping[str, Any], gradient: Mapping[str, Any], device: Any,
                  count: int = 8, pool: int = 16) -> tuple[list, list]:
    """Directions for a central-difference check of ``gradient``: ``count`` smooth random directions of
    (theta', u, v), each field's scaled by its max-norm in ``fields``, with the gradient's directional derivative
    g.d.  Of ``pool`` draws, one whose |g.d| is below 5 % of the pool's median is nearly orthogonal to the
    gradient -- a derivative of cancelling terms that binary32 central differences cannot resolve -- and gives
    way to the next draw.  Returns ([(d, g.d)] accepted, [(draw index, g.d)] replaced)."""
    draws = []
    for _ in range(pool):
        d = {n: torch.from_numpy((smooth_direction(rng, tuple(fields[n].shape))
                                  * float(np.abs(np.asarray(fields[n])).max())).astype(np.float32)).to(device)
             for n in ("th", "u", "v")}
        draws.append((d, sum(float(torch.sum(gradient[n].double() * d[n].double())) for n in d)))
    floor = 0.05 * float(np.median([abs(ad) for _, ad in draws]))
    kept, replaced = [], []
    for k, (d, ad) in enumerate(draws):
        if len(kept) == count:
            break
        if abs(ad) < floor:
            replaced.append((k, ad))
        else:
            kept.append((d, ad))
    if len(kept) < count:
        raise ValueError("too many directions nearly orthogonal to the gradient")
    return kept, replaced


def storm_intensity_weights(cfg, metrics, zlo: float = 2000.0, zhi: float = 8000.0) -> np.ndarray:
    """Per-level weights of the storm-intensity loss: rf0 on the w levels with height in [zlo, zhi], normalised
    so that J = sum(weight * w^2) is the base-density-weighted mean of w^2 over that layer."""
    rf0 = np.asarray(metrics["rf0"], np.float64)[:, 0, 0]
    z = np.arange(cfg.nk + 1) * float(cfg.dz)
    wt = np.where((z >= zlo) & (z <= zhi), rf0, 0.0)
    return (wt / (wt.sum() * cfg.ni * cfg.nj)).astype(np.float32)


class WindowGradient:
    """Gradient of the storm-intensity loss after ``nsteps`` large steps with respect to the initial state."""

    def __init__(self, step: DryPeriodicStep, nsteps: int, snapshots: int = 8, zlo: float = 2000.0,
                 zhi: float = 8000.0, metrics=None):
        self.step = step
        self.nsteps = nsteps
        self.snapshots = max(0, min(snapshots, nsteps))
        self.table, self.recompute_steps = revolve_schedule(nsteps, self.snapshots)
        m = metrics
        wt = storm_intensity_weights(step.cfg, m, zlo, zhi)
        with step._on_stream():
            self.weights = torch.from_numpy(wt).to(step.dev)[:, None, None]
        self.forward_steps = 0
        self.adjoint_steps = 0

    def loss(self, S: dict) -> torch.Tensor:
        with self.step._on_stream():
            return torch.sum(self.weights * S["w"] * S["w"])

    def loss_adjoint(self, S: dict) -> dict:
        Sb = self.step.empty_state()
        with self.step._on_stream():
            Sb["w"] = 2.0 * self.weights * S["w"]
        return Sb

    def advance(self, S: dict, n: int) -> dict:
        for _ in range(n):
            S = self.step.forward(S)
            self.forward_steps += 1
        return S

    def value(self, S0: dict) -> torch.Tensor:
        return self.loss(self.advance(S0, self.nsteps))

    def _reverse(self, S: dict, n: int, s: int, Sb_end: dict | None):
        """(adjoint at S, loss) for the n steps from S; ``Sb_end`` None means the loss is taken at the end."""
        if n == 1:
            J = None
            if Sb_end is None:
                end = self.step.forward(S)
                self.forward_steps += 1
                J = self.loss(end)
                Sb_end = self.loss_adjoint(end)
            self.adjoint_steps += 1
            return self.step.adjoint(S, Sb_end), J
        m = self.table[(n, max(s, 0))]
        mid = self.advance(S, m)
        Sb_mid, J = self._reverse(mid, n - m, s - 1, Sb_end)
        del mid
        Sb, _ = self._reverse(S, m, s, Sb_mid)
        return Sb, J

    def gradient(self, S0: dict) -> tuple[torch.Tensor, dict]:
        """(J, adjoint of the initial state) by binomial checkpointing."""
        self.forward_steps = 0
        self.adjoint_steps = 0
        Sb, J = self._reverse(S0, self.nsteps, self.snapshots, None)
        return J, Sb


class StormIntensity(torch.autograd.Function):
    """J(theta', u, v) after a window of large steps, differentiable by the device adjoint.

    ``apply(window, base_state, th, u, v)``: ``window`` a WindowGradient,
    ``base_state`` the initial halo-free state whose w, pp and ppx are held
    fixed; returns the scalar loss.
    """

    @staticmethod
    def forward(ctx, window, base_state, th, u, v):
        S0 = {**base_state, "th": th.detach(), "u": u.detach(), "v": v.detach()}
        ctx.window = window
        ctx.S0 = S0
        return window.value(S0)

    @staticmethod
    def backward(ctx, gJ):
        _, Sb = ctx.window.gradient(ctx.S0)
        return None, None, gJ * Sb["th"], gJ * Sb["u"], gJ * Sb["v"]


def from_padded(a: dict, ni: int, nj: int, nk: int, ng: int = 3) -> dict:
    """Halo-free state from a padded large-step state (torch tensors or arrays)."""
    return {"u": a["u3d"][1 : nk + 1, ng : ng + nj, ng : ng + ni],
            "v": a["v3d"][1 : nk + 1, ng : ng + nj, ng : ng + ni],
            "w": a["w3d"][1 : nk + 2, ng : ng + nj, ng : ng + ni],
            "th": a["th3d"],
            "pp": a["pp3d"][1 : nk + 1, 1 : nj + 1, 1 : ni + 1],
            "ppx": a["ppx"][1 : nk + 1, 1 : nj + 1, 1 : ni + 1]}


__all__ = ["DryPeriodicStep", "STATE", "StormIntensity", "WindowGradient", "from_padded", "pack_columns",
           "revolve_schedule", "smooth_direction", "stage_consts", "storm_intensity_weights"]
