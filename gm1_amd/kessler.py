"""Kessler warm rain and its large-step transition on gfx950.

Host mirror of ``cuda_gm1_kessler_driver.py`` (``Gm1KesslerDriverSelection``
:100-160, ``gm1_kessler_driver_shapes`` :180-212,
``create_cuda_gm1_kessler_driver_workspace`` :806,
``launch_cuda_gm1_kessler_driver`` :907-1087), of the transition
``CudaGm1KesslerLargeStepTransition`` (cuda_gm1_kessler_large_step.py:295-560)
and of the dry-air mass correction (cuda_apmasscon.py).  The device work is
three C-ABI calls; budgets (asq, qbudget) and the satadj status words stay in
device buffers.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Mapping

import torch

from . import _native
from ._device import no_alias, ptr, stream_handle
from ._native import D, Fl, I, P, signature
from .errors import refuse

I64 = ctypes.c_int64
signature("gm1hip_kessler_scratch_bytes", I, I, I, restype=I64)
signature("gm1hip_dry_mass_scratch_bytes", I, I, I, restype=I64)
signature("gm1hip_kessler_driver", *([P] * 25), I64, I, I, I, Fl, Fl, Fl, Fl, I, P)
signature("gm1hip_kessler_driver_decisions", *([P] * 26), I64, I, I, I, Fl, Fl, Fl, Fl, I, P)
signature("gm1hip_kessler_driver_noalloc", *([P] * 24), I, I, I, Fl, Fl, Fl, Fl, I, P)
signature("gm1hip_kessler_driver_efall", *([P] * 28), I64, I, I, I, Fl, Fl, Fl, Fl, I, P)

# Size of the no-allocation driver's device workspace (a file-scope array in the library).
KESSLER_NOALLOC_MAX_BYTES = 64 << 20
signature("gm1hip_kessler_transition_pre", *([P] * 14), I, I, I, Fl, I, I, P)
signature("gm1hip_apmasscon", *([P] * 13), D, P, I64, I, I, I, P)
signature("gm1hip_kessler_transition_pre_apmasscon", *([P] * 19), D, P, I64, I, I, I, Fl, I, P)


@dataclass(frozen=True)
class Gm1KesslerDriverSelection:
    """The sealed native branch (cuda_gm1_kessler_driver.py:100-160), efall 0 or 1."""

    ptype: int = 1
    equation_set: int = 1
    testcase: int = 0
    pdscheme: int = 1
    efall: int = 0
    axisymmetric: int = 0
    moving_domain: int = 0
    rain_bins: int = 1
    nrk: int = 4
    input_phase: str = "pre-pdefq"
    output_phase: str = "post-satadj"


DEFAULT_GM1_KESSLER_DRIVER_SELECTION = Gm1KesslerDriverSelection()


def validate_selection(selection: Gm1KesslerDriverSelection) -> None:
    if not isinstance(selection, Gm1KesslerDriverSelection):
        raise TypeError("selection must be a Gm1KesslerDriverSelection")
    fixed = {"ptype": 1, "testcase": 0, "pdscheme": 1, "axisymmetric": 0, "moving_domain": 0, "rain_bins": 1,
             "nrk": 4, "input_phase": "pre-pdefq", "output_phase": "post-satadj"}
    for name, wanted in fixed.items():
        value = getattr(selection, name)
        if type(value) is not type(wanted) or value != wanted:
            refuse(code="GM1-KESSLER-DRIVER-SELECTION", gate="physics-seam", option=name, value=value,
                   reason="the complete Kessler operator is sealed to one exact native branch",
                   remedy=f"use {name}={wanted!r}")
    if type(selection.efall) is not int or selection.efall not in (0, 1):
        refuse(code="GM1-KESSLER-DRIVER-EFALL", gate="physics-seam", option="efall", value=selection.efall,
               reason="the authority admits exactly the efall=0 and efall=1 branches",
               remedy="use efall=0 or efall=1")
    if type(selection.equation_set) is not int or selection.equation_set not in (1, 2):
        refuse(code="GM1-KESSLER-DRIVER-EQTSET", gate="physics-seam", option="equation_set",
               value=selection.equation_set,
               reason="the authority has exactly traditional and energy-conserving branches",
               remedy="use equation_set=1 or equation_set=2")


def gm1_kessler_driver_shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    mass = (nk, nj, ni)
    out = {n: mass for n in ("rmh", "mh", "pi0", "th0", "rho", "rr", "pp3d", "th3d", "prs", "qv", "qc", "qr",
                             "fall_velocity")}
    out.update({"ruh": (ni,), "rvh": (nj,), "rain": (1, nj, ni), "prate": (nj, ni), "asq": (3,), "qbudget": (10,)})
    return out


@dataclass
class Gm1KesslerDriverWorkspace:
    scratch: torch.Tensor
    status: torch.Tensor  # int32 [max satadj iterations, non-converged cells]
    ni: int
    nj: int
    nk: int
    temperature: torch.Tensor | None = None  # efall=1: the held temperature of the last call


def create_cuda_gm1_kessler_driver_workspace(ni: int, nj: int, nk: int, device: Any = "cuda") -> Gm1KesslerDriverWorkspace:
    nbytes = int(_native.query("gm1hip_kessler_scratch_bytes", ni, nj, nk))
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=device)
    status = torch.zeros(2, dtype=torch.int32, device=device)
    return Gm1KesslerDriverWorkspace(scratch, status, ni, nj, nk)


def _check(arrays: Mapping[str, Any], ni: int, nj: int, nk: int, efall: int = 0) -> dict[str, torch.Tensor]:
    out = {}
    shapes = gm1_kessler_driver_shapes(ni, nj, nk)
    if efall:
        if nk < 2:
            raise ValueError("efall=1 requires nk >= 2")
        shapes["mf"] = (nk + 1, nj, ni)
    for name, shape in shapes.items():
        if name not in arrays:
            raise ValueError(f"missing GM1 Kessler array: {name}")
        t = arrays[name]
        dtype = torch.float64 if name in ("asq", "qbudget") else torch.float32
        if not isinstance(t, torch.Tensor) or t.dtype != dtype or tuple(t.shape) != shape or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous {dtype} tensor of shape {shape}")
        if not t.is_cuda:
            raise ValueError(f"{name} must be on the HIP device")
        out[name] = t
    return out


def launch_cuda_gm1_kessler_driver(
    arrays: Mapping[str, Any], *, ni: int, nj: int, nk: int, dx: float, dy: float, dz: float, dt: float,
    selection: Gm1KesslerDriverSelection = DEFAULT_GM1_KESSLER_DRIVER_SELECTION, stream: Any | None = None,
    workspace: Gm1KesslerDriverWorkspace | None = None, held: Mapping[str, Any] | None = None,
    decisions: torch.Tensor | None = None, temperature: torch.Tensor | None = None, **_: Any,
) -> Gm1KesslerDriverWorkspace:
    """Every ptype=1 phase (pdefq, fall speed, [efall,] fallout, local, satadj) in place.

    ``selection.efall == 1`` needs ``arrays["mf"]`` (nk+1, nj, ni) and runs the
    energy fallout before the mass fallout (cuda_gm1_kessler_driver.py:482-551);
    the held temperature it leaves lands in ``temperature`` (float32 (nk, nj,
    ni), allocated when not given) and is kept on the workspace.

    With ``held`` (qpten/qtten/qvten/qcten holding the snapshot of the
    eqtset=2 transition) the held rates are recaptured in the same pass.
    ``decisions`` (int32 (nk, nj, ni), optional) receives each cell's branch
    word: bits 0..6 the local-conversion branches (qc > 0, autoconversion,
    accretion, evaporation, evaporation capped at saturation, the ar + cr > qc
    limiter, the er > 1e-7 update), bit 7 the saturation adjustment ran, bits
    8..15 its iteration count, bit 16 it did not converge.
    """
    validate_selection(selection)
    a = _check(arrays, ni, nj, nk, selection.efall)
    if decisions is not None and (not isinstance(decisions, torch.Tensor) or decisions.dtype != torch.int32
                                  or tuple(decisions.shape) != (nk, nj, ni) or not decisions.is_contiguous()
                                  or decisions.device != a["rho"].device):
        raise ValueError(f"decisions must be a contiguous int32 tensor of shape {(nk, nj, ni)} on the device")
    if workspace is None:
        workspace = create_cuda_gm1_kessler_driver_workspace(ni, nj, nk, a["rho"].device)
    h = held or {}
    if selection.efall == 1:
        if temperature is None:
            temperature = workspace.temperature
        if temperature is None or tuple(temperature.shape) != (nk, nj, ni):
            temperature = torch.empty((nk, nj, ni), dtype=torch.float32, device=a["rho"].device)
        if (temperature.dtype != torch.float32 or not temperature.is_contiguous()
                or temperature.device != a["rho"].device):
            raise ValueError(f"temperature must be a contiguous float32 tensor of shape {(nk, nj, ni)}")
        no_alias({**a, "temperature": temperature}, "Kessler arrays and the held temperature must not alias")
        workspace.temperature = temperature
        _native.call(
            "gm1hip_kessler_driver_efall",
            *(ptr(a[n]) for n in ("ruh", "rvh", "rmh", "mh", "mf", "pi0", "th0", "rho", "rr", "pp3d", "th3d", "prs",
                                  "qv", "qc", "qr", "fall_velocity", "rain", "prate")),
            ptr(temperature), ptr(a["asq"]), ptr(a["qbudget"]),
            ptr(workspace.status), ptr(h.get("qpten")), ptr(h.get("qtten")), ptr(h.get("qvten")),
            ptr(h.get("qcten")), ptr(decisions), ptr(workspace.scratch), workspace.scratch.numel(), ni, nj, nk,
            float(dx), float(dy), float(dz), float(dt), selection.equation_set, stream_handle(stream))
        return workspace
    _native.call(
        "gm1hip_kessler_driver_decisions",
        *(ptr(a[n]) for n in ("ruh", "rvh", "rmh", "mh", "pi0", "th0", "rho", "rr", "pp3d", "th3d", "prs", "qv",
                              "qc", "qr", "fall_velocity", "rain", "prate", "asq", "qbudget")),
        ptr(workspace.status), ptr(h.get("qpten")), ptr(h.get("qtten")), ptr(h.get("qvten")), ptr(h.get("qcten")),
        ptr(decisions), ptr(workspace.scratch), workspace.scratch.numel(), ni, nj, nk, float(dx), float(dy),
        float(dz), float(dt), selection.equation_set, stream_handle(stream))
    return workspace


def launch_cuda_gm1_kessler_driver_noalloc(
    arrays: Mapping[str, Any], status: torch.Tensor, *, ni: int, nj: int, nk: int, dx: float, dy: float, dz: float,
    dt: float, selection: Gm1KesslerDriverSelection = DEFAULT_GM1_KESSLER_DRIVER_SELECTION, stream: Any | None = None,
    held: Mapping[str, Any] | None = None,
) -> None:
    """``launch_cuda_gm1_kessler_driver`` with no workspace: nothing is allocated.

    The driver's scratch is the library's file-scope device array
    (``KESSLER_NOALLOC_MAX_BYTES``); ``status`` (int32 [2]) is the caller's.
    Calls that overlap on one GPU share that array and race.
    """
    validate_selection(selection)
    if selection.efall != 0:
        refuse(code="GM1-KESSLER-NOALLOC-EFALL", gate="physics-seam", option="efall", value=selection.efall,
               reason="the no-allocation driver runs the efall=0 branch (efall=1 needs a held-temperature buffer)",
               remedy="use launch_cuda_gm1_kessler_driver for efall=1")
    a = _check(arrays, ni, nj, nk)
    need = int(_native.query("gm1hip_kessler_scratch_bytes", ni, nj, nk))
    if need > KESSLER_NOALLOC_MAX_BYTES:
        raise ValueError(f"a {ni}x{nj}x{nk} Kessler driver needs {need} bytes of scratch; the no-allocation "
                         f"driver holds {KESSLER_NOALLOC_MAX_BYTES} (use launch_cuda_gm1_kessler_driver)")
    if (not isinstance(status, torch.Tensor) or status.dtype != torch.int32 or tuple(status.shape) != (2,)
            or status.device != a["rho"].device):
        raise ValueError("status must be an int32 tensor of shape (2,) on the device")
    h = held or {}
    _native.call(
        "gm1hip_kessler_driver_noalloc",
        *(ptr(a[n]) for n in ("ruh", "rvh", "rmh", "mh", "pi0", "th0", "rho", "rr", "pp3d", "th3d", "prs", "qv",
                              "qc", "qr", "fall_velocity", "rain", "prate", "asq", "qbudget")),
        ptr(status), ptr(h.get("qpten")), ptr(h.get("qtten")), ptr(h.get("qvten")), ptr(h.get("qcten")),
        ni, nj, nk, float(dx), float(dy), float(dz), float(dt), selection.equation_set, stream_handle(stream))


@dataclass
class DryMassWorkspace:
    scratch: torch.Tensor
    ledger: torch.Tensor  # float64 [mass before, sum(pi0+pp3d) before]
    correction: torch.Tensor  # float32 [1]


def create_dry_mass_workspace(ni: int, nj: int, nk: int, device: Any = "cuda") -> DryMassWorkspace:
    nbytes = int(_native.query("gm1hip_dry_mass_scratch_bytes", ni, nj, nk))
    return DryMassWorkspace(torch.empty(nbytes, dtype=torch.uint8, device=device),
                            torch.zeros(2, dtype=torch.float64, device=device),
                            torch.zeros(1, dtype=torch.float32, device=device))


def launch_cuda_apmasscon(arrays: Mapping[str, Any], *, ni: int, nj: int, nk: int, mass1: float,
                          workspace: DryMassWorkspace, stream: Any | None = None) -> DryMassWorkspace:
    """Exner correction restoring the dry-air mass ledger (cuda_apmasscon.py:256-327).

    ``mass1 <= 0`` only measures the ledger into ``workspace.ledger``.
    """
    a = arrays
    _native.call("gm1hip_apmasscon", *(ptr(a.get(n)) for n in ("ruh", "rvh", "rmh", "pi0", "th0", "th3d", "qv",
                                                                  "pp3d", "prs", "rho", "rr")),
                 ptr(workspace.ledger), ptr(workspace.correction), float(mass1), ptr(workspace.scratch),
                 workspace.scratch.numel(), ni, nj, nk, stream_handle(stream))
    return workspace


def launch_cuda_gm1_kessler_transition(
    arrays: Mapping[str, Any], held: Mapping[str, Any], *, ni: int, nj: int, nk: int, dx: float, dy: float,
    dz: float, dt: float, selection: Gm1KesslerDriverSelection, workspace: Gm1KesslerDriverWorkspace,
    pressure_family: Any | None = None, mass1: float | None = None, dry_mass: DryMassWorkspace | None = None,
    stream: Any | None = None,
) -> None:
    """cpu_gm1_kessler_large_step_inplace on the device (cuda_gm1_kessler_large_step.py:183-292).

    eqtset=2 held remove/snapshot and the pre-physics EOS (pp3d taken from
    ``pressure_family`` when given), the dry-mass hook when ``mass1`` is
    given, ``rr <- rho``, the driver, and the held recapture.
    """
    validate_selection(selection)
    a = _check(arrays, ni, nj, nk, selection.efall)
    eqt = selection.equation_set
    hook = mass1 is not None
    if hook and dry_mass is None:
        raise ValueError("the dry-mass hook needs a DryMassWorkspace")
    pre = (ptr(pressure_family), ptr(a["pp3d"]), ptr(a["th3d"]), ptr(a["qv"]), ptr(a["qc"]), ptr(held["qpten"]),
           ptr(held["qtten"]), ptr(held["qvten"]), ptr(held["qcten"]), ptr(a["pi0"]), ptr(a["th0"]), ptr(a["prs"]),
           ptr(a["rho"]), ptr(a["rr"]))
    if hook and float(mass1) > 0.0:
        # the transition's cells and the ledger's partials of the rho and pp3d they write in one pass
        _native.call("gm1hip_kessler_transition_pre_apmasscon", *pre, ptr(a["ruh"]), ptr(a["rvh"]), ptr(a["rmh"]),
                     ptr(dry_mass.ledger), ptr(dry_mass.correction), float(mass1), ptr(dry_mass.scratch),
                     dry_mass.scratch.numel(), ni, nj, nk, float(dt), eqt, stream_handle(stream))
    else:
        _native.call("gm1hip_kessler_transition_pre", *pre, ni, nj, nk, float(dt), eqt, int(not hook),
                     stream_handle(stream))
        if hook:
            launch_cuda_apmasscon(a, ni=ni, nj=nj, nk=nk, mass1=float(mass1), workspace=dry_mass, stream=stream)
    launch_cuda_gm1_kessler_driver(a, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz, dt=dt, selection=selection,
                                   stream=stream, workspace=workspace, held=held if eqt == 2 else None)


# ----------------------------------------------------------------------------- the persistent transition
# CudaGm1KesslerLargeStepTransition (cuda_gm1_kessler_large_step.py:295-560): one object per device
# state carries the eqtset=2 held rates across large steps and applies, after the final RK stage,
# the held remove/snapshot, the pre-physics equation of state, an optional hook (the dry-mass
# correction), rr <- rho, the ptype=1 driver and the held recapture.

GM1_KESSLER_LARGE_STEP_RESTART_SCHEMA = "gm1_amd.kessler-large-step-restart/v1"
_RESTART_NAMES = ("held.qpten", "held.qtten", "held.qvten", "held.qcten", "persistent.rain", "persistent.prate",
                  "persistent.asq", "persistent.qbudget")


@dataclass
class Gm1KesslerHeldCarriers:
    qpten: torch.Tensor
    qtten: torch.Tensor
    qvten: torch.Tensor
    qcten: torch.Tensor


@dataclass(frozen=True)
class Gm1KesslerLargeStepRestart:
    schema: str
    completed_transitions: int
    arrays: Mapping[str, Any]


@dataclass(frozen=True)
class Gm1KesslerLargeStepResult:
    ready_event: Any
    completed_transitions: int
    receipt: Mapping[str, Any]
    restart: Gm1KesslerLargeStepRestart
    driver_launch: Any


class CudaGm1KesslerLargeStepTransition:
    """Persistent source-ordered Kessler transition for one device state."""

    def __init__(self, *, arrays: dict[str, Any], carriers: Gm1KesslerHeldCarriers, ni: int, nj: int, nk: int,
                 dx: float, dy: float, dz: float, dt: float, selection: Gm1KesslerDriverSelection, stream: Any,
                 workspace: Gm1KesslerDriverWorkspace, completed_transitions: int = 0, **_: Any) -> None:
        self.arrays = arrays
        self.carriers = carriers
        self.ni, self.nj, self.nk = ni, nj, nk
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        self.selection = selection
        self.stream = stream
        self.workspace = workspace
        self.completed_transitions = completed_transitions
        self._last_receipt: dict[str, Any] | None = None

    def _restart_arrays(self) -> dict[str, Any]:
        c, a = self.carriers, self.arrays
        return {"held.qpten": c.qpten, "held.qtten": c.qtten, "held.qvten": c.qvten, "held.qcten": c.qcten,
                "persistent.rain": a["rain"], "persistent.prate": a["prate"], "persistent.asq": a["asq"],
                "persistent.qbudget": a["qbudget"]}

    def capture_restart(self) -> Gm1KesslerLargeStepRestart:
        values = self._restart_arrays()
        if tuple(values) != _RESTART_NAMES:
            raise RuntimeError("GM1 Kessler restart inventory drifted")
        return Gm1KesslerLargeStepRestart(GM1_KESSLER_LARGE_STEP_RESTART_SCHEMA, self.completed_transitions,
                                          dict(values))

    def clone_restart(self) -> Gm1KesslerLargeStepRestart:
        with torch.cuda.stream(self.stream) if self.stream is not None else _NullContext():
            values = {k: v.clone() for k, v in self._restart_arrays().items()}
        return Gm1KesslerLargeStepRestart(GM1_KESSLER_LARGE_STEP_RESTART_SCHEMA, self.completed_transitions, values)

    def receipt_dict(self) -> dict[str, Any]:
        if self._last_receipt is None:
            raise RuntimeError("GM1 Kessler large-step transition has not completed")
        return dict(self._last_receipt)

    def apply(self, *, final_rk: bool = True, refl_10cm_due: bool = False,
              pre_component_hook: Any | None = None) -> Gm1KesslerLargeStepResult:
        """cuda_gm1_kessler_large_step.py:380-555: held remove/snapshot, EOS, hook, rr <- rho, driver, recapture."""
        if type(final_rk) is not bool or not final_rk:
            refuse(code="GM1-KESSLER-LARGE-STEP-CADENCE", gate="physics-seam", option="final_rk", value=final_rk,
                   reason="CM1 r21.1 calls mp_driver once after the completed RK solve",
                   remedy="invoke the transition only after final RK stage 3")
        if type(refl_10cm_due) is not bool:
            raise TypeError("refl_10cm_due must be boolean")
        if refl_10cm_due:
            refuse(code="GM1-KESSLER-REFLECTIVITY-UNPORTED", gate="physics-seam", option="refl_10cm_due",
                   value=True, reason="CM1 does not admit the selected reflectivity consumer for ptype=1",
                   remedy="disable reflectivity or select an admitted microphysics consumer")
        if pre_component_hook is not None and not callable(pre_component_hook):
            raise TypeError("pre_component_hook must be callable or None")
        a = _check(self.arrays, self.ni, self.nj, self.nk)
        c = self.carriers
        eqt = self.selection.equation_set
        h = stream_handle(self.stream)
        _native.call("gm1hip_kessler_transition_pre", None, ptr(a["pp3d"]), ptr(a["th3d"]), ptr(a["qv"]),
                     ptr(a["qc"]), ptr(c.qpten), ptr(c.qtten), ptr(c.qvten), ptr(c.qcten), ptr(a["pi0"]),
                     ptr(a["th0"]), ptr(a["prs"]), ptr(a["rho"]), ptr(a["rr"]), self.ni, self.nj, self.nk,
                     float(self.dt), eqt, int(pre_component_hook is None), h)
        hook_result = None
        if pre_component_hook is not None:
            hook_result = pre_component_hook(self.carriers)
            with torch.cuda.stream(self.stream) if self.stream is not None else _NullContext():
                a["rr"].copy_(a["rho"])
        held = {"qpten": c.qpten, "qtten": c.qtten, "qvten": c.qvten, "qcten": c.qcten} if eqt == 2 else None
        driver = launch_cuda_gm1_kessler_driver(a, ni=self.ni, nj=self.nj, nk=self.nk, dx=self.dx, dy=self.dy,
                                                dz=self.dz, dt=self.dt, selection=self.selection,
                                                stream=self.stream, workspace=self.workspace, held=held)
        self.completed_transitions += 1
        restart = self.capture_restart()
        self._last_receipt = {
            "completed_transitions": self.completed_transitions,
            "configuration": {"ni": self.ni, "nj": self.nj, "nk": self.nk, "dx_m": self.dx, "dy_m": self.dy,
                              "dz_m": self.dz, "dt_s": self.dt, "ptype": 1, "equation_set": eqt,
                              "efall": self.selection.efall, "final_rk": True},
            "source_order": ["held-rate remove/snapshot" if eqt == 2 else "held-rate lifecycle inactive",
                             "pre-physics EOS",
                             "dry-mass correction hook" if pre_component_hook is not None else "dry-mass hook absent",
                             "mp_driver ptype=1", "held-rate recapture" if eqt == 2 else "held-rate recapture inactive"],
            "restart": {"inventory": list(restart.arrays), "persistent": ["rain", "prate", "asq", "qbudget"]},
            "hook_result": None if hook_result is None else repr(hook_result),
            "target_arch": "gfx950",
        }
        event = None
        if not torch.cuda.is_current_stream_capturing():
            event = torch.cuda.Event()
            event.record(self.stream if self.stream is not None else torch.cuda.current_stream())
        return Gm1KesslerLargeStepResult(event, self.completed_transitions, self._last_receipt, restart, driver)


def create_cuda_gm1_kessler_large_step_transition(
    arrays: Mapping[str, Any], *, ni: int, nj: int, nk: int, dx: float, dy: float, dz: float, dt: float,
    selection: Gm1KesslerDriverSelection, stream: Any | None = None, workspace: Gm1KesslerDriverWorkspace | None = None,
    restart: Gm1KesslerLargeStepRestart | None = None, **_: Any,
) -> CudaGm1KesslerLargeStepTransition:
    """Bind a Kessler-initialised device state (cuda_gm1_kessler_large_step.py:558-640).

    The held rates are allocated zero, or taken from ``restart`` (its arrays
    copied in place, its transition count carried on).
    """
    validate_selection(selection)
    for name, value, minimum in (("ni", ni, 1), ("nj", nj, 1), ("nk", nk, 2)):
        if type(value) is not int or value < minimum:
            raise ValueError(f"{name} must be an integer >= {minimum}")
    a = _check(arrays, ni, nj, nk, selection.efall)
    dev = a["rho"].device
    held = Gm1KesslerHeldCarriers(*(torch.zeros((nk, nj, ni), dtype=torch.float32, device=dev) for _ in range(4)))
    done = 0
    if restart is not None:
        if restart.schema != GM1_KESSLER_LARGE_STEP_RESTART_SCHEMA or tuple(restart.arrays) != _RESTART_NAMES:
            raise ValueError("the restart is not a GM1 Kessler large-step restart of this inventory")
        for name, dst in (("held.qpten", held.qpten), ("held.qtten", held.qtten), ("held.qvten", held.qvten),
                          ("held.qcten", held.qcten), ("persistent.rain", a["rain"]), ("persistent.prate", a["prate"]),
                          ("persistent.asq", a["asq"]), ("persistent.qbudget", a["qbudget"])):
            dst.copy_(restart.arrays[name])
        done = int(restart.completed_transitions)
    workspace = workspace or create_cuda_gm1_kessler_driver_workspace(ni, nj, nk, dev)
    return CudaGm1KesslerLargeStepTransition(arrays=dict(arrays), carriers=held, ni=ni, nj=nj, nk=nk, dx=float(dx),
                                             dy=float(dy), dz=float(dz), dt=float(dt), selection=selection,
                                             stream=stream, workspace=workspace, completed_transitions=done)


class _NullContext:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


__all__ = [
    "CudaGm1KesslerLargeStepTransition", "GM1_KESSLER_LARGE_STEP_RESTART_SCHEMA", "Gm1KesslerHeldCarriers",
    "Gm1KesslerLargeStepRestart", "Gm1KesslerLargeStepResult", "create_cuda_gm1_kessler_large_step_transition",
    "DEFAULT_GM1_KESSLER_DRIVER_SELECTION", "DryMassWorkspace", "Gm1KesslerDriverSelection",
    "Gm1KesslerDriverWorkspace", "create_cuda_gm1_kessler_driver_workspace", "create_dry_mass_workspace",
    "gm1_kessler_driver_shapes", "launch_cuda_apmasscon", "launch_cuda_gm1_kessler_driver",
    "launch_cuda_gm1_kessler_transition",
]
