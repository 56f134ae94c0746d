"""Theta and moisture RK stage on gfx950 (host mirror of cuda_scalar_moist_rk.py).

Per RK stage (``launch_cuda_scalar_moist_rk_stage``, :1041-1499) each scalar
is packed into a padded carrier with its lateral-family halos, advected, and
composed A-anchored: ``out = A + dttmp*((held - divterm) + advection)`` with
``|out| < 1e-30 -> 0``.  Here advection and compose are ONE LDS-tiled kernel
(``gm1hip_scalar_rk_stage``); the advection tendency never reaches HBM.  On
the final stage each moisture row instead takes the positive-definite chain
(gated WENO fluxes, directional tendencies, donor repair in x, y and z,
total) fused with its compose (``gm1hip_pdef_moisture_stage``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any, Mapping

import numpy as np
import torch

from . import _native
from ._device import NG, lateral_code, no_alias, ptr, record, require, shapes, stream_handle, terrain_static_shapes
from ._native import D, Fl, I, P, U, signature

signature("gm1hip_scalar_rk_stage", *([P] * 13), I, I, I, I, D, Fl, Fl, Fl, U, Fl, P)
signature("gm1hip_pdef_moisture_stage", *([P] * 17), I, I, I, I, D, Fl, Fl, Fl, U, Fl, P)
signature("gm1hip_pack_scalar", P, P, P, I, I, I, I, U, P)
signature("gm1hip_scalar_rk_stage_src", *([P] * 15), I, I, I, I, D, Fl, Fl, Fl, U, Fl, P)
signature("gm1hip_pdef_moisture_stage_src", *([P] * 18), I, I, I, I, D, Fl, Fl, Fl, U, Fl, P)

CM1_NRKMAX = 3
CM1_TH0R_F32 = 300.0
CM1_SMEPS_F32 = 1.0e-30
CM1_EPSILON_F32 = 1.0e-18
CM1_QMAG_F32 = (1.0e-2,) * 6
CM1_SCALAR_MOIST_SPECIES = ("qv", "qc", "qr", "qi", "qs", "qg")


def scalar_moist_rk_stage_timestep_f32(dt: float, nrk: int) -> float:
    """CM1's default-REAL dt/(nrkmax+1-nrk) (cuda_scalar_moist_rk.py:141-151)."""
    if isinstance(nrk, bool) or not isinstance(nrk, int) or not 1 <= nrk <= CM1_NRKMAX:
        raise ValueError("nrk must be 1, 2 or 3")
    return float(np.float32(np.float32(dt) / np.float32(CM1_NRKMAX + 1 - nrk)))


def theta_weps_f64() -> float:
    """weps = f32(10)*f32(epsilon) for theta (cuda_scalar_moist_rk.py:238-240)."""
    return float(np.float32(10.0) * np.float32(CM1_EPSILON_F32))


def moisture_weps_f64(species: int, qmag: tuple[float, ...] = CM1_QMAG_F32) -> float:
    """weps = f32(qmag(n))*f32(epsilon) (cuda_scalar_moist_rk.py:168-192, solve2.F:1472)."""
    if isinstance(species, bool) or not isinstance(species, int):
        raise TypeError("species must be an integer")
    if not 0 <= species < len(qmag):
        raise ValueError(f"species must index one of {len(qmag)} CM1 q3d scalars")
    return float(np.float32(qmag[species]) * np.float32(CM1_EPSILON_F32))


def _mass(mass_flux: Mapping[str, Any], ni, nj, nk):
    sh = shapes(ni, nj, nk)
    return (require(mass_flux["rru"], "rru", sh["u"]), require(mass_flux["rrv"], "rrv", sh["v"]),
            require(mass_flux["rrw"], "rrw", sh["w"]))


def pack_scalar(th0: Any, source: Any, padded: Any, *, ni: int, nj: int, nk: int, theta: bool,
                lateral_sides, stream: Any | None = None) -> None:
    """Interior pack ((th0-300)+th3d for theta) and the family halo fill."""
    _native.call("gm1hip_pack_scalar", ptr(th0), ptr(source), ptr(padded), ni, nj, nk, int(theta),
                 lateral_code(lateral_sides), stream_handle(stream))


def launch_scalar_rk_compose(
    padded: Any, mass_flux: Mapping[str, Any], metrics: Mapping[str, Any], divx: Any, a0: Any, held: Any,
    divterm: Any | None, out: Any, *, ni: int, nj: int, nk: int, nrk: int, dt: float, dx: float, dy: float,
    dz: float, weps: float, lateral_sides, stream: Any | None = None,
) -> None:
    """Advection (5th order, WENO5 at nrk=3) and the A-anchored compose of one scalar."""
    sh = shapes(ni, nj, nk)
    cell = (nk, nj, ni)
    padded = require(padded, "padded", sh["s"])
    rru, rrv, rrw = _mass(mass_flux, ni, nj, nk)
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("c1", "c2", "rr0", "gz")}
    for name, t in (("divx", divx), ("a0", a0), ("held", held), ("out", out)):
        require(t, name, cell)
    if divterm is not None:
        require(divterm, "divterm", cell)
    use_weno = int(nrk == CM1_NRKMAX)
    _native.call("gm1hip_scalar_rk_stage", ptr(padded), ptr(rru), ptr(rrv), ptr(rrw), ptr(m["c1"]), ptr(m["c2"]),
                 ptr(m["rr0"]), ptr(m["gz"]), ptr(divx), ptr(a0), ptr(held), ptr(divterm), ptr(out), ni, nj, nk,
                 use_weno, float(weps), float(dx), float(dy), float(dz), lateral_code(lateral_sides),
                 scalar_moist_rk_stage_timestep_f32(dt, nrk), stream_handle(stream))


def launch_pdef_moisture(
    padded: Any, mass_flux: Mapping[str, Any], metrics: Mapping[str, Any], divx: Any, qa: Any, held: Any, out: Any,
    scratch: Mapping[str, Any], *, species: int, ni: int, nj: int, nk: int, dt: float, dx: float, dy: float,
    dz: float, lateral_sides, stream: Any | None = None,
) -> None:
    """Final-stage moisture row: positive-definite chain fused with the compose."""
    sh = shapes(ni, nj, nk)
    cell = (nk, nj, ni)
    padded = require(padded, "padded", sh["s"])
    rru, rrv, rrw = _mass(mass_flux, ni, nj, nk)
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("c1", "c2", "rho0_3d", "rr0", "gz", "rgz")}
    for name, t in (("divx", divx), ("qa", qa), ("held", held), ("out", out), ("advx", scratch["advx"]),
                    ("advy", scratch["advy"]), ("advz", scratch["advz"])):
        require(t, name, cell)
    _native.call("gm1hip_pdef_moisture_stage", ptr(padded), ptr(rru), ptr(rrv), ptr(rrw), ptr(m["c1"]),
                 ptr(m["c2"]), ptr(m["rho0_3d"]), ptr(m["rr0"]), ptr(m["gz"]), ptr(m["rgz"]), ptr(divx), ptr(qa),
                 ptr(held), ptr(out), ptr(scratch["advx"]), ptr(scratch["advy"]), ptr(scratch["advz"]), ni, nj, nk,
                 int(species != 0), moisture_weps_f64(species), float(dx), float(dy), float(dz),
                 lateral_code(lateral_sides), scalar_moist_rk_stage_timestep_f32(dt, CM1_NRKMAX),
                 stream_handle(stream))


def launch_scalar_stage_from_carrier(
    carrier: Any, th0: Any | None, mass_flux: Mapping[str, Any], metrics: Mapping[str, Any], divx: Any, a0: Any,
    held: Any, divterm: Any | None, scratch: Any, *, ni: int, nj: int, nk: int, nrk: int, dt: float, dx: float,
    dy: float, rdz: float, lateral_sides, flat_column: Any | None = None, pdef_species: int | None = None,
    pdef_scratch: Mapping[str, Any] | None = None, weps: float | None = None, stream: Any | None = None,
    c12_columns: bool = False, copy_back: bool = True,
) -> None:
    """One scalar's RK stage read straight from its unpadded carrier (nk, nj, ni), result back in it.

    The pack into a padded carrier and its lateral-family fill
    (cuda_scalar_moist_rk.py:255-280, 1041-1160) are folded into the tiles'
    loads (``gm1hip_scalar_rk_stage_src`` / ``gm1hip_pdef_moisture_stage_src``);
    ``th0`` makes the carrier theta' (packed as (th0 - 300) + theta').  The
    kernel writes ``scratch`` (cell-shaped) -- a tile reads its neighbours'
    halo words of the carrier, so the carrier is updated by a copy after it.
    Terrain columns take ``metrics["gz"]``/``["rgz"]`` and ``rdz = 1/dz``; flat
    stretched columns ``flat_column`` and the metric rdz.  ``pdef_species``
    selects the final-stage positive-definite moisture chain.  ``c12_columns``
    states that c1 and c2 hold one word per level at every column (checked by
    ``verify_horizontally_uniform``); the kernels then read the first column's.
    ``copy_back=False`` leaves the result in ``scratch`` (a caller alternating
    two carriers across stages).
    """
    cell = (nk, nj, ni)
    sh = shapes(ni, nj, nk)
    for name, t in (("carrier", carrier), ("divx", divx), ("a0", a0), ("held", held), ("scratch", scratch)):
        require(t, name, cell)
    if divterm is not None:
        require(divterm, "divterm", cell)
    if carrier.data_ptr() == scratch.data_ptr():
        raise ValueError("scratch must not alias the carrier")
    rru, rrv, rrw = (require(mass_flux["rru"], "rru", sh["u"]), require(mass_flux["rrv"], "rrv", sh["v"]),
                     require(mass_flux["rrw"], "rrw", sh["w"]))
    terrain = flat_column is None
    gz = metrics["gz"] if terrain else None
    obc = lateral_code(lateral_sides) | (OBC_C12_COLUMN if c12_columns else 0)
    dttmp = scalar_moist_rk_stage_timestep_f32(dt, nrk)
    h = stream_handle(stream)
    if pdef_species is None:
        use_weno = int(nrk == CM1_NRKMAX)
        _native.call("gm1hip_scalar_rk_stage_src", ptr(carrier), ptr(th0), ptr(rru), ptr(rrv), ptr(rrw),
                     ptr(metrics["c1"]), ptr(metrics["c2"]), ptr(metrics["rr0"]), ptr(gz), ptr(flat_column),
                     ptr(divx), ptr(a0), ptr(held), ptr(divterm), ptr(scratch), ni, nj, nk, use_weno, float(weps),
                     float(dx), float(dy), float(rdz), obc, dttmp, h)
    else:
        if nrk != CM1_NRKMAX or th0 is not None:
            raise ValueError("the positive-definite chain is the final stage's moisture row")
        rgz = metrics["rgz"] if terrain else None
        ps = pdef_scratch or {}
        for name in ("advx", "advy", "advz"):
            require(ps[name], name, cell)
        _native.call("gm1hip_pdef_moisture_stage_src", ptr(carrier), ptr(rru), ptr(rrv), ptr(rrw),
                     ptr(metrics["c1"]), ptr(metrics["c2"]), ptr(metrics["rho0_3d"]), ptr(metrics["rr0"]), ptr(gz),
                     ptr(rgz), ptr(flat_column), ptr(divx), ptr(a0), ptr(held), ptr(scratch), ptr(ps["advx"]),
                     ptr(ps["advy"]), ptr(ps["advz"]), ni, nj, nk, int(pdef_species != 0),
                     moisture_weps_f64(pdef_species), float(dx), float(dy), float(rdz), obc, dttmp, h)
    if copy_back:
        with torch.cuda.stream(stream) if stream is not None else _NoStream():
            carrier.copy_(scratch)


signature("gm1hip_scalar_rk_stage_src_batch", I, P, P, *([P] * 9), P, P, P, P, I, I, I, Fl, Fl, Fl, U, Fl, P)
BATCH_MAX = 4


def launch_scalar_stages_from_carriers(
    carriers: list, th0s: list, mass_flux: Mapping[str, Any], metrics: Mapping[str, Any], divx: Any, a0s: list,
    helds: list, divterms: list, outs: list, *, ni: int, nj: int, nk: int, nrk: int, dt: float, dx: float, dy: float,
    rdz: float, lateral_sides, flat_column: Any | None = None, stream: Any | None = None, c12_columns: bool = False,
) -> None:
    """Several scalars' stage before the last (fifth-order fluxes) in one pass, each into its ``outs`` entry.

    Each output holds the words :func:`launch_scalar_stage_from_carrier` with
    ``copy_back=False`` leaves for that scalar; the mass fluxes, divx and
    the tile addressing are shared.  ``th0s`` / ``divterms`` entries may be
    None.  The reference advects theta and each species in passes of their
    own (cuda_scalar_moist_rk.py:1296, 1314-1340).
    """
    ns = len(carriers)
    if not 1 <= ns <= BATCH_MAX or not all(len(x) == ns for x in (th0s, a0s, helds, divterms, outs)):
        raise ValueError(f"a batch holds 1..{BATCH_MAX} scalars, with one entry per scalar in every list")
    if nrk >= CM1_NRKMAX:
        raise ValueError("the batched compose is the fifth-order form of the stages before the last")
    cell = (nk, nj, ni)
    sh = shapes(ni, nj, nk)
    require(divx, "divx", cell)
    for q in range(ns):
        for name, t in (("carrier", carriers[q]), ("a0", a0s[q]), ("held", helds[q]), ("out", outs[q])):
            require(t, f"{name}[{q}]", cell)
        if divterms[q] is not None:
            require(divterms[q], f"divterm[{q}]", cell)
        if th0s[q] is not None:
            require(th0s[q], f"th0[{q}]", (nk,))
    written = {int(t.data_ptr()) for t in outs}
    if len(written) != ns or written & {int(t.data_ptr()) for t in carriers}:
        raise ValueError("each output must be its own array, apart from every carrier")
    rru, rrv, rrw = (require(mass_flux["rru"], "rru", sh["u"]), require(mass_flux["rrv"], "rrv", sh["v"]),
                     require(mass_flux["rrw"], "rrw", sh["w"]))
    terrain = flat_column is None
    obc = lateral_code(lateral_sides) | (OBC_C12_COLUMN if c12_columns else 0)
    dttmp = scalar_moist_rk_stage_timestep_f32(dt, nrk)

    def arr(ts):
        return (ctypes.c_void_p * ns)(*[ptr(t) for t in ts])

    _native.call("gm1hip_scalar_rk_stage_src_batch", ns, arr(carriers), arr(th0s), ptr(rru), ptr(rrv), ptr(rrw),
                 ptr(metrics["c1"]), ptr(metrics["c2"]), ptr(metrics["rr0"]), ptr(metrics["gz"] if terrain else None),
                 ptr(flat_column), ptr(divx), arr(a0s), arr(helds), arr(divterms), arr(outs), ni, nj, nk,
                 float(dx), float(dy), float(rdz), obc, dttmp, stream_handle(stream))


# bit 16 of the lateral-family word: c1 and c2 are horizontally uniform
OBC_C12_COLUMN = 1 << 16
_uniform_checked: set[tuple] = set()


def verify_horizontally_uniform(metrics: Mapping[str, Any], names: tuple[str, ...] = ("c1", "c2")) -> bool:
    """Check once on the device that every level of each named (levels, nj, ni) metric holds one word at every
    column (the first column's, sign and NaN payloads included).  Remembered for these tensors until they are
    written again, so the check runs outside any step."""
    key = tuple((int(metrics[n].data_ptr()), int(metrics[n]._version)) for n in names)
    if key in _uniform_checked:
        return True
    for n in names:
        t = metrics[n]
        if t.dim() != 3 or t.dtype != torch.float32:
            return False
        w = t.contiguous().view(torch.int32)
        if not bool(torch.equal(w, w[:, :1, :1].expand_as(w))):
            return False
    _uniform_checked.add(key)
    return True


class _NoStream:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


# ----------------------------------------------------------------------------- the stateful stage API
# create_/begin_/launch_cuda_scalar_moist_rk_stage (cuda_scalar_moist_rk.py:823-1499): a state binds
# the live theta and moisture carriers (A level, current, held) and the stage scratch; stages run
# in the order 1, 2, 3 and a new large step begins only after stage 3.


@dataclass
class CudaScalarMoistRKState:
    ni: int
    nj: int
    nk: int
    th0: torch.Tensor
    theta_a: torch.Tensor
    theta_current: torch.Tensor
    moisture_a: torch.Tensor
    moisture_current: torch.Tensor
    theta_held: torch.Tensor
    moisture_held: torch.Tensor
    scratch: torch.Tensor
    advx: torch.Tensor
    advy: torch.Tensor
    advz: torch.Tensor
    stream: Any
    device_id: int
    next_nrk: int = 1
    completed: bool = False
    qmag: tuple[float, ...] = CM1_QMAG_F32
    names: tuple[str, ...] = CM1_SCALAR_MOIST_SPECIES
    lateral_sides: tuple[int, int, int, int] = (2, 2, 2, 2)


@dataclass(frozen=True)
class CudaScalarMoistRKStageLaunch:
    state: CudaScalarMoistRKState
    ready_event: Any
    receipt: dict[str, Any] = field(default_factory=dict)


def create_cuda_scalar_moist_rk_state(
    carriers: Mapping[str, Any], held: Mapping[str, Any], th0: Any, *, ni: int, nj: int, nk: int,
    stream: Any | None = None, qmag: tuple[float, ...] = CM1_QMAG_F32,
    names: tuple[str, ...] = CM1_SCALAR_MOIST_SPECIES, lateral_sides: Any = (2, 2, 2, 2),
) -> CudaScalarMoistRKState:
    """Bind the live carriers (tha, th3d, qa, q3d; thten1, qten) and allocate the stage scratch (:856-966).

    ``lateral_sides`` is the family the scalars' halos follow (the reference
    binds it through its compiled open-boundary image).
    """
    for name, value in (("ni", ni), ("nj", nj), ("nk", nk)):
        if isinstance(value, bool) or not isinstance(value, int) or value < 10:
            raise ValueError(f"{name} must be an integer >= 10")
    if not qmag or any(isinstance(v, bool) or not isinstance(v, (int, float)) or v <= 0.0 for v in qmag):
        raise ValueError("qmag must be CM1's positive per-scalar magnitude table")
    if len(names) != len(qmag):
        raise ValueError(f"names {tuple(names)!r} must give CM1's qname for each of the {len(qmag)} rows qmag "
                         "describes")
    physical = (nk, nj, ni)
    moisture = (len(qmag), nk, nj, ni)
    bound = {
        "th0": require(th0, "th0", (nk,)), "tha": require(carriers["tha"], "tha", physical),
        "th3d": require(carriers["th3d"], "th3d", physical), "qa": require(carriers["qa"], "qa", moisture),
        "q3d": require(carriers["q3d"], "q3d", moisture), "thten1": require(held["thten1"], "thten1", physical),
        "qten": require(held["qten"], "qten", moisture),
    }
    no_alias(bound, "live scalar/moisture RK carriers must not alias")
    dev = bound["th3d"].device
    z = lambda: torch.zeros(physical, dtype=torch.float32, device=dev)  # noqa: E731
    return CudaScalarMoistRKState(
        ni, nj, nk, bound["th0"], bound["tha"], bound["th3d"], bound["qa"], bound["q3d"], bound["thten1"],
        bound["qten"], z(), z(), z(), z(), stream, dev.index if dev.index is not None else torch.cuda.current_device(),
        qmag=tuple(float(v) for v in qmag), names=tuple(names), lateral_sides=tuple(int(v) for v in lateral_sides))


def begin_cuda_scalar_moist_rk_step(state: CudaScalarMoistRKState) -> None:
    """Reset the cadence guard after the caller advances the A and held fields (:969-978)."""
    if not isinstance(state, CudaScalarMoistRKState):
        raise TypeError("state must be a CudaScalarMoistRKState")
    if not state.completed or state.next_nrk != 4:
        raise ValueError("cannot begin a new large step before stage 3 completes")
    state.next_nrk = 1
    state.completed = False


def launch_cuda_scalar_moist_rk_stage(
    state: CudaScalarMoistRKState, mass_flux: Mapping[str, Any], metrics: Mapping[str, Any], divx: Any,
    theta_divergence_term: Any, *, nrk: int, dt: float, dx: float, dy: float, dz: float | None = None,
    vertical: Any | None = None, stream: Any | None = None, compiled: Any = None, advection_compiled: Any = None,
    open_compiled: Any = None,
) -> CudaScalarMoistRKStageLaunch:
    """One theta + moisture RK stage (:1041-1499): 5th order below nrk=3, WENO5 theta and the
    positive-definite moisture chain at nrk=3, each composed A-anchored into the current carrier.

    Exactly one vertical coordinate: ``dz`` (the uniform terrain-coordinate
    column) or ``vertical`` (a flat stretched column,
    ``gm1_amd.stretched.FlatStretchedColumn``).
    """
    if (dz is None) == (vertical is None):
        raise ValueError("the scalar/moist stage takes exactly one vertical coordinate: dz for the uniform "
                         "column, or vertical for a stretched one")
    if not isinstance(state, CudaScalarMoistRKState):
        raise TypeError("state must be a CudaScalarMoistRKState")
    if isinstance(nrk, bool) or nrk != state.next_nrk or state.completed:
        raise ValueError(f"scalar/moist RK cadence requires stage {state.next_nrk}, received {nrk}")
    for name, value in (("dx", dx), ("dy", dy)) + ((("dz", dz),) if dz is not None else ()):
        if isinstance(value, bool) or not isinstance(value, (int, float)):
            raise TypeError(f"{name} must be numeric")
        if not np.isfinite(float(value)) or float(value) <= 0.0:
            raise ValueError(f"{name} must be finite and > 0")
    if torch.cuda.current_device() != state.device_id:
        raise ValueError("scalar/moist RK state is bound to another HIP device")
    stream = stream if stream is not None else state.stream
    if stream_handle(stream) != stream_handle(state.stream):
        raise ValueError("scalar/moist RK state must remain on its bound stream")
    physical = (state.nk, state.nj, state.ni)
    require(theta_divergence_term, "theta_divergence_term", physical)
    require(divx, "divx", physical)
    if vertical is None:
        flat_column, rdz = None, float(np.float32(1.0) / np.float32(dz))
    else:
        flat_column, rdz = vertical.column, float(vertical.rdz)
    grid = dict(ni=state.ni, nj=state.nj, nk=state.nk, nrk=nrk, dt=dt, dx=dx, dy=dy, rdz=rdz,
                lateral_sides=state.lateral_sides, flat_column=flat_column, stream=stream)
    launch_scalar_stage_from_carrier(state.theta_current, state.th0, mass_flux, metrics, divx, state.theta_a,
                                     state.theta_held, theta_divergence_term, state.scratch, weps=theta_weps_f64(),
                                     **grid)
    for sp in range(len(state.qmag)):
        if nrk < CM1_NRKMAX:
            launch_scalar_stage_from_carrier(state.moisture_current[sp], None, mass_flux, metrics, divx,
                                             state.moisture_a[sp], state.moisture_held[sp], None, state.scratch,
                                             weps=moisture_weps_f64(sp, state.qmag), **grid)
        else:
            if state.qmag[sp] != CM1_QMAG_F32[0]:
                raise ValueError("the positive-definite chain's weps is CM1's 1e-2 qmag")
            launch_scalar_stage_from_carrier(state.moisture_current[sp], None, mass_flux, metrics, divx,
                                             state.moisture_a[sp], state.moisture_held[sp], None, state.scratch,
                                             pdef_species=sp, pdef_scratch={"advx": state.advx, "advy": state.advy,
                                                                            "advz": state.advz}, **grid)
    state.next_nrk = nrk + 1
    state.completed = nrk == CM1_NRKMAX
    receipt = {"nrk": nrk, "dttmp": scalar_moist_rk_stage_timestep_f32(dt, nrk), "species": list(state.names),
               "theta": "weno5" if nrk == CM1_NRKMAX else "flux5",
               "moisture": "positive-definite" if nrk == CM1_NRKMAX else "flux5",
               "vertical": "uniform" if vertical is None else "stretched", "target_arch": "gfx950"}
    return CudaScalarMoistRKStageLaunch(state, record(stream), receipt)


__all__ = [
    "CM1_QMAG_F32", "CM1_SCALAR_MOIST_SPECIES", "CudaScalarMoistRKStageLaunch", "CudaScalarMoistRKState",
    "begin_cuda_scalar_moist_rk_step", "create_cuda_scalar_moist_rk_state", "launch_cuda_scalar_moist_rk_stage",
    "launch_pdef_moisture", "launch_scalar_rk_compose", "launch_scalar_stage_from_carrier",
    "moisture_weps_f64", "pack_scalar", "scalar_moist_rk_stage_timestep_f32", "theta_weps_f64",
]
