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
from typing import Any, Mapping

import numpy as np
import torch

from . import _native
from ._device import NG, lateral_code, ptr, require, shapes, stream_handle, terrain_static_shapes
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
    selects the final-stage positive-definite moisture chain.
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
    obc = lateral_code(lateral_sides)
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
    with torch.cuda.stream(stream) if stream is not None else _NoStream():
        carrier.copy_(scratch)


class _NoStream:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


__all__ = [
    "CM1_QMAG_F32", "CM1_SCALAR_MOIST_SPECIES", "launch_pdef_moisture", "launch_scalar_rk_compose",
    "launch_scalar_stage_from_carrier",
    "moisture_weps_f64", "pack_scalar", "scalar_moist_rk_stage_timestep_f32", "theta_weps_f64",
]
