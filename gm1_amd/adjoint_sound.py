"""Adjoint of one psolver=3 acoustic RK stage on gfx950 (dry, flat terrain, periodic sides).

The forward is ``launch_cuda_psolver3_terrain_open_stage``
(cuda_psolver3_terrain_open.py:1436-1507; CPU form :774-893) on flat metrics
(gz = rgzu = rgzv = 1, dzdx = dzdy = gx* = gy* = 0, a horizontally uniform
base state) with CM1's (1, 1, 1, 1) lateral sides.  For fixed thv, ppterm
and piadv the stage is linear in its state and forcing; the device adjoint
(``gm1hip_psolver3_stage_adjoint``) is the transpose of that map, taken
statement by statement in reverse, in binary32.

GM1 has no adjoint: this is the hand-written adjoint of the dry acoustic
stage, reachable as a plain launcher and as ``PeriodicAcousticStage``, a
``torch.autograd.Function`` whose forward is the device stage and whose
backward is the device adjoint.
"""

from __future__ import annotations

import ctypes
from typing import Any, Mapping

import numpy as np
import torch

from . import _native
from ._device import check_spacing, no_alias, record, require, stream_handle, terrain_static_shapes
from ._native import Fl, I, P, signature
from .errors import refuse
from .grid import VerticalProfiles, sound3_stage_schedule_f32, terrain_coordinate_profiles
from .psolver3 import (CM1_ALPH_F32, CM1_KDIV_F32, device_profiles, launch_cuda_psolver3_terrain_open_stage,
                       terrain_open_sound_shapes)

signature("gm1hip_psolver3_stage_adjoint", P, I, I, I, I, I, Fl, Fl, Fl, Fl, Fl, P)

# output adjoints of one stage (what the stage writes)
STATE = ("u3d", "v3d", "w3d", "pp3d", "ppd", "ppx", "rru", "rrv", "rrw")
# inputs of one stage that it reads
INPUTS = ("u3d", "v3d", "w3d", "pp3d", "ppx", "ua", "va", "wa", "ppi", "uten", "vten", "wten", "ppten")
COEFFICIENTS = ("thv", "ppterm", "piadv")
FLAT_ONES = ("gz", "rgzu", "rgzv")
FLAT_ZEROS = ("dzdx", "dzdy", "gx", "gy", "gxu", "gyv")
PERIODIC = (1, 1, 1, 1)

# slot order of the device entry point's pointer table
_TABLE = ("thv", "ppterm", "piadv", "rho0_3d", "rr0", "rf0", "c1", "c2", "profiles", "qk", "pk1", "pk2", "mm",
          "fpkb", "elim_c", "elim_d", "u3d", "v3d", "w3d", "pp3d", "ppd", "ppx", "rru", "rrv", "rrw",
          "uten", "vten", "wten", "ppten", "ua", "va", "wa", "ppi")
_LIKE = {"uten": "u3d", "vten": "v3d", "wten": "w3d", "ppten": "pp3d", "ua": "u3d", "va": "v3d", "wa": "w3d",
         "ppi": "pp3d", "qk": "pp3d", "pk1": "pp3d", "pk2": "pp3d", "mm": "pp3d", "fpkb": "pp3d",
         "elim_c": "pp3d", "elim_d": "pp3d"}


def _check_sides(lateral_sides) -> None:
    if tuple(lateral_sides) != PERIODIC:
        refuse(code="GM1-ADJOINT-GEOMETRY", gate="adjoint", option="lateral_sides", value=tuple(lateral_sides),
               reason="the acoustic-stage adjoint transposes the doubly periodic stage only",
               remedy="use lateral_sides=(1, 1, 1, 1)")


def check_flat_periodic(metrics: Mapping[str, torch.Tensor], lateral_sides) -> None:
    """Refuse geometries the adjoint does not transpose (synchronises the stream)."""
    _check_sides(lateral_sides)
    for name in FLAT_ONES:
        if not bool(torch.all(metrics[name] == 1.0)):
            refuse(code="GM1-ADJOINT-GEOMETRY", gate="adjoint", option=name, value="not 1",
                   reason="the acoustic-stage adjoint transposes flat-terrain metrics only",
                   remedy="use flat terrain (gz = rgzu = rgzv = 1)")
    for name in FLAT_ZEROS:
        if bool(torch.any(metrics[name] != 0.0)):
            refuse(code="GM1-ADJOINT-GEOMETRY", gate="adjoint", option=name, value="not 0",
                   reason="the acoustic-stage adjoint transposes flat-terrain metrics only",
                   remedy="use flat terrain (zero slopes)")
    rho = metrics["rho0_3d"]
    if not bool(torch.all(rho == rho[:, :1, :1])):
        refuse(code="GM1-ADJOINT-GEOMETRY", gate="adjoint", option="rho0_3d", value="varies horizontally",
               reason="the pressure-advection adjoint reads one base density per level",
               remedy="use a horizontally uniform base state")


def launch_cuda_psolver3_stage_adjoint(
    bar: Mapping[str, torch.Tensor], coefficients: Mapping[str, torch.Tensor], metrics: Mapping[str, torch.Tensor],
    *, ni: int, nj: int, nk: int, nrk: int, dt: float, nsound: int, dx: float, dy: float, dz: float,
    lateral_sides=PERIODIC, profiles: VerticalProfiles | None = None, alph: float | None = None,
    kdiv: float | None = None, check_geometry: bool = True, stream: Any | None = None,
) -> tuple[dict[str, torch.Tensor], Any]:
    """Input adjoints of one acoustic RK stage from its output adjoints.

    ``bar`` holds adjoints of the stage's outputs (``STATE``, the forward's
    shapes); it is not modified.  ``coefficients`` holds the thv, ppterm and
    piadv the forward stage ran with.  Returned, with the event recorded after
    the launches: adjoints of u3d, v3d, w3d, pp3d (the state the stage starts
    from at nrk = 1, or leaves untouched at nrk > 1), ppx, wa, uten, vten,
    wten, ppten, and at nrk > 1 ua, va, ppi.
    """
    if isinstance(nrk, bool) or not isinstance(nrk, int) or nrk not in (1, 2, 3):
        raise ValueError("nrk must select explicit CM1 RK stage 1, 2, or 3")
    _check_sides(lateral_sides)
    dt, dx, dy, dz = (check_spacing(n, v) for n, v in (("dt", dt), ("dx", dx), ("dy", dy), ("dz", dz)))
    shapes = terrain_open_sound_shapes(ni, nj, nk)
    b = {n: require(bar[n], f"bar[{n}]", shapes[n]) for n in STATE}
    c = {n: require(coefficients[n], n, shapes[n]) for n in COEFFICIENTS}
    ms = terrain_static_shapes(ni=ni, nj=nj, nk=nk)
    m = {n: require(metrics[n], n, ms[n]) for n in ("rho0_3d", "rr0", "rf0", "c1", "c2", *FLAT_ONES, *FLAT_ZEROS)}
    if check_geometry:
        check_flat_periodic(m, lateral_sides)
    dev = b["u3d"].device
    prof = profiles or terrain_coordinate_profiles(nk, dz)
    nloop, dts = sound3_stage_schedule_f32(dt, nsound, nrk)
    alph32 = float(np.float32(CM1_ALPH_F32 if alph is None else alph))
    kdiv32 = float(np.float32(CM1_KDIV_F32 if kdiv is None else kdiv))
    h = stream_handle(stream)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    with torch.cuda.stream(s):
        t = {n: b[n].clone() for n in STATE}
        for n, like in _LIKE.items():
            t[n] = torch.zeros(shapes[like], dtype=torch.float32, device=dev)
    t.update(c)
    t.update({n: m[n] for n in ("rho0_3d", "rr0", "rf0", "c1", "c2")})
    t["profiles"] = device_profiles(prof, dev)
    no_alias({n: t[n] for n in _TABLE if n not in c and n not in m and n != "profiles"},
             "adjoint buffers must not alias")
    table = (ctypes.c_void_p * len(_TABLE))(*(t[n].data_ptr() for n in _TABLE))
    _native.call("gm1hip_psolver3_stage_adjoint", ctypes.cast(table, ctypes.c_void_p), ni, nj, nk, nrk, nloop, dts,
                 dx, dy, alph32, kdiv32, h)
    out = {n: t[n] for n in ("u3d", "v3d", "w3d", "pp3d", "ppx", "wa", "uten", "vten", "wten", "ppten")}
    if nrk != 1:
        out.update({n: t[n] for n in ("ua", "va", "ppi")})
    return out, record(stream)


class PeriodicAcousticStage(torch.autograd.Function):
    """One acoustic RK stage as a differentiable map of its state and forcing.

    ``apply(stage, *tensors)`` with ``tensors`` in ``INPUTS`` order and
    ``stage`` a mapping of the forward's fixed arguments: ``arrays`` (every
    other array of the stage, coefficients included), ``metrics`` and the
    keyword sizes ``ni nj nk nrk dt nsound dx dy dz``.  Returns the stage's
    outputs in ``STATE`` order; thv, ppterm and piadv are held fixed.
    """

    @staticmethod
    def forward(ctx, stage, *tensors):
        kw = {k: stage[k] for k in ("ni", "nj", "nk", "nrk", "dt", "nsound", "dx", "dy", "dz")}
        a = dict(stage["arrays"])
        for n, x in zip(INPUTS, tensors):
            a[n] = x
        # the stage works in place: run it on copies of everything it writes or reads
        a = {n: (v.clone() if isinstance(v, torch.Tensor) else v) for n, v in a.items()}
        launch_cuda_psolver3_terrain_open_stage(a, stage["metrics"], lateral_sides=PERIODIC,
                                                profiles=stage.get("profiles"), **kw)
        ctx.stage = stage
        ctx.coef = {n: a[n] for n in COEFFICIENTS}
        return tuple(a[n] for n in STATE)

    @staticmethod
    def backward(ctx, *grads):
        st = ctx.stage
        kw = {k: st[k] for k in ("ni", "nj", "nk", "nrk", "dt", "nsound", "dx", "dy", "dz")}
        shapes = terrain_open_sound_shapes(st["ni"], st["nj"], st["nk"])
        bar = {}
        for n, g in zip(STATE, grads):
            bar[n] = (torch.zeros(shapes[n], dtype=torch.float32, device=ctx.coef["thv"].device) if g is None
                      else g.contiguous())
        xbar, _ = launch_cuda_psolver3_stage_adjoint(bar, ctx.coef, st["metrics"], profiles=st.get("profiles"),
                                                     check_geometry=st.get("check_geometry", True), **kw)
        grads_in = []
        for n in INPUTS:
            g = xbar.get(n)
            if g is None and ctx.needs_input_grad[1 + INPUTS.index(n)]:
                g = torch.zeros(shapes[n], dtype=torch.float32, device=ctx.coef["thv"].device)
            grads_in.append(g)
        return (None, *grads_in)


__all__ = ["COEFFICIENTS", "INPUTS", "PeriodicAcousticStage", "STATE", "check_flat_periodic",
           "launch_cuda_psolver3_stage_adjoint"]
