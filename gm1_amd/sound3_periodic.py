"""The flat, uniform, doubly periodic psolver=3 acoustic stage on gfx950.

Host mirror of ``cuda_sound3_periodic.py``: the array inventory and shapes
(``SCALAR_NAMES`` … ``ALL_ARRAY_NAMES`` :59-94, ``array_shapes`` :449-491),
``default_periodic_face_weights`` (:494-503), ``validate_periodic_contract``
(:569-602, on device tensors) and ``launch_cuda_sound3_periodic_three_stage``
(:1216-1460).  Each stage is one C-ABI call, ``gm1hip_sound3_periodic_stage``,
that takes the 47 device pointers as one table in ``ALL_ARRAY_NAMES`` order and
enqueues prepare, the small steps and the stage fills on the caller's stream.
"""

from __future__ import annotations

import ctypes
from collections.abc import Mapping
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import _native
from ._device import record, require, stream_handle
from ._native import Fl, I, P, signature
from .grid import rk_stage_dt_f32, sound3_stage_schedule_f32

CM1_SOUND3_PERIODIC_ALPH_F32 = 0.60
CM1_SOUND3_PERIODIC_KDIV_F32 = 0.10
CM1_SOUND3_PERIODIC_CP_F32 = 1005.7
CM1_SOUND3_PERIODIC_SMEPS_F32 = 1.0e-30

SCALAR_NAMES = ("mh", "rr0", "rf0", "c1", "c2", "mf", "piadv", "ppterm", "thv", "ppten", "ppi", "pp3d", "ppx",
                "ppd", "qk", "pk1", "pk2", "mm", "fpk", "divh", "hadvp", "elim_c", "elim_d", "tri_aa", "tri_bb",
                "tri_cc", "tri_dd", "tri_pivot", "tri_residual")
W_NAMES = ("wten", "wa", "w3d", "rrw")
U_NAMES = ("ua", "u3d", "uten", "rru")
V_NAMES = ("va", "v3d", "vten", "rrv")
METRIC_NAMES = ("dtu0", "dtv0", "uf", "vf", "uh", "vh")
ALL_ARRAY_NAMES = SCALAR_NAMES + W_NAMES + U_NAMES + V_NAMES + METRIC_NAMES
COMPARE_OUTPUT_NAMES = ("u3d", "v3d", "w3d", "pp3d", "ppx", "ppd", "rru", "rrv", "rrw", "qk", "pk1", "pk2", "mm",
                        "fpk", "divh", "hadvp", "elim_c", "elim_d", "tri_aa", "tri_bb", "tri_cc", "tri_dd",
                        "tri_pivot", "tri_residual")

signature("gm1hip_sound3_periodic_stage", P, I, I, I, I, I, Fl, Fl, Fl, Fl, Fl, Fl, Fl, P)


def _validate_dimensions(ni: int, nj: int, nk: int) -> None:
    for name, value in (("ni", ni), ("nj", nj), ("nk", nk)):
        if isinstance(value, bool) or not isinstance(value, int) or value < 3:
            raise ValueError(f"{name} must be an integer >= 3")


def array_shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    _validate_dimensions(ni, nj, nk)
    shapes = {name: (nk + 2, nj + 2, ni + 2) for name in SCALAR_NAMES}
    shapes.update({name: (nk + 2, nj, ni) for name in W_NAMES})
    shapes.update({name: (nk + 2, nj + 2, ni + 1) for name in U_NAMES})
    shapes.update({name: (nk + 2, nj + 1, ni + 2) for name in V_NAMES})
    shapes.update({"dtu0": (nj + 2, ni + 1), "dtv0": (nj + 1, ni + 2), "uf": (ni + 2,), "vf": (nj + 2,),
                   "uh": (ni + 2,), "vh": (nj + 2,)})
    return shapes


def default_periodic_face_weights(ni: int, nj: int) -> tuple[np.ndarray, np.ndarray]:
    """Every live face weighs one; serial periodic boundaries zero no normal face."""
    _validate_dimensions(ni, nj, 3)
    dtu0 = np.zeros((nj + 2, ni + 1), dtype=np.float32)
    dtv0 = np.zeros((nj + 1, ni + 2), dtype=np.float32)
    dtu0[1 : nj + 1, :] = np.float32(1.0)
    dtv0[:, 1 : ni + 1] = np.float32(1.0)
    return dtu0, dtv0


def _require_device_arrays(arrays: Mapping[str, Any], ni: int, nj: int, nk: int) -> dict[str, torch.Tensor]:
    shapes = array_shapes(ni, nj, nk)
    missing = [name for name in ALL_ARRAY_NAMES if name not in arrays]
    if missing:
        raise ValueError(f"missing periodic sound3 arrays: {', '.join(missing)}")
    result = {name: require(arrays[name], name, shapes[name]) for name in ALL_ARRAY_NAMES}
    if len({int(t.data_ptr()) for t in result.values()}) != len(result):
        raise ValueError("periodic sound3 device arrays must not alias")
    return result


def _bits_equal(left: torch.Tensor, right: torch.Tensor) -> bool:
    return bool(torch.equal(left.contiguous().view(torch.int32), right.contiguous().view(torch.int32)))


def validate_periodic_contract(arrays: Mapping[str, Any], *, ni: int, nj: int, nk: int) -> None:
    """The serial-periodic preconditions (cuda_sound3_periodic.py:569-602), read on the device."""
    a = _require_device_arrays(arrays, ni, nj, nk)
    dev = a["dtu0"].device
    eu, ev = (torch.from_numpy(x).to(dev) for x in default_periodic_face_weights(ni, nj))
    if not _bits_equal(a["dtu0"], eu) or not _bits_equal(a["dtv0"], ev):
        raise ValueError("dtu0/dtv0 differ from CM1 serial-periodic default face weights")
    for name in ("uf", "vf", "uh", "vh"):
        if not _bits_equal(a[name], torch.ones_like(a[name])):
            raise ValueError(f"{name} must be exactly one for this uniform flat-grid slice")
    K = slice(1, nk + 1)
    for name in ("piadv", "thv", "ppi", "ppx"):
        v = a[name]
        if not _bits_equal(v[K, :, 0], v[K, :, ni]):
            raise ValueError(f"{name} west periodic halo is invalid")
        if not _bits_equal(v[K, :, ni + 1], v[K, :, 1]):
            raise ValueError(f"{name} east periodic halo is invalid")
        if not _bits_equal(v[K, 0, :], v[K, nj, :]):
            raise ValueError(f"{name} south periodic halo is invalid")
        if not _bits_equal(v[K, nj + 1, :], v[K, 1, :]):
            raise ValueError(f"{name} north periodic halo is invalid")
    if not _bits_equal(a["ua"][K, 1 : nj + 1, 0], a["ua"][K, 1 : nj + 1, ni]):
        raise ValueError("ua duplicate periodic x faces differ")
    if not _bits_equal(a["va"][K, 0, 1 : ni + 1], a["va"][K, nj, 1 : ni + 1]):
        raise ValueError("va duplicate periodic y faces differ")


@dataclass(frozen=True)
class CudaSound3PeriodicLaunch:
    ready_event: Any
    compiled: Any
    receipt: dict[str, Any]


def _table(a: dict[str, torch.Tensor]) -> ctypes.Array:
    table = (ctypes.c_void_p * len(ALL_ARRAY_NAMES))()
    for slot, name in enumerate(ALL_ARRAY_NAMES):
        table[slot] = a[name].data_ptr()
    return table


def launch_cuda_sound3_periodic_stage(
    arrays: Mapping[str, Any], *, ni: int, nj: int, nk: int, nrk: int, dt: float, nsound: int, rdx: float,
    rdy: float, rdz: float, cp: float = CM1_SOUND3_PERIODIC_CP_F32, alph: float = CM1_SOUND3_PERIODIC_ALPH_F32,
    kdiv: float = CM1_SOUND3_PERIODIC_KDIV_F32, stream: Any | None = None, _checked: dict | None = None,
) -> dict[str, Any]:
    """One RK/acoustic stage (cpu_sound3_periodic_stage_inplace's order), enqueued without synchronising."""
    if isinstance(nrk, bool) or not isinstance(nrk, int) or nrk not in (1, 2, 3):
        raise ValueError("nrk must select explicit CM1 RK stage 1, 2, or 3")
    a = _checked if _checked is not None else _require_device_arrays(arrays, ni, nj, nk)
    nloop, dts = sound3_stage_schedule_f32(dt, nsound, nrk)
    f = lambda x: float(np.float32(x))  # noqa: E731
    table = _table(a)
    _native.call("gm1hip_sound3_periodic_stage", ctypes.cast(table, ctypes.c_void_p), ni, nj, nk, nrk, nloop,
                 f(dts), f(rdx), f(rdy), f(rdz), f(cp), f(alph), f(kdiv), stream_handle(stream))
    return {"nrk": nrk, "dttmp_f32": rk_stage_dt_f32(dt, nrk), "nloop": nloop, "dts_f32": f(dts),
            "ppd_boundary_fills": nloop - 1, "stage_velocity_boundary_fills": 1,
            "stage_pressure_boundary_fills": 1 if nrk < 3 else 0, "ppx_written": nrk == 3}


def launch_cuda_sound3_periodic_three_stage(
    arrays: Mapping[str, Any], *, ni: int, nj: int, nk: int, dt: float, nsound: int, rdx: float, rdy: float,
    rdz: float, cp: float = CM1_SOUND3_PERIODIC_CP_F32, alph: float = CM1_SOUND3_PERIODIC_ALPH_F32,
    kdiv: float = CM1_SOUND3_PERIODIC_KDIV_F32, stream: Any | None = None, compiled: Any | None = None,
) -> CudaSound3PeriodicLaunch:
    """Enqueue one three-stage periodic acoustic transition without synchronising (:1216-1460)."""
    a = _require_device_arrays(arrays, ni, nj, nk)
    cadence = [launch_cuda_sound3_periodic_stage(a, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=dt, nsound=nsound, rdx=rdx,
                                                 rdy=rdy, rdz=rdz, cp=cp, alph=alph, kdiv=kdiv, stream=stream,
                                                 _checked=a)
               for nrk in (1, 2, 3)]
    return CudaSound3PeriodicLaunch(record(stream), compiled, {
        "domain": {"ni": ni, "nj": nj, "nk": nk, "layout": "scalar (nk+2,nj+2,ni+2), u (nk+2,nj+2,ni+1), "
                                                           "v (nk+2,nj+1,ni+2), w (nk+2,nj,ni)"},
        "cadence": cadence,
        "acoustic_reals": {"cp": float(np.float32(cp)), "alph": float(np.float32(alph)),
                           "kdiv": float(np.float32(kdiv)), "smeps": CM1_SOUND3_PERIODIC_SMEPS_F32},
        "target_arch": "gfx950",
    })


__all__ = [
    "ALL_ARRAY_NAMES", "COMPARE_OUTPUT_NAMES", "CudaSound3PeriodicLaunch", "array_shapes",
    "default_periodic_face_weights", "launch_cuda_sound3_periodic_stage", "launch_cuda_sound3_periodic_three_stage",
    "validate_periodic_contract",
]
