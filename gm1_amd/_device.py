"""Device-array checks and stream plumbing for the operator entry points.

Arrays are ROCm ``torch.Tensor`` objects (float32, C-contiguous, on the HIP
device), the stand-in for the reference's ``cupy.ndarray`` carriers.  Checks
raise the reference launchers' errors: ``TypeError`` for the wrong type or
dtype, ``ValueError`` for the wrong shape/layout or for aliasing
(e.g. ``cuda_advection5.py:798-809,1001-1007``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Iterable

import torch

NG = 3


def require(value: Any, name: str, shape: tuple[int, ...], *, dtype=torch.float32) -> torch.Tensor:
    if not isinstance(value, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor on the HIP device")
    if value.dtype != dtype:
        raise TypeError(f"{name} must have dtype {str(dtype).replace('torch.', '')}")
    if not value.is_cuda:
        raise TypeError(f"{name} must live on the HIP device")
    if tuple(value.shape) != tuple(shape) or not value.is_contiguous():
        raise ValueError(f"{name} must have C-order shape {tuple(shape)}")
    return value


def no_alias(named: dict[str, torch.Tensor], message: str) -> None:
    spans = []
    for name, t in named.items():
        if t is None:
            continue
        start = t.data_ptr()
        spans.append((start, start + t.numel() * t.element_size(), name))
    spans.sort()
    for (s0, e0, n0), (s1, e1, n1) in zip(spans, spans[1:]):
        if s1 < e0:
            raise ValueError(f"{message} ({n0} overlaps {n1})")


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def stream_handle(stream: Any | None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def record(stream: Any | None) -> torch.cuda.Event | None:
    """A ready event recorded after the launches; None while the stream is being captured."""
    if torch.cuda.is_current_stream_capturing():
        return None
    event = torch.cuda.Event()
    event.record(stream if stream is not None else torch.cuda.current_stream())
    return event


@dataclass(frozen=True)
class Launch:
    """What every launcher returns: the reference's ``...Launch`` triple."""

    ready_event: Any
    compiled: Any
    receipt: dict[str, Any]


def check_grid(ni: int, nj: int, nk: int, minimum: int = 8) -> tuple[int, int, int]:
    for name, value in (("ni", ni), ("nj", nj), ("nk", nk)):
        if isinstance(value, bool) or not isinstance(value, int) or value < minimum:
            raise ValueError(f"{name} must be an integer >= {minimum}")
    return ni, nj, nk


def check_spacing(name: str, value: Any) -> float:
    import math

    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise TypeError(f"{name} must be numeric")
    result = float(value)
    if not math.isfinite(result) or result <= 0.0:
        raise ValueError(f"{name} must be finite and > 0")
    return result


def shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    ng = NG
    return {
        "u": (nk + 2, nj + 2 * ng, ni + 1 + 2 * ng),
        "v": (nk + 2, nj + 1 + 2 * ng, ni + 2 * ng),
        "w": (nk + 3, nj + 2 * ng, ni + 2 * ng),
        "s": (nk + 2, nj + 2 * ng, ni + 2 * ng),
        "p": (nk + 2, nj + 2, ni + 2),
        "cell": (nk, nj, ni),
    }


def terrain_static_shapes(*, ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    """Metric shapes of the reference (cuda_terrain_open_boundary.py:429-459)."""
    return {
        "terrain_msl": (nj, ni), "zh": (nk, nj, ni), "zf": (nk + 1, nj, ni),
        "mh": (nk, nj, ni), "rmh": (nk, nj, ni), "mf": (nk + 1, nj, ni),
        "rmf": (nk + 1, nj, ni), "c1": (nk + 1, nj, ni), "c2": (nk + 1, nj, ni),
        "rho0_3d": (nk, nj, ni), "rr0": (nk, nj, ni), "rf0": (nk + 1, nj, ni),
        "rrf0": (nk + 1, nj, ni), "gz": (nj, ni), "rgz": (nj, ni),
        "gzu": (nj, ni + 1), "rgzu": (nj, ni + 1), "gzv": (nj + 1, ni),
        "rgzv": (nj + 1, ni), "dzdx": (nj, ni), "dzdy": (nj, ni),
        "gx": (nk + 1, nj, ni), "gxu": (nk + 1, nj, ni + 1),
        "gy": (nk + 1, nj, ni), "gyv": (nk + 1, nj + 1, ni),
    }


def lateral_code(lateral_sides: Iterable[int] | None) -> int:
    """Pack CM1's (wbc, ebc, sbc, nbc) as wbc|ebc<<4|sbc<<8|nbc<<12.

    ``None`` is a tile-local launch (0), cuda_advection5.py:825-846.
    """
    if lateral_sides is None:
        return 0
    sides = tuple(lateral_sides)
    if len(sides) != 4 or any(
        isinstance(v, bool) or not isinstance(v, int) or v not in (1, 2, 3, 4) for v in sides
    ):
        raise ValueError("lateral_sides must be CM1's (wbc, ebc, sbc, nbc), each 1, 2, 3 or 4")
    return sides[0] | (sides[1] << 4) | (sides[2] << 8) | (sides[3] << 12)
