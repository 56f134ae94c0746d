"""Choice of the acoustic column pass per grid shape, by timing it (one GPU).

The column pass has word-identical variants (``psolver3.COLUMN_PASSES``,
``psolver3.FLAT_COLUMN_PASSES``): 32- or 16-column workgroups (the tile
width across i), LDS-staged or pipelined (loads issued a level ahead), and on
the flat column the fused forms that step u and v inside the pipelined pass.
Which is fastest depends on the column count and height, so
``tune_column_pass`` runs each candidate's acoustic RK stage on a scratch
copy of a large-step state, times its small steps (pressure gradient and
column pass) with the library's event-pair timer, and returns the fastest.  Results are kept per (ni, nj, nk, sides, stretched) for the
process, and in a JSON file when ``GM1_AUTOTUNE_CACHE`` names one.

Choosing a variant never changes a word of the result, so a tuned run
equals an untuned one bit for bit.
"""

from __future__ import annotations

import json
import os
import warnings
from pathlib import Path
from typing import Any

import torch

from . import timing
from .psolver3 import FLAT_COLUMN_PASSES, TerrainOpenSoundAblations, launch_cuda_psolver3_terrain_open_stage
from .stretched import launch_cuda_gm1_flat_acoustic_stage

CANDIDATES = ("pipelined", "staged", "pipelined16", "staged16")
FLAT_CANDIDATES = tuple(FLAT_COLUMN_PASSES)
_memo: dict[str, dict[str, Any]] = {}


def _library_tag() -> str:
    """The built library's source hash and arch (a cache written by another build is not reused)."""
    from .build import RECEIPT

    try:
        r = json.loads(RECEIPT.read_text())
        return f"{r.get('arch', '?')}:{r.get('sources_sha256', '?')[:16]}"
    except (OSError, ValueError):
        return "unknown"


def _key(st) -> str:
    cfg = st.cfg
    device = torch.cuda.get_device_name() if torch.cuda.is_available() else "no-device"
    return json.dumps([_library_tag(), device, cfg.ni, cfg.nj, cfg.nk,
                       list(cfg.lateral_sides), st.flat is not None])


def _candidates(flat: bool) -> tuple[str, ...]:
    return FLAT_CANDIDATES if flat else CANDIDATES


def _valid(entry: Any, flat: bool) -> bool:
    return isinstance(entry, dict) and entry.get("choice") in _candidates(flat)


def _cache_path() -> Path | None:
    p = os.environ.get("GM1_AUTOTUNE_CACHE")
    return Path(p) if p else None


def _load_cache() -> None:
    path = _cache_path()
    if path and path.exists():
        try:
            _memo.update(json.loads(path.read_text()))
        except (OSError, ValueError) as exc:  # an unreadable cache is re-measured, not trusted
            warnings.warn(f"ignoring autotune cache {path}: {exc}", RuntimeWarning, stacklevel=2)


def _save_cache() -> None:
    path = _cache_path()
    if path:
        path.write_text(json.dumps(_memo, indent=1, sort_keys=True))


def _stage(st, a, nrk: int, column_pass: str) -> None:
    cfg = st.cfg
    kw = dict(ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, nrk=nrk, dt=cfg.dt, nsound=cfg.nsound, dx=cfg.dx, dy=cfg.dy,
              stream=st.stream, lateral_sides=cfg.lateral_sides, column_pass=column_pass)
    if st.flat is not None:
        launch_cuda_gm1_flat_acoustic_stage(a, st.metrics, vertical=st.flat, alternates=st.alt, **kw)
    else:
        launch_cuda_psolver3_terrain_open_stage(a, st.metrics, dz=cfg.dz, profiles=st.prof,
                                                ablations=TerrainOpenSoundAblations(),
                                                terrain_slopes_zero=st.slopes_zero, **kw)


def tune_column_pass(st, repeats: int = 2, candidates=None) -> dict[str, Any]:
    """Fastest column pass for ``st``'s grid: {"choice", "us": {variant: mean small step in us}}.

    A small step is timed from its pressure-gradient launch to the end of its
    column pass (one launch for the fused forms), so every form is charged the
    same work.
    """
    if not _memo:
        _load_cache()
    key = _key(st)
    if key in _memo and _valid(_memo[key], st.flat is not None):
        return _memo[key]
    allowed = _candidates(st.flat is not None)
    candidates = allowed if candidates is None else tuple(candidates)
    for c in candidates:
        if c not in allowed:
            raise ValueError(f"unknown column pass {c!r}")
    stream = st.stream if st.stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(stream):  # the clones are ordered after the state's own pending work
        scratch = {k: (v.clone() if isinstance(v, torch.Tensor) else v) for k, v in st.a.items()}
    torch.cuda.synchronize()
    us = {}
    for c in candidates:
        _stage(st, scratch, 3, c)  # first launch of a variant: code object load, caches
        torch.cuda.synchronize()
        # a small step's pressure-gradient and column work, one pass in the fused forms and two in the others
        timing.arm("k_snd_small_step", 64 * repeats + 8)
        for _ in range(repeats):
            _stage(st, scratch, 3, c)
        torch.cuda.synchronize()
        total_ms, launches = timing.collect()
        us[c] = total_ms * 1e3 / max(launches, 1)
    timing.disarm()
    del scratch
    choice = min(us, key=us.get)
    _memo[key] = {"choice": choice, "us": us}
    _save_cache()
    return _memo[key]


def autotune_state(st, **kw) -> str:
    """Set ``st.column_pass`` to the tuned choice for its grid and return it."""
    st.column_pass = tune_column_pass(st, **kw)["choice"]
    return st.column_pass


__all__ = ["CANDIDATES", "FLAT_CANDIDATES", "autotune_state", "tune_column_pass"]
