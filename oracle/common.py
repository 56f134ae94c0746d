"""Shared binary32 helpers and layout facts for the oracle (test infrastructure).

Layout follows the reference: ``(z, y, x)`` C-order float32, winds and
scalars padded by ``ng = 3`` lateral cells (``cuda_sound3_open_boundary.py:36``
``CM1_OPEN_NGXY``), the pressure family by one cell, metrics unpadded per
``terrain_static_shapes`` (``cuda_terrain_open_boundary.py:429-459``).
"""

from __future__ import annotations

import os

import numpy as np

# The restatement's arithmetic type: binary32 (the reference's), or binary64
# when GM1_ORACLE_BINARY64=1 is set at import -- the same statements in double
# precision, the yardstick for the binary32 rounding error of a result.
F = np.float64 if os.environ.get("GM1_ORACLE_BINARY64") == "1" else np.float32
NG = 3

# Constant weights, each the binary64 quotient rounded once to binary32
# (cuda_advection5.py:170-213 spells them np.float32(a / b)).
C2_60 = F(2.0 / 60.0)
CM13_60 = F(-13.0 / 60.0)
C47_60 = F(47.0 / 60.0)
C27_60 = F(27.0 / 60.0)
CM3_60 = F(-3.0 / 60.0)
CM1_6 = F(-1.0 / 6.0)
C1_6 = F(1.0 / 6.0)
C2_6 = F(2.0 / 6.0)
C5_6 = F(5.0 / 6.0)
CM7_6 = F(-7.0 / 6.0)
C11_6 = F(11.0 / 6.0)
C13_12 = F(13.0 / 12.0)


def terrain_static_shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    """Metric shapes (cuda_terrain_open_boundary.py:429-459)."""

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


def padded_shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    """Wind/scalar/pressure carrier shapes (cuda_psolver3_terrain_open.py:162-187)."""

    ng = NG
    return {
        "u": (nk + 2, nj + 2 * ng, ni + 1 + 2 * ng),
        "v": (nk + 2, nj + 1 + 2 * ng, ni + 2 * ng),
        "w": (nk + 3, nj + 2 * ng, ni + 2 * ng),
        "s": (nk + 2, nj + 2 * ng, ni + 2 * ng),
        "p": (nk + 2, nj + 2, ni + 2),
    }


def flx3(s1, s2, s3):
    """adv_routines.F flx3 as cuda_advection5.py:170-173 spells it."""
    r = CM1_6 * s1
    r = r + C5_6 * s2
    return r + C2_6 * s3


def flx5(s1, s2, s3, s4, s5):
    """adv_routines.F flx5 as cuda_advection5.py:176-182 spells it."""
    r = C2_60 * s1
    r = r + CM13_60 * s2
    r = r + C47_60 * s3
    r = r + C27_60 * s4
    return r + CM3_60 * s5


def weno5(s1, s2, s3, s4, s5, weps: float):
    """WENO-Z siform 2 (cuda_advection5.py:185-213): f32 indicators, f64 weights."""
    d1 = (s1 - F(2.0) * s2) + s3
    e1 = (s1 - F(4.0) * s2) + F(3.0) * s3
    b1r = C13_12 * (d1 * d1) + F(0.25) * (e1 * e1)
    d2 = (s2 - F(2.0) * s3) + s4
    e2 = s2 - s4
    b2r = C13_12 * (d2 * d2) + F(0.25) * (e2 * e2)
    d3 = (s3 - F(2.0) * s4) + s5
    e3 = (F(3.0) * s3 - F(4.0) * s4) + s5
    b3r = C13_12 * (d3 * d3) + F(0.25) * (e3 * e3)
    b1 = np.asarray(b1r, dtype=np.float64)
    b2 = np.asarray(b2r, dtype=np.float64)
    b3 = np.asarray(b3r, dtype=np.float64)
    eps = float(weps)
    tau = np.abs(b1 - b3)
    with np.errstate(all="ignore"):
        q1 = np.fmin(1.0e30, tau / (b1 + eps))
        q2 = np.fmin(1.0e30, tau / (b2 + eps))
        q3 = np.fmin(1.0e30, tau / (b3 + eps))
        w1 = float(F(0.1)) * (1.0 + q1 * q1)
        w2 = float(F(0.6)) * (1.0 + q2 * q2)
        w3 = float(F(0.3)) * (1.0 + q3 * q3)
    p1 = (C2_6 * s1 + CM7_6 * s2) + C11_6 * s3
    p2 = (CM1_6 * s2 + C5_6 * s3) + C2_6 * s4
    p3 = (C2_6 * s3 + C5_6 * s4) + CM1_6 * s5
    with np.errstate(all="ignore"):
        num = (w1 * p1.astype(np.float64) + w2 * p2.astype(np.float64)) + w3 * p3.astype(np.float64)
        out = num / ((w1 + w2) + w3)
    return out.astype(F)


def upstrpd(s1, s2, s3, weps: float):
    """Positive-definite lower closure (cuda_advection5.py:244-253)."""
    dd = s2 - s1
    se = np.copysign(F(abs(F(weps))), dd).astype(F)
    with np.errstate(all="ignore"):
        rr = (s3 - s2) / (se + dd)
    cand = C1_6 + C2_6 * rr
    phi = np.maximum(F(0.0), np.minimum(rr, np.minimum(cand, F(1.0))))
    return s2 + phi * (s2 - s1)


def recon5(s1, s2, s3, s4, s5, use_weno: bool, weps: float):
    return weno5(s1, s2, s3, s4, s5, weps) if use_weno else flx5(s1, s2, s3, s4, s5)


def lateral_code(lateral_sides) -> int:
    """CM1 (wbc, ebc, sbc, nbc) packed as wbc|ebc<<4|sbc<<8|nbc<<12; None is 0."""
    if lateral_sides is None:
        return 0
    w, e, s, n = (int(v) for v in lateral_sides)
    return w | (e << 4) | (s << 8) | (n << 12)
