"""Vertical coordinate profiles and the RK3/acoustic cadence (host side).

* ``terrain_coordinate_profiles`` follows ``cuda_terrain_open_boundary.py:486-549``:
  binary32 ``sigma``/``sigmaf`` with CM1 indices, ``rds``/``rdsf`` as
  reciprocal spacings and the ground/top quadratic extrapolation weights
  from Fornberg's finite-difference weight recursion (Fornberg 1988, Math.
  Comp. 51:699-706), as ``cm1_vertical_grid.py:387-454`` applies it.
* ``sound3_stage_schedule_f32`` / ``rk_stage_dt_f32`` follow
  ``cuda_sound3_column.py:391-424``: the small-step count and ``dts`` of each
  RK stage (2+3+6 at nsound=6, 3+4+8 at 8, 4+5+10 at 10).

The device copy of a profile is one float32 buffer laid out as
``sigma | sigmaf | rds | rdsf`` (``nk+2`` words each) then ``cgs[3] | cgt[3]``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

F = np.float32


def sound3_stage_schedule_f32(dt: float, nsound: int, nrk: int) -> tuple[int, float]:
    if isinstance(nsound, bool) or not isinstance(nsound, int) or nsound < 4:
        raise ValueError("nsound must be an integer >= 4")
    if nsound % 2:
        raise ValueError("native diagnostic cadence requires even nsound")
    if isinstance(nrk, bool) or not isinstance(nrk, int) or nrk not in (1, 2, 3):
        raise ValueError("nrk must be one of 1, 2, 3")
    dt32 = F(dt)
    if nrk == 1:
        nloop = int(np.floor(F(nsound) / F(3.0) + F(0.5)))
        dts = F(dt32 / F(F(nloop) * F(3.0)))
        if dts > F(dt32 / F(nsound)):
            nloop += 1
            dts = F(dt32 / F(F(nloop) * F(3.0)))
    elif nrk == 2:
        nloop = nsound // 2
        dts = F(dt32 / F(nsound))
    else:
        nloop = nsound
        dts = F(dt32 / F(nsound))
    return nloop, float(dts)


def rk_stage_dt_f32(dt: float, nrk: int) -> float:
    if nrk not in (1, 2, 3):
        raise ValueError("nrk must be one of 1, 2, 3")
    return float(F(F(dt) / F(4 - nrk)))


def small_steps_per_large_step(dt: float, nsound: int) -> int:
    return sum(sound3_stage_schedule_f32(dt, nsound, nrk)[0] for nrk in (1, 2, 3))


def _fornberg(alpha: list[float], x0: float) -> list[list[list[float]]]:
    # weights delta[n][m+1][nu] of the m-th derivative at x0 over nodes 0..n
    d = [[[0.0] * 4 for _ in range(6)] for _ in range(4)]
    d[0][1][0] = 1.0
    c1 = 1.0
    for n in range(1, 4):
        c2 = 1.0
        for nu in range(n):
            c3 = alpha[n] - alpha[nu]
            c2 *= c3
            d[n - 1][n + 1][nu] = 0.0
            for m in range(0, min(n, 4) + 1):
                d[n][m + 1][nu] = ((alpha[n] - x0) * d[n - 1][m + 1][nu] - m * d[n - 1][m][nu]) / c3
        for m in range(0, min(n, 4) + 1):
            d[n][m + 1][n] = (c1 / c2) * (m * d[n - 1][m][n - 1] - (alpha[n - 1] - x0) * d[n - 1][m + 1][n - 1])
        c1 = c2
    return d


def _boundary_weights(sigma, sigmaf, nk: int, top: bool) -> np.ndarray:
    if top:
        spacing = F(sigmaf[nk + 1] - sigmaf[nk])
        origin = sigmaf[nk + 1]
        nodes = (sigma[nk], sigma[nk - 1], sigma[nk - 2], sigma[nk - 3])
    else:
        spacing = F(sigmaf[2] - sigmaf[1])
        origin = sigmaf[1]
        nodes = (sigma[1], sigma[2], sigma[3], sigma[4])
    if not np.isfinite(spacing) or float(spacing) <= 0.0:
        raise ValueError("extrapolation spacing must be positive")
    x0 = float(F(F(origin) / spacing))
    alpha = [float(F(F(v) / spacing)) for v in nodes]
    d = _fornberg(alpha, x0)
    return np.array([F(d[2][1][nu]) for nu in range(3)], np.float32)


@dataclass(frozen=True)
class VerticalProfiles:
    nk: int
    sigma: np.ndarray
    sigmaf: np.ndarray
    rds: np.ndarray
    rdsf: np.ndarray
    cgs: np.ndarray
    cgt: np.ndarray
    ztop: float  # binary32 model top used by the terrain height factor

    def packed(self) -> np.ndarray:
        return np.concatenate([self.sigma, self.sigmaf, self.rds, self.rdsf, self.cgs, self.cgt]).astype(np.float32)


def terrain_coordinate_profiles(nk: int, dz: float, *, sigma=None, sigmaf=None) -> VerticalProfiles:
    if isinstance(nk, bool) or not isinstance(nk, int) or nk < 4:
        raise ValueError("nk must be an integer >= 4")
    dz32 = F(dz)
    s = np.zeros(nk + 2, np.float32)
    sf = np.zeros(nk + 2, np.float32)
    if sigma is None:
        for k in range(1, nk + 1):
            s[k] = (F(k) - F(0.5)) * dz32
        for k in range(1, nk + 2):
            sf[k] = F(k - 1) * dz32
        ztop = float(F(nk * float(dz)))
    else:
        s[1 : nk + 1] = np.asarray(sigma, np.float32)[1 : nk + 1]
        sf[1 : nk + 2] = np.asarray(sigmaf, np.float32)[1 : nk + 2]
        ztop = float(sf[nk + 1])
    rds = np.zeros(nk + 2, np.float32)
    rdsf = np.zeros(nk + 2, np.float32)
    for k in range(2, nk + 1):
        rds[k] = F(1.0) / (s[k] - s[k - 1])
    for k in range(1, nk + 1):
        rdsf[k] = F(1.0) / (sf[k + 1] - sf[k])
    return VerticalProfiles(nk, s, sf, rds, rdsf, _boundary_weights(s, sf, nk, False),
                            _boundary_weights(s, sf, nk, True), ztop)
