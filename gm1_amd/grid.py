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


# ----------------------------------------------------------------------------- CM1 vertical levels


class VerticalGridError(ValueError):
    """A vertical coordinate CM1 would stop on (param.F's feasibility checks)."""


@dataclass(frozen=True)
class VerticalLevels:
    """One CM1 vertical coordinate, CM1-indexed binary32 (cm1_vertical_grid.py:616-652).

    ``zf``/``sigmaf`` (nk+3), ``zh``/``sigma``/``mh``/``rmh`` (nk+2),
    ``mf``/``rmf`` (nk+3); ``mh = dz/(zf(k+1)-zf(k))`` and
    ``mf = dz/(zh(k)-zh(k-1))`` against the metric reference ``dz``.
    """

    stretch_z: int
    nk: int
    dz: float
    ztop: float
    zf: np.ndarray
    zh: np.ndarray
    sigmaf: np.ndarray
    sigma: np.ndarray
    mh: np.ndarray
    mf: np.ndarray
    rmh: np.ndarray
    rmf: np.ndarray

    @property
    def rdz(self) -> float:
        """param.F:6327's 1/dz of the metric reference, binary32."""
        return float(F(F(1.0) / F(self.dz)))

    @property
    def spacings(self) -> np.ndarray:
        return (self.zf[2 : self.nk + 2] - self.zf[1 : self.nk + 1]).astype(np.float64)

    def profiles(self) -> VerticalProfiles:
        """The terrain-coordinate profile set of these levels (sigma/sigmaf and Fornberg weights)."""
        return terrain_coordinate_profiles(self.nk, self.dz, sigma=self.sigma, sigmaf=self.sigmaf)

    def flat_column(self) -> np.ndarray:
        """Device packing of the flat column: ``mh`` (nk+2) then ``mf`` (nk+3)."""
        return np.concatenate([self.mh, self.mf]).astype(np.float32)


def _stretch_z1_derive(nk, dz, ztop, str_bot, str_top, dz_bot, dz_top):
    """param.F:7819-7892: segment counts, nominal spacing and the quadratic's c1/c2 (cuda_stretch_z1.py:130-201)."""
    dz, ztop, str_bot, str_top, dz_bot, dz_top = (F(x) for x in (dz, ztop, str_bot, str_top, dz_bot, dz_top))
    if not all(np.isfinite(x) for x in (dz, ztop, str_bot, str_top, dz_bot, dz_top)):
        raise VerticalGridError("stretch_z=1 parameters must be finite")
    if dz <= 0 or ztop <= 0 or dz_bot <= 0 or dz_top <= 0:
        raise VerticalGridError("dz, ztop, dz_bot and dz_top must be positive")
    if str_bot < 0 or str_top <= str_bot or str_top > ztop:
        raise VerticalGridError("require 0 <= str_bot < str_top <= ztop")
    nint = lambda x: int(np.floor(float(F(x)) + 0.5))  # noqa: E731
    nominal = F(F(0.5) * F(dz_bot + dz_top))
    nk1 = nint(F(str_bot / dz_bot))
    nk3 = nint(F(F(ztop - str_top) / dz_top))
    nk2 = nk - (nk1 + nk3)
    if nk1 < 0 or nk2 < 2 or nk3 < 0:
        raise VerticalGridError("stretch_z=1 requires nk1>=0, nk2>=2 and nk3>=0")
    if (abs(float(F(F(F(nk1) * dz_bot) - str_bot))) > 0.01
            or abs(float(F(F(F(nk2) * nominal) - F(str_top - str_bot)))) >= 0.01
            or abs(float(F(F(F(nk3) * dz_top) - F(ztop - str_top)))) > 0.01):
        raise VerticalGridError("stretch_z=1 segment depths fail the 0.01 m feasibility checks")
    c2 = F(F(nominal - dz_bot) / F(F(nominal * nominal) * F(nk2 - 1)))
    c1 = F(F(dz_bot / nominal) - F(c2 * nominal))
    test_len = F(F(F(dz_bot * F(nk1)) + F(nominal * F(nk2))) + F(dz_top * F(nk3)))
    if abs(float(F(ztop - test_len))) > float(F(F(1.0e-6) * ztop)):
        raise VerticalGridError("stretch_z=1 total depth fails the relative feasibility check")
    return nk1, nk2, nk3, nominal, c1, c2


def build_vertical_levels(*, stretch_z: int, nk: int, dz: float, ztop: float | None = None, str_bot: float = 0.0,
                          str_top: float = 0.0, dz_bot: float = 0.0, dz_top: float = 0.0) -> VerticalLevels:
    """CM1's stretch_z=0 (uniform) and stretch_z=1 (Wilhelmson-Chen) coordinates.

    cm1_vertical_grid.build_vertical_levels (:780-919) with
    cuda_stretch_z1.build_stretch_z1_grid (:204-263): stretch_z=0 runs the
    stretch_z=1 generator on its degenerate parameters (str_bot=0,
    str_top=ztop=dz*nk, dz_bot=dz_top=dz), which yields zf(k)=(k-1)*dz.
    """
    if isinstance(nk, bool) or not isinstance(nk, int) or nk < 4:
        raise VerticalGridError("nk must be an integer >= 4")
    if stretch_z == 0:
        top = float(F(F(dz) * F(nk)))
        if ztop is not None and float(F(ztop)) != top:
            raise VerticalGridError(f"stretch_z=0 fixes the model top at dz*nk = {top} m")
        ztop, str_bot, str_top, dz_bot, dz_top = top, 0.0, top, dz, dz
    elif stretch_z == 1:
        if ztop is None:
            raise VerticalGridError("stretch_z=1 needs ztop")
    elif stretch_z == 2:
        raise VerticalGridError("stretch_z=2 is refused, as CM1 r21.1 refuses it for this lane")
    else:
        raise VerticalGridError(f"stretch_z={stretch_z} is not built (0 and 1 are)")
    nk1, nk2, nk3, nominal, c1, c2 = _stretch_z1_derive(nk, dz, ztop, str_bot, str_top, dz_bot, dz_top)
    dz32, dz_bot32, dz_top32 = F(dz), F(dz_bot), F(dz_top)
    zf = np.zeros(nk + 3, np.float32)
    for k in range(1, nk1 + 2):  # param.F:7925-7927
        zf[k] = F(F(k - 1) * dz_bot32)
    for k in range(nk1 + 1, nk1 + nk2 + 2):  # param.F:7928-7931
        delta = F(k - 1 - nk1)
        factor = F(c1 + F(F(c2 * delta) * nominal))
        zf[k] = F(zf[nk1 + 1] + F(F(factor * delta) * nominal))
    for k in range(nk1 + nk2 + 2, nk + 2):  # param.F:7932-7934
        zf[k] = F(zf[k - 1] + dz_top32)
    zf[0] = F(-zf[2])
    zf[nk + 2] = F(zf[nk + 1] + F(zf[nk + 1] - zf[nk]))
    sigmaf = zf.copy()
    zh = np.empty(nk + 2, np.float32)
    mh = np.empty(nk + 2, np.float32)
    for k in range(nk + 2):  # param.F:8168-8171
        zh[k] = F(F(0.5) * F(zf[k + 1] + zf[k]))
        mh[k] = F(dz32 / F(zf[k + 1] - zf[k]))
    zh[0] = F(-zh[1])
    zh[nk + 1] = F(zh[nk] + F(F(2.0) * F(zf[nk + 1] - zh[nk])))
    mf = np.empty(nk + 3, np.float32)
    for k in range(1, nk + 2):  # param.F:8175-8179
        mf[k] = F(dz32 / F(zh[k] - zh[k - 1]))
    mf[0], mf[nk + 2] = mf[1], mf[nk + 1]
    sigma = np.empty(nk + 2, np.float32)
    for k in range(nk + 2):  # param.F:8189-8191
        sigma[k] = F(F(0.5) * F(sigmaf[k] + sigmaf[k + 1]))
    rmh = (F(1.0) / mh).astype(np.float32)
    rmf = (F(1.0) / mf).astype(np.float32)
    if not np.all(np.diff(zf[1 : nk + 2]) > 0):
        raise VerticalGridError("stretch_z=1 generated a non-positive layer depth")
    return VerticalLevels(stretch_z, nk, float(dz32), float(F(ztop)), zf, zh, sigmaf, sigma, mh, mf, rmh, rmf)
