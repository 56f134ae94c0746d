"""Synthetic initial states for the benchmark and parity configurations (host side).

The reference builds its analytic storm environments in Rust
(``rust/cm1-state/src/storm_initialization.rs:40-330``, a Weisman-Klemp
hydrostatic fixed point) and perturbs them with a cos^2 warm bubble
(``initial_perturbation.rs:448-467``) and seeded noise
(``ensemble_branch.py:116-161``).  This module writes its own NumPy
generator of the same kind of state:

* potential temperature theta(z) = 300 + 43 (z/12 km)^1.25 K below 12 km,
  isothermal at 213 K above (the Weisman-Klemp form); surface pressure
  1000 hPa; Exner function integrated hydrostatically in binary64;
* moist soundings: relative humidity 0.9 below 1.2 km falling linearly to
  0.25 at 12 km, surface qv capped at 14 g/kg, saturation from Bolton's
  (1980) vapour-pressure fit, virtual potential temperature in the balance;
* winds: the configuration's analytic profile;
* bubble: amplitude * cos^2(pi*beta/2) for beta < 1 with beta the scaled
  distance from the centre;
* noise: splitmix64 (Steele, Lea & Flood 2014) uniform in [-a, a) in k, j, i
  order, the stream the reference's ensemble branch uses.

Everything is computed in binary64 and rounded once to binary32.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .grid import VerticalLevels, VerticalProfiles, build_vertical_levels, terrain_coordinate_profiles

F = np.float32
NG = 3
G = 9.81
RD = 287.04
RV = 461.5
CP = 1005.7
P00 = 1.0e5
EPS = RD / RV

_GOLDEN = 0x9E3779B97F4A7C15
_MIX1 = 0xBF58476D1CE4E5B9
_MIX2 = 0x94D049BB133111EB


def splitmix64_words(seed: int, count: int) -> np.ndarray:
    """``count`` consecutive splitmix64 outputs after ``seed`` (uint64)."""
    if seed < 0 or seed >= 1 << 64 or count < 1:
        raise ValueError("seed must be a uint64 and count >= 1")
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * np.uint64(_GOLDEN)
        z = z ^ (z >> np.uint64(30))
        z = z * np.uint64(_MIX1)
        z = z ^ (z >> np.uint64(27))
        z = z * np.uint64(_MIX2)
        return z ^ (z >> np.uint64(31))


def uniform_noise(seed: int, amplitude: float, shape: tuple[int, ...]) -> np.ndarray:
    """amplitude*(2u-1) with u the top 53 bits of each splitmix64 word."""
    n = int(np.prod(shape))
    u = (splitmix64_words(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return (float(amplitude) * (2.0 * u - 1.0)).astype(np.float32).reshape(shape)


@dataclass
class CaseConfig:
    """One benchmark/parity case: grid, time step, family and environment."""

    name: str
    ni: int
    nj: int
    nk: int
    dx: float
    dy: float
    dz: float
    dt: float
    nsound: int
    lateral_sides: tuple[int, int, int, int]
    moist: bool
    wind: str  # "linear15" or "hodograph"
    bubbles: tuple[tuple[float, float, float, float], ...]  # (amp K, x m, y m, z m)
    bubble_radius: tuple[float, float] = (10000.0, 1400.0)
    noise: float = 0.05
    seed: int = 1
    # CM1 stretch_z=1 parameters (ztop, str_bot, str_top, dz_bot, dz_top); None is the uniform column of
    # spacing dz.  ``dz`` stays the metric reference of mh = dz/(zf(k+1)-zf(k)).
    stretched: tuple[float, float, float, float, float] | None = None
    extra: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.ni * self.nj * self.nk

    def levels(self) -> VerticalLevels:
        if self.stretched is None:
            return build_vertical_levels(stretch_z=0, nk=self.nk, dz=self.dz)
        ztop, str_bot, str_top, dz_bot, dz_top = self.stretched
        return build_vertical_levels(stretch_z=1, nk=self.nk, dz=self.dz, ztop=ztop, str_bot=str_bot,
                                     str_top=str_top, dz_bot=dz_bot, dz_top=dz_top)


def dry_c1(ni=64, nj=64, nk=40) -> CaseConfig:
    """64x64x40, dx 1 km, dz 500 m, periodic, dt 6 s, 6 substeps, dry."""
    return CaseConfig("dry-periodic-64x64x40", ni, nj, nk, 1000.0, 1000.0, 500.0, 6.0, 6, (1, 1, 1, 1), False,
                      "linear15", ((2.0, ni * 500.0, nj * 500.0, 1400.0),), noise=0.05, seed=1)


def moist_c2(ni=320, nj=320, nk=80) -> CaseConfig:
    """320x320x80, dx=dz=250 m, open-radiative, dt 1.5 s, 8 substeps, Kessler."""
    return CaseConfig("moist-kessler-320x320x80", ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 8, (2, 2, 2, 2), True,
                      "hodograph", ((3.0, ni * 125.0, nj * 125.0, 1400.0),), noise=0.1, seed=2)


def large_stretched_c3(ni=640, nj=768, nk=80) -> CaseConfig:
    """640x768x80, dx 100 m, stretch_z=1 column 20 m -> 280 m, two bubbles 8 km apart, dt 0.6 s, 10 substeps.

    CM1's stretch_z=1 coordinate with dz_bot 20 m up to 200 m, a quadratic
    stretch to dz_top 280 m at 2450 m and 280 m layers to a 17.85 km top: the
    feasible stretch_z=1 column of 80 levels nearest to "20 m at the ground,
    280 m aloft, top near 18 km" (20 m up to 500 m and an 18 km top do not fit
    80 levels with 280 m aloft).
    """
    x0, y0 = ni * 50.0, nj * 50.0
    return CaseConfig("moist-kessler-stretched-640x768x80", ni, nj, nk, 100.0, 100.0, 100.0, 0.6, 10, (2, 2, 2, 2),
                      True, "hodograph", ((3.0, x0 - 4000.0, y0, 1400.0), (3.0, x0 + 4000.0, y0, 1400.0)),
                      bubble_radius=(4000.0, 1400.0), noise=0.1, seed=3,
                      stretched=(17850.0, 200.0, 2450.0, 20.0, 280.0))


def adjoint_c5(ni=192, nj=192, nk=64) -> CaseConfig:
    """192x192x64, dx 500 m, dz 250 m, periodic, dt 3 s, 6 substeps, dry; the dry case's bubble, noise 0.1 K."""
    return CaseConfig(f"dry-periodic-adjoint-{ni}x{nj}x{nk}", ni, nj, nk, 500.0, 500.0, 250.0, 3.0, 6, (1, 1, 1, 1),
                      False, "linear15", ((2.0, ni * 250.0, nj * 250.0, 1400.0),), noise=0.1, seed=5)


ENSEMBLE_SIZES = (128, 192, 256, 320)
ENSEMBLE_LEVELS = (64, 80)


def ensemble_members(seed: int = 2024, count: int = 24) -> list[CaseConfig]:
    """Perturbed moist warm bubbles on uneven grids, every draw from ``seed`` (splitmix64).

    Per member: ni, nj from {128, 192, 256, 320}, nk from {64, 80}, dx = dy = dz
    = 250 m, open-radiative laterals, dt 1.5 s with 8 substeps; bubble
    amplitude U(1.5, 3.5) K, centre offset U(-5, 5) km in x and y from the
    domain centre, the moisture profile scaled by U(0.9, 1.1), noise
    U(-0.1, 0.1) K seeded per member.
    """
    words = splitmix64_words(seed, 8 * count).reshape(count, 8)
    unit = (words >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    out = []
    for m in range(count):
        w, u = words[m], unit[m]
        ni, nj = ENSEMBLE_SIZES[int(w[0] % 4)], ENSEMBLE_SIZES[int(w[1] % 4)]
        nk = ENSEMBLE_LEVELS[int(w[2] % 2)]
        amp = 1.5 + 2.0 * u[3]
        x0 = ni * 125.0 + (2.0 * u[4] - 1.0) * 5000.0
        y0 = nj * 125.0 + (2.0 * u[5] - 1.0) * 5000.0
        out.append(CaseConfig(f"member-{m:02d}-{ni}x{nj}x{nk}", ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 8,
                              (2, 2, 2, 2), True, "hodograph", ((amp, x0, y0, 1400.0),), noise=0.1,
                              seed=seed * 1000 + m, extra={"qv_scale": 0.9 + 0.2 * u[6]}))
    return out


def profiles_for(cfg: CaseConfig) -> VerticalProfiles:
    if cfg.stretched is None:
        return terrain_coordinate_profiles(cfg.nk, cfg.dz)
    return cfg.levels().profiles()


def _theta_env(z: np.ndarray) -> np.ndarray:
    ztr, thtr, ttr = 12000.0, 343.0, 213.0
    th = np.where(z <= ztr, 300.0 + 43.0 * (np.maximum(z, 0.0) / ztr) ** 1.25, thtr * np.exp(G / (CP * ttr) * (z - ztr)))
    return th


def _rh_env(z: np.ndarray) -> np.ndarray:
    return np.where(z <= 1200.0, 0.9, np.where(z >= 12000.0, 0.25, 0.9 - 0.65 * (z - 1200.0) / 10800.0))


def base_state(cfg: CaseConfig, prof: VerticalProfiles) -> dict[str, np.ndarray]:
    """Hydrostatic column at scalar and w levels (binary64 then binary32)."""
    nk = cfg.nk
    zs = prof.sigma[1 : nk + 1].astype(np.float64)
    zf = prof.sigmaf[1 : nk + 2].astype(np.float64)
    th = _theta_env(zs)
    qv = np.zeros(nk)
    # integrate Exner from the surface (pi=1 at 1000 hPa) at scalar levels; the first half
    # step averages theta_v with the surface's (300 K, saturated at 1000 hPa), as CM1's
    # base.F does (storm_initialization.rs:62-79, 281-286)
    t_sfc = 300.0
    es_sfc = min(611.2 * math.exp(17.67 * (t_sfc - 273.15) / (t_sfc - 29.65)), 0.5 * P00)
    qv_sfc = EPS * es_sfc / (P00 - es_sfc)
    thv_sfc = 300.0 * (1.0 + qv_sfc / EPS) / (1.0 + qv_sfc)
    pi = np.empty(nk)
    for it in range(20):
        thv = th * (1.0 + qv / EPS) / (1.0 + qv)
        pi[0] = 1.0 - G * zs[0] / (CP * 0.5 * (thv_sfc + thv[0]))
        for k in range(1, nk):
            pi[k] = pi[k - 1] - G / CP * (zs[k] - zs[k - 1]) * 2.0 / (thv[k - 1] + thv[k])
        if not cfg.moist:
            break
        t = th * pi
        p = P00 * pi ** (CP / RD)
        es = np.minimum(611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65)), 0.5 * p)
        qvs = EPS * es / (p - es)
        qv = np.minimum(_rh_env(zs) * qvs, 0.014) * float(cfg.extra.get("qv_scale", 1.0))
    thv = th * (1.0 + qv / EPS) / (1.0 + qv)
    p = P00 * pi ** (CP / RD)
    rho = p / (RD * thv * pi)
    # w-level density by linear interpolation in height, extrapolated at the ends
    rhof = np.interp(zf, zs, rho)
    rhof[0] = rho[0] + (rho[0] - rho[1]) * (zs[0] - zf[0]) / (zs[1] - zs[0])
    rhof[-1] = rho[-1] + (rho[-1] - rho[-2]) * (zf[-1] - zs[-1]) / (zs[-1] - zs[-2])
    c2 = np.zeros(nk + 1)
    c2[1:nk] = (zf[1:nk] - zs[:-1]) / (zs[1:] - zs[:-1])
    c1 = 1.0 - c2
    return {"th0": th.astype(F), "pi0": pi.astype(F), "qv0": qv.astype(F), "rho0": rho.astype(F),
            "rf0": rhof.astype(F), "c1": c1.astype(F), "c2": c2.astype(F), "zs": zs, "zf": zf,
            "binary64": {"th": th, "pi": pi, "qv": qv, "thv_sfc": thv_sfc}}


def metrics_for(cfg: CaseConfig, base: dict[str, np.ndarray]) -> dict[str, np.ndarray]:
    """Flat-terrain metric carriers in the reference's terrain_static_shapes layout."""
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    m: dict[str, np.ndarray] = {}
    col = lambda v: np.broadcast_to(np.asarray(v, np.float32)[:, None, None], (len(v), nj, ni)).copy()  # noqa: E731
    m["rho0_3d"] = col(base["rho0"])
    m["rr0"] = (F(1.0) / m["rho0_3d"]).astype(F)
    m["rf0"] = col(base["rf0"])
    m["rrf0"] = (F(1.0) / m["rf0"]).astype(F)
    m["c1"] = col(base["c1"])
    m["c2"] = col(base["c2"])
    m["zh"] = col(base["zs"].astype(F))
    m["zf"] = col(base["zf"].astype(F))
    for name, shape in (("mh", (nk, nj, ni)), ("rmh", (nk, nj, ni)), ("mf", (nk + 1, nj, ni)), ("rmf", (nk + 1, nj, ni)),
                        ("gz", (nj, ni)), ("rgz", (nj, ni)), ("gzu", (nj, ni + 1)), ("rgzu", (nj, ni + 1)),
                        ("gzv", (nj + 1, ni)), ("rgzv", (nj + 1, ni))):
        m[name] = np.ones(shape, np.float32)
    if cfg.stretched is not None:
        levels = cfg.levels()
        for name, lo, hi in (("mh", 1, nk + 1), ("rmh", 1, nk + 1), ("mf", 1, nk + 2), ("rmf", 1, nk + 2)):
            m[name] = col(getattr(levels, name)[lo:hi])
    for name, shape in (("terrain_msl", (nj, ni)), ("dzdx", (nj, ni)), ("dzdy", (nj, ni)), ("gx", (nk + 1, nj, ni)),
                        ("gxu", (nk + 1, nj, ni + 1)), ("gy", (nk + 1, nj, ni)), ("gyv", (nk + 1, nj + 1, ni))):
        m[name] = np.zeros(shape, np.float32)
    return m


def wind_profile(cfg: CaseConfig, z: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    if cfg.wind == "linear15":
        return np.minimum(z / 6000.0, 1.0) * 15.0, np.zeros_like(z)
    # quarter circle of radius 10 m/s below 2 km, straight shear to 30 m/s at 6 km
    ang = np.clip(z / 2000.0, 0.0, 1.0) * (0.5 * math.pi)
    u = np.where(z < 2000.0, 10.0 - 10.0 * np.cos(ang), 10.0 + 20.0 * np.clip((z - 2000.0) / 4000.0, 0.0, 1.0))
    v = np.where(z < 2000.0, 10.0 * np.sin(ang), 10.0)
    return u, v


def _bubble_box(x, y, zs, xc, yc, zc, xr, zr) -> tuple[slice, slice, slice]:
    """(k, j, i) slices holding every point with |x - xc| <= xr, |y - yc| <= xr, |z - zc| <= zr."""

    def span(c: np.ndarray, centre: float, radius: float) -> slice:
        idx = np.nonzero(np.abs(c - centre) <= radius)[0]
        return slice(int(idx[0]), int(idx[-1]) + 1) if idx.size else slice(0, 0)

    return span(zs, zc, zr), span(y, yc, xr), span(x, xc, xr)


def initial_fields(cfg: CaseConfig, prof: VerticalProfiles, base: dict[str, np.ndarray]) -> dict[str, np.ndarray]:
    """theta', u, v, w, pressure carriers and moisture of the initial state."""
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    zs = base["zs"]
    x = (np.arange(ni) + 0.5) * cfg.dx
    y = (np.arange(nj) + 0.5) * cfg.dy
    th3d = np.zeros((nk, nj, ni), np.float64)
    xr, zr = cfg.bubble_radius
    for amp, xc, yc, zc in cfg.bubbles:
        # beta < 1 only inside the bubble's bounding box; outside it the perturbation is +0 either way
        b = _bubble_box(x, y, zs, xc, yc, zc, xr, zr)
        beta = np.sqrt(((x[None, None, b[2]] - xc) / xr) ** 2 + ((y[None, b[1], None] - yc) / xr) ** 2
                       + ((zs[b[0], None, None] - zc) / zr) ** 2)
        th3d[b] += np.where(beta < 1.0, amp * np.cos(0.5 * math.pi * beta) ** 2, 0.0)
    th3d = th3d.astype(F) + uniform_noise(cfg.seed, cfg.noise, (nk, nj, ni))
    u, v = wind_profile(cfg, zs)
    out = {"th3d": th3d.astype(F), "u_profile": u.astype(F), "v_profile": v.astype(F)}
    if cfg.moist:
        q = np.zeros((3, nk, nj, ni), np.float32)
        q[0] = np.broadcast_to(base["qv0"][:, None, None], (nk, nj, ni))
        rh = cfg.extra.get("bubble_rh")
        if rh is not None:
            # raise the bubble's vapour to rh times saturation at its temperature (Bolton 1980)
            for amp, xc, yc, zc in cfg.bubbles:
                b = _bubble_box(x, y, zs, xc, yc, zc, xr, zr)
                pi = base["pi0"].astype(np.float64)[b[0], None, None]
                t = (base["th0"].astype(np.float64)[b[0], None, None] + th3d[b].astype(np.float64)) * pi
                p = P00 * pi ** (CP / RD)
                es = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))
                qvs = EPS * es / (p - es)
                beta = np.sqrt(((x[None, None, b[2]] - xc) / xr) ** 2 + ((y[None, b[1], None] - yc) / xr) ** 2
                               + ((zs[b[0], None, None] - zc) / zr) ** 2)
                qb = q[0][b]
                q[0][b] = np.where(beta < 1.0, np.maximum(qb, (rh * qvs).astype(F)), qb)
        out["q3d"] = q
    return out


def initial_thermodynamics(base: dict[str, np.ndarray], th3d: np.ndarray, qv: np.ndarray | None) -> dict[str, np.ndarray]:
    """Pressure and density of the initial state (pi' = 0) by the moist equation of state.

    prs = p00 pi0^(cp/rd); rho = prs / ((th0 + th') pi0 (rd + rv max(qv, 0))), in binary32.
    """
    nk, nj, ni = th3d.shape
    pi0_k = base["pi0"].astype(F)[:, None, None]
    th0_k = base["th0"].astype(F)[:, None, None]
    prs_k = (np.power(pi0_k.astype(np.float64), float(F(CP) / F(RD))) * P00).astype(F)
    vap = np.zeros_like(th3d) if qv is None else np.maximum(qv, F(0.0))
    rho = prs_k / (((th0_k + th3d) * pi0_k) * (F(RD) + F(RV) * vap))
    shape = (nk, nj, ni)
    return {"prs": np.broadcast_to(prs_k, shape).astype(F), "rho": rho.astype(F, copy=False),
            "pi0_3d": np.broadcast_to(pi0_k, shape).astype(F), "th0_3d": np.broadcast_to(th0_k, shape).astype(F)}
