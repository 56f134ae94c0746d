"""One CM1 large step composed from the oracle's operators (test infrastructure).

Follows the source order of ``cuda_gm1_forecast.launch_cuda_gm1_large_step``
(:1569-1627), ``launch_cuda_gm1_rk_stage`` (:1241-1345) and
``finish_cuda_gm1_large_step`` / ``_finish_solve3`` (:1414-1568), i.e.
``gm1_runtime.large_step_schedule`` (``gm1_runtime.py:1969-2007``):

  solve1   held carriers zeroed (no closure); held compose (moist)
  nrk=1..3 stage terms -> held momentum seed (+ bcwsfc, bc2d at nrk>=2)
           -> momentum advection -> buoyancy -> acoustic stage (stage fills)
           -> getdiv / getdivx -> theta RK stage -> moisture RK stage
           (nrk=3: positive-definite chain)
  finish   Kessler transition (moist) -> solve3: ppx += pp3d, ppi = pp3d,
           A <- current, density

The composition itself has no CPU form in the reference (its forecast
composer is device-only); each operator is pinned on its own against the
reference's golden vectors or closed forms, and the composition is pinned by
invariants (a state at rest stays at rest, periodic translation
equivariance).  For comparison with the device the Kessler phase runs with
binary64-rounded transcendentals and may take the dry-mass correction the
device computed (its ledger is a reduction in another order).
"""

from __future__ import annotations

import numpy as np

from . import boundary as bc
from . import kessler as ok
from . import stage as st
from .advection5 import cpu_scalar_advection5_inplace
from .common import F, NG
from .momentum5 import cpu_terrain_open_momentum_advection5_inplace
from .pdef import pdef_advection
from .sound import cpu_flat_stage_inplace, cpu_psolver3_stage_inplace

THETA_WEPS = float(F(10.0) * F(1.0e-18))
MOISTURE_WEPS = float(F(1.0e-2) * F(1.0e-18))
KESSLER_HELD = ("qpten", "qtten", "qvten", "qcten")


def _flux_scratch(ni, nj, nk):
    return {"x": np.zeros((nk + 2, nj + 2 * NG, ni + 1 + 2 * NG), F),
            "y": np.zeros((nk + 2, nj + 1 + 2 * NG, ni + 2 * NG), F),
            "z": np.zeros((nk + 3, nj + 2 * NG, ni + 2 * NG), F)}


def _species6(q3d):
    if q3d is None:
        return None
    q = np.zeros((6,) + q3d.shape[1:], F)
    q[: q3d.shape[0]] = q3d
    return q


def rk_stage(a, m, base, prof, cfg, nrk, *, zt, flat=None):
    """One RK stage; ``flat`` (mh, mf, rdz) runs the flat stretched column's operators.

    The flat column is cuda_cartesian_dynamics.py's stretched lane
    (:1070-1100, 1150-1192, 1213-1224, 1242-1262): no bcwsfc/bc2d, the flat
    momentum, buoyancy, acoustic stage, getdiv/getdivx and scalar texts.
    """
    ni, nj, nk = cfg["ni"], cfg["nj"], cfg["nk"]
    dx, dy, dz, dt = cfg["dx"], cfg["dy"], cfg["dz"], cfg["dt"]
    sides = tuple(cfg["lateral_sides"])
    q3d = a.get("q3d")
    st.stage_terms(a, base, a["th3d"], _species6(q3d), ni=ni, nj=nj, nk=nk, sides=sides)
    st.seed_held((a["uten1"], a["vten1"], a["wten1"]), (a["uten"], a["vten"], a["wten"]), ni=ni, nj=nj, nk=nk)
    if nrk >= 2 and flat is None:
        bc.bcwsfc(a, m, ni=ni, nj=nj, nk=nk)
        bc.bc2d_open_w_surface(a["w3d"], ni=ni, nj=nj)
    cpu_terrain_open_momentum_advection5_inplace(a, m, a["divx"], ni=ni, nj=nj, nk=nk, nrk=nrk, dx=dx, dy=dy, dz=dz,
                                                 lateral_sides=sides, flat=flat)
    if flat is None:
        st.buoyancy(a, m, ni=ni, nj=nj, nk=nk)
        cpu_psolver3_stage_inplace(a, m, ni=ni, nj=nj, nk=nk, dt=dt, nsound=cfg["nsound"], nrk=nrk, dx=dx, dy=dy,
                                   dz=dz, lateral_sides=sides, prof=prof, zt=zt)
        st.getdiv(a, m, a["divergence"], a["theta_div"], a["thterm"], ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
        st.getdivx(a, m, a["divx"], ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    else:
        st.buoyancy_flat(a, m, ni=ni, nj=nj, nk=nk)
        cpu_flat_stage_inplace(a, m, flat, ni=ni, nj=nj, nk=nk, dt=dt, nsound=cfg["nsound"], nrk=nrk, dx=dx, dy=dy,
                               lateral_sides=sides)
        st.getdiv_flat(a, a["divergence"], a["theta_div"], a["thterm"], flat, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy)
        st.getdivx_flat(a, m, a["divx"], flat, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy)
    dttmp = F(F(dt) / F(4 - nrk))
    padded = np.zeros((nk + 2, nj + 2 * NG, ni + 2 * NG), F)
    adv = np.zeros((nk, nj, ni), F)
    st.pack_theta(base["th0"], a["th3d"], padded, ni=ni, nj=nj, nk=nk)
    bc.fill(padded, bc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
    cpu_scalar_advection5_inplace(padded, a, m, a["divx"], adv, _flux_scratch(ni, nj, nk), ni=ni, nj=nj, nk=nk,
                                  nrk=nrk, dx=dx, dy=dy, dz=dz, weps=THETA_WEPS, lateral_sides=sides, flat=flat)
    a["th3d"][...] = st.compose_theta(a["tha"], a["thten1"], a["theta_div"], adv, dttmp)
    if q3d is None:
        return
    for sp in range(q3d.shape[0]):
        padded = np.zeros((nk + 2, nj + 2 * NG, ni + 2 * NG), F)
        st.pack_scalar(q3d[sp], padded, ni=ni, nj=nj, nk=nk)
        bc.fill(padded, bc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
        weps = float(F(F(1.0e-2) * F(1.0e-18)))
        if nrk < 3:
            adv = np.zeros((nk, nj, ni), F)
            cpu_scalar_advection5_inplace(padded, a, m, a["divx"], adv, _flux_scratch(ni, nj, nk), ni=ni, nj=nj,
                                          nk=nk, nrk=nrk, dx=dx, dy=dy, dz=dz, weps=weps, lateral_sides=sides,
                                          flat=flat)
        else:
            adv = pdef_advection(a["qa"][sp], padded, a, m, a["divx"], ni=ni, nj=nj, nk=nk, dt=dttmp, dx=dx, dy=dy,
                                 dz=dz, weps=weps, pdefweno=sp != 0, sides=sides, flat=flat)
        q3d[sp] = st.compose_moisture(a["qa"][sp], a["qten"][sp], adv, dttmp)


def kessler_arrays(a):
    return {"ruh": a["ruh"], "rvh": a["rvh"], "rmh": a["rmh"], "mh": a["mh"], "pi0": a["pi0_3d"],
            "th0": a["th0_3d"], "rho": a["rho"], "rr": a["rr"], "pp3d": a["pp_mass"], "th3d": a["th3d"],
            "prs": a["prs"], "qv": a["q3d"][0], "qc": a["q3d"][1], "qr": a["q3d"][2],
            "fall_velocity": a["fall_velocity"], "rain": a["rain"], "prate": a["prate"], "asq": a["asq"],
            "qbudget": a["qbudget"]}


def large_step(a, m, base, prof, cfg, *, zt, mass_correction=None, mass1=None, mass_hook=None, equation_set=2,
               budgets=True, flat=None):
    """Advance the state ``a`` by one large step in place.

    The dry-mass hook: ``mass_correction`` applies a given binary32 Exner
    correction; ``mass1`` computes it from the ledger (cpu_apmasscon_inplace);
    ``mass_hook`` is called with the Kessler arrays in the hook's place;
    none of them skips the hook.
    """
    ni, nj, nk = cfg["ni"], cfg["nj"], cfg["nk"]
    sides = tuple(cfg["lateral_sides"])
    moist = "q3d" in a
    for name in ("uten1", "vten1", "wten1", "thten1", "ppten"):
        a[name][...] = F(0.0)
    if moist:
        a["qten"][...] = F(0.0)
        rates = {"qpten": a["qpten"], "qtten": a["qtten"], "qvten": a["qvten"], "qcten": a["qcten"]}
        ppi = a["ppi"][1 : nk + 1, 1 : nj + 1, 1 : ni + 1]
        st.predict_and_compose_held(base, a["rho"], ppi, a["tha"], a["qa"], a["thten1"], a["qten"], a["ppten"],
                                    rates, cfg["dt"], ni=ni, nj=nj, nk=nk)
    for nrk in (1, 2, 3):
        rk_stage(a, m, base, prof, cfg, nrk, zt=zt, flat=flat)
    if moist:
        a["pp_mass"][...] = a["pp3d"][1 : nk + 1, 1 : nj + 1, 1 : ni + 1]
        hook = mass_hook
        if mass_correction is not None:
            hook = lambda arr: ok.apply_mass_correction(arr, F(mass_correction))  # noqa: E731
        elif mass1 is not None:
            hook = lambda arr: ok.apmasscon(arr, mass1)  # noqa: E731
        held = {n: a[n] for n in KESSLER_HELD}
        with ok.transcendentals("rounded"):
            ok.large_step_transition(kessler_arrays(a), held, dx=cfg["dx"], dy=cfg["dy"], dz=cfg["dz"],
                                     dt=cfg["dt"], equation_set=equation_set, hook=hook, budgets=budgets)
        a["pp3d"][1 : nk + 1, 1 : nj + 1, 1 : ni + 1] = a["pp_mass"]
    st.solve3_pressure(a, sides, ni=ni, nj=nj, nk=nk)
    a["ua"][...] = a["u3d"]
    a["va"][...] = a["v3d"]
    a["wa"][...] = a["w3d"]
    a["tha"][...] = a["th3d"]
    if moist:
        a["qa"][...] = a["q3d"]
        rr, rf = st.solve3_density(a["rho"], m["c1"], m["c2"])
        a["rr"][...] = rr
        a["rf"][...] = rf
