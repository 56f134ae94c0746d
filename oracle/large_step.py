"""One CM1 large step composed from the oracle's operators (test infrastructure).

Follows the source order of ``cuda_gm1_forecast.launch_cuda_gm1_large_step``
(:1569-1627) and ``launch_cuda_gm1_rk_stage`` (:1241-1345), i.e.
``gm1_runtime.large_step_schedule`` (``gm1_runtime.py:1969-2007``):

  solve1   held carriers zeroed (no closure); held compose (moist only)
  nrk=1..3 stage terms -> held momentum seed (+ bcwsfc, bc2d at nrk>=2)
           -> momentum advection -> buoyancy -> acoustic stage (stage fills)
           -> getdiv / getdivx -> theta RK stage
  solve3   ppx += pp3d, ppi = pp3d, halos, A <- current

The composition itself has no CPU form in the reference (its forecast
composer is device-only); each operator here is pinned on its own against
the reference's golden vectors or closed forms, and the composition is
pinned by invariants (a resting isothermal-free state stays at rest, a
periodic domain's mass budget, translation equivariance).
"""

from __future__ import annotations

import numpy as np

from . import boundary as bc
from . import stage as st
from .advection5 import cpu_scalar_advection5_inplace
from .common import F, NG
from .momentum5 import cpu_terrain_open_momentum_advection5_inplace
from .sound import cpu_psolver3_stage_inplace

THETA_WEPS = float(F(10.0) * F(1.0e-18))


def _flux_scratch(ni, nj, nk):
    return {"x": np.zeros((nk + 2, nj + 2 * NG, ni + 1 + 2 * NG), np.float32),
            "y": np.zeros((nk + 2, nj + 1 + 2 * NG, ni + 2 * NG), np.float32),
            "z": np.zeros((nk + 3, nj + 2 * NG, ni + 2 * NG), np.float32)}


def rk_stage(a, m, base, prof, cfg, nrk, *, zt):
    ni, nj, nk = cfg["ni"], cfg["nj"], cfg["nk"]
    dx, dy, dz, dt = cfg["dx"], cfg["dy"], cfg["dz"], cfg["dt"]
    sides = tuple(cfg["lateral_sides"])
    st.stage_terms(a, base, a["th3d"], None, ni=ni, nj=nj, nk=nk, sides=sides)
    st.seed_held((a["uten1"], a["vten1"], a["wten1"]), (a["uten"], a["vten"], a["wten"]), ni=ni, nj=nj, nk=nk)
    if nrk >= 2:
        bc.bcwsfc(a, m, ni=ni, nj=nj, nk=nk)
        bc.bc2d_open_w_surface(a["w3d"], ni=ni, nj=nj)
    cpu_terrain_open_momentum_advection5_inplace(a, m, a["divx"], ni=ni, nj=nj, nk=nk, nrk=nrk, dx=dx, dy=dy, dz=dz,
                                                 lateral_sides=sides)
    st.buoyancy(a, m, ni=ni, nj=nj, nk=nk)
    cpu_psolver3_stage_inplace(a, m, ni=ni, nj=nj, nk=nk, dt=dt, nsound=cfg["nsound"], nrk=nrk, dx=dx, dy=dy, dz=dz,
                               lateral_sides=sides, prof=prof, zt=zt)
    st.getdiv(a, m, a["divergence"], a["theta_div"], a["thterm"], ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    st.getdivx(a, m, a["divx"], ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    dttmp = F(F(dt) / F(4 - nrk))
    st.pack_theta(base["th0"], a["th3d"], a["theta_padded"], ni=ni, nj=nj, nk=nk)
    bc.fill(a["theta_padded"], bc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
    cpu_scalar_advection5_inplace(a["theta_padded"], a, m, a["divx"], a["advection"], _flux_scratch(ni, nj, nk),
                                  ni=ni, nj=nj, nk=nk, nrk=nrk, dx=dx, dy=dy, dz=dz, weps=THETA_WEPS,
                                  lateral_sides=sides)
    a["th3d"][...] = st.compose_theta(a["tha"], a["thten1"], a["theta_div"], a["advection"], dttmp)


def large_step(a, m, base, prof, cfg, *, zt):
    """Advance the dry state ``a`` by one large step in place."""
    ni, nj, nk = cfg["ni"], cfg["nj"], cfg["nk"]
    sides = tuple(cfg["lateral_sides"])
    for name in ("uten1", "vten1", "wten1", "thten1", "ppten"):
        a[name][...] = F(0.0)
    for nrk in (1, 2, 3):
        rk_stage(a, m, base, prof, cfg, nrk, zt=zt)
    st.solve3_pressure(a, sides, ni=ni, nj=nj, nk=nk)
    a["ua"][...] = a["u3d"]
    a["va"][...] = a["v3d"]
    a["wa"][...] = a["w3d"]
    a["tha"][...] = a["th3d"]
