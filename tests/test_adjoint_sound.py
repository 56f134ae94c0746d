"""Adjoint of the acoustic RK stage on the C5 geometry (dry, flat, uniform, periodic).

CPU: the binary64 adjoint (oracle/adjoint_sound.py) against the binary64
restatement of the forward stage by the dot-product identity
<ybar, F(x + d) - F(x)> = <F^T ybar, d>: for fixed coefficients the stage is
linear in its state and forcing, so the identity holds to binary64 rounding
for any d.  Each RK stage (nrk = 1, 2, 3) is checked with random output
adjoints over every word of every output (halos included) and random
perturbations of the stage's inputs.
"""

from __future__ import annotations

import numpy as np
import pytest

from gm1_amd.fields import CaseConfig
from gm1_amd.large_step import host_initial_state
from test_binary64 import oracle_binary64, promote

PERIODIC = (1, 1, 1, 1)


def c5_like(ni=14, nj=12, nk=10, seed=11):
    """C5's geometry and state family at a small size: dx 500 m, dz 250 m, dt 3 s, 6 substeps."""
    return CaseConfig(f"c5-{ni}x{nj}x{nk}", ni, nj, nk, 500.0, 500.0, 250.0, 3.0, 6, PERIODIC, False, "linear15",
                      ((2.0, ni * 250.0, nj * 250.0, 1400.0),), bubble_radius=(ni * 150.0, 1400.0), noise=0.1,
                      seed=seed)


def stage_case(nrk, seed=0):
    """A binary64 state after the stage terms, with nonzero winds, pressure and forcing everywhere."""
    o64 = oracle_binary64()
    cfg = c5_like()
    host = host_initial_state(cfg)
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    a = promote({k: v.copy() for k, v in host["arrays"].items()})
    m = promote(host["metrics"])
    base = promote(host["base"])
    rng = np.random.default_rng(100 + nrk + seed)
    for name, scale in (("u3d", 2.0), ("v3d", 2.0), ("w3d", 0.5), ("ua", 2.0), ("va", 2.0), ("wa", 0.5),
                        ("uten", 1e-3), ("vten", 1e-3), ("wten", 1e-3), ("ppten", 1e-6), ("ppx", 1e-4),
                        ("ppi", 1e-4), ("pp3d", 1e-4)):
        a[name] = a[name] + scale * rng.standard_normal(a[name].shape)
    o64.stage.stage_terms(a, base, a["th3d"], None, ni=ni, nj=nj, nk=nk, sides=PERIODIC)
    prof = o64.sound.terrain_coordinate_profiles(nk, cfg.dz)
    return o64, cfg, a, m, prof


def run_stage(o64, cfg, a, m, prof, nrk):
    b = {k: v.copy() for k, v in a.items()}
    o64.sound.cpu_psolver3_stage_inplace(b, m, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, dt=cfg.dt, nsound=cfg.nsound,
                                         nrk=nrk, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, lateral_sides=PERIODIC,
                                         prof=prof, zt=np.float64(cfg.nk * cfg.dz), radiation=False)
    return b


def inputs_of(nrk):
    if nrk == 1:
        return ("u3d", "v3d", "w3d", "pp3d", "ppx", "wa", "uten", "vten", "wten", "ppten")
    return ("ua", "va", "wa", "ppi", "ppx", "uten", "vten", "wten", "ppten")


@pytest.mark.parametrize("nrk", [1, 2, 3])
def test_stage_adjoint_dot_product_binary64(nrk):
    o64, cfg, a, m, prof = stage_case(nrk)
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    rng = np.random.default_rng(7 + nrk)
    y0 = run_stage(o64, cfg, a, m, prof, nrk)
    d = {n: rng.standard_normal(a[n].shape) * np.abs(a[n]).max() * 1e-3 for n in inputs_of(nrk)}
    a1 = {k: v.copy() for k, v in a.items()}
    for n, dv in d.items():
        a1[n] = a1[n] + dv
    y1 = run_stage(o64, cfg, a1, m, prof, nrk)
    ybar = {n: rng.standard_normal(y0[n].shape) for n in o64.adjoint_sound.STATE}
    lhs = sum(float(np.sum(ybar[n] * (y1[n] - y0[n]))) for n in ybar)
    coef = {n: a[n] for n in ("thv", "ppterm", "piadv")}
    xbar = o64.adjoint_sound.stage_adjoint(ybar, coef, m, ni=ni, nj=nj, nk=nk, dt=cfg.dt, nsound=cfg.nsound,
                                           nrk=nrk, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, prof=prof)
    rhs = sum(float(np.sum(xbar[n] * d[n])) for n in d)
    scale = sum(float(np.sum(np.abs(ybar[n] * (y1[n] - y0[n])))) for n in ybar)
    print(f"nrk={nrk} lhs={lhs:.12e} rhs={rhs:.12e} scale={scale:.3e}")
    assert abs(lhs - rhs) <= 1e-9 * scale, (lhs, rhs, scale)
