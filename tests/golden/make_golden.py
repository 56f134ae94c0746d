"""Write golden input/output vectors from the reference's own CPU functions.

Run in the build container only (the reference is not present elsewhere):

    python tests/golden/make_golden.py [name ...]

Each fixture is data: the seeded inputs and the outputs the reference's
``cpu_*_inplace`` transcriptions produce from them, stored as compressed
``.npz`` next to this script.  Nothing of the reference's code is stored.
The reference is imported from ``/root/reference/src`` (read-only).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REFERENCE_SRC = Path("/root/reference/src")
F = np.float32
NG = 3


def _ref():
    if str(REFERENCE_SRC) not in sys.path:
        sys.path.insert(0, str(REFERENCE_SRC))


def _save(name: str, **data) -> None:
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **data)
    print(f"wrote {path.name}: {sum(v.nbytes for v in data.values() if hasattr(v, 'nbytes'))} bytes raw")


def _shapes(ni, nj, nk):
    return {
        "u": (nk + 2, nj + 2 * NG, ni + 1 + 2 * NG),
        "v": (nk + 2, nj + 1 + 2 * NG, ni + 2 * NG),
        "w": (nk + 3, nj + 2 * NG, ni + 2 * NG),
        "s": (nk + 2, nj + 2 * NG, ni + 2 * NG),
        "p": (nk + 2, nj + 2, ni + 2),
    }


def _metric_shapes(ni, nj, nk):
    _ref()
    from cm1core.cuda_terrain_open_boundary import terrain_static_shapes

    return terrain_static_shapes(ni=ni, nj=nj, nk=nk)


def seeded_metrics(rng, ni, nj, nk, *, terrain=True):
    """Smooth positive metrics with mild seeded variation (our own construction)."""
    shapes = _metric_shapes(ni, nj, nk)
    m = {name: np.ones(shape, F) for name, shape in shapes.items()}
    kk = np.arange(nk, dtype=np.float64)
    kf = np.arange(nk + 1, dtype=np.float64)
    rho = 1.15 * np.exp(-kk / 30.0)
    m["rho0_3d"][:] = (rho[:, None, None] * (1.0 + 0.01 * rng.standard_normal((nk, nj, ni)))).astype(F)
    m["rr0"][:] = (F(1.0) / m["rho0_3d"]).astype(F)
    rf = 1.18 * np.exp(-kf / 30.0)
    m["rf0"][:] = (rf[:, None, None] * (1.0 + 0.01 * rng.standard_normal((nk + 1, nj, ni)))).astype(F)
    m["rrf0"][:] = (F(1.0) / m["rf0"]).astype(F)
    c2 = (0.5 + 0.05 * rng.standard_normal((nk + 1, nj, ni))).astype(F)
    m["c2"][:] = c2
    m["c1"][:] = (F(1.0) - c2).astype(F)
    if terrain:
        m["dzdx"][:] = (0.01 * rng.standard_normal((nj, ni))).astype(F)
        m["dzdy"][:] = (0.01 * rng.standard_normal((nj, ni))).astype(F)
        m["gz"][:] = (1.0 + 0.02 * rng.standard_normal((nj, ni))).astype(F)
        m["rgz"][:] = (F(1.0) / m["gz"]).astype(F)
        m["gzu"][:] = (1.0 + 0.02 * rng.standard_normal((nj, ni + 1))).astype(F)
        m["rgzu"][:] = (F(1.0) / m["gzu"]).astype(F)
        m["gzv"][:] = (1.0 + 0.02 * rng.standard_normal((nj + 1, ni))).astype(F)
        m["rgzv"][:] = (F(1.0) / m["gzv"]).astype(F)
        m["gxu"][:] = (0.01 * rng.standard_normal((nk + 1, nj, ni + 1))).astype(F)
        m["gyv"][:] = (0.01 * rng.standard_normal((nk + 1, nj + 1, ni))).astype(F)
        m["gx"][:] = (0.01 * rng.standard_normal((nk + 1, nj, ni))).astype(F)
        m["gy"][:] = (0.01 * rng.standard_normal((nk + 1, nj, ni))).astype(F)
    else:
        for name in ("dzdx", "dzdy", "gxu", "gyv", "gx", "gy"):
            m[name][:] = F(0.0)
    return m


def advection5():
    """cpu_terrain_mass_flux5_inplace + cpu_scalar_advection5_inplace."""
    _ref()
    from cm1core.cuda_advection5 import cpu_scalar_advection5_inplace, cpu_terrain_mass_flux5_inplace

    out = {}
    for case, (ni, nj, nk, seed) in enumerate(((12, 10, 9, 11), (9, 8, 8, 12))):
        rng = np.random.default_rng(seed)
        sh = _shapes(ni, nj, nk)
        m = seeded_metrics(rng, ni, nj, nk)
        a = {
            "u3d": (3.0 * rng.standard_normal(sh["u"])).astype(F),
            "v3d": (3.0 * rng.standard_normal(sh["v"])).astype(F),
            "w3d": (1.0 * rng.standard_normal(sh["w"])).astype(F),
            "rru": np.zeros(sh["u"], F),
            "rrv": np.zeros(sh["v"], F),
            "rrw": np.zeros(sh["w"], F),
        }
        scalar = (300.0 + 2.0 * rng.standard_normal(sh["s"])).astype(F)
        # a sharp front so WENO weights move
        scalar[:, :, : NG + ni // 2] += F(5.0)
        divx = np.zeros((nk, nj, ni), F)
        pre = {f"c{case}_in_{k}": v.copy() for k, v in a.items()}
        cpu_terrain_mass_flux5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, dx=250.0, dy=200.0, dz=150.0)
        out.update(pre)
        out.update({f"c{case}_metric_{k}": v for k, v in m.items() if k in (
            "rho0_3d", "rf0", "rgzu", "rgzv", "dzdx", "dzdy", "gz", "c1", "c2", "rr0")})
        out.update({f"c{case}_out_{k}": a[k].copy() for k in ("rru", "rrv", "rrw")})
        out[f"c{case}_out_divx"] = divx.copy()
        out[f"c{case}_scalar"] = scalar
        for nrk in (1, 3):
            fl = {"x": np.zeros_like(a["rru"]), "y": np.zeros_like(a["rrv"]), "z": np.zeros_like(a["rrw"])}
            tend = np.zeros((nk, nj, ni), F)
            cpu_scalar_advection5_inplace(
                scalar, a, m, divx, tend, fl, ni=ni, nj=nj, nk=nk, nrk=nrk,
                dx=250.0, dy=200.0, dz=150.0, weps=1.0e-17,
            )
            out[f"c{case}_nrk{nrk}_tend"] = tend
            for d in "xyz":
                out[f"c{case}_nrk{nrk}_flux_{d}"] = fl[d]
        out[f"c{case}_dims"] = np.array([ni, nj, nk], np.int64)
    _save("advection5", **out)


def momentum5():
    """cpu_terrain_open_momentum_advection5_inplace for three lateral families."""
    _ref()
    from cm1core.cuda_momentum_advection5 import cpu_terrain_open_momentum_advection5_inplace

    out = {}
    cases = (((10, 9, 8, 21), (2, 2, 2, 2)), ((8, 11, 9, 22), (1, 1, 1, 1)), ((9, 8, 10, 23), (1, 1, 2, 3)))
    for case, ((ni, nj, nk, seed), sides) in enumerate(cases):
        rng = np.random.default_rng(seed)
        sh = _shapes(ni, nj, nk)
        m = seeded_metrics(rng, ni, nj, nk)
        a = {
            "u3d": (3.0 * rng.standard_normal(sh["u"])).astype(F),
            "v3d": (3.0 * rng.standard_normal(sh["v"])).astype(F),
            "w3d": (1.0 * rng.standard_normal(sh["w"])).astype(F),
            "uten": (0.01 * rng.standard_normal(sh["u"])).astype(F),
            "vten": (0.01 * rng.standard_normal(sh["v"])).astype(F),
            "wten": (0.01 * rng.standard_normal(sh["w"])).astype(F),
            "rru": np.zeros(sh["u"], F), "rrv": np.zeros(sh["v"], F), "rrw": np.zeros(sh["w"], F),
            "uh": np.ones(ni + 2 * NG, F), "vh": np.ones(nj + 2 * NG, F),
        }
        divx = np.zeros((nk, nj, ni), F)
        out.update({f"c{case}_in_{k}": v.copy() for k, v in a.items()})
        out.update({f"c{case}_metric_{k}": v for k, v in m.items() if k in (
            "rho0_3d", "rf0", "rr0", "rrf0", "c1", "c2", "gz", "gzu", "gzv", "rgzu", "rgzv", "dzdx", "dzdy")})
        cpu_terrain_open_momentum_advection5_inplace(
            a, m, divx, ni=ni, nj=nj, nk=nk, nrk=2, dx=250.0, dy=200.0, dz=150.0, lateral_sides=sides)
        out.update({f"c{case}_out_{k}": a[k].copy() for k in ("rru", "rrv", "rrw", "uten", "vten", "wten")})
        out[f"c{case}_out_divx"] = divx
        out[f"c{case}_dims"] = np.array([ni, nj, nk], np.int64)
        out[f"c{case}_sides"] = np.array(sides, np.int64)
    _save("momentum5", **out)


def psolver_inputs(rng, ni, nj, nk, *, terrain=True):
    """Seeded, physically scaled acoustic-stage carriers (our own construction)."""
    _ref()
    from cm1core.cuda_psolver3_terrain_open import terrain_open_sound_shapes

    shapes = terrain_open_sound_shapes(ni, nj, nk)
    a = {name: np.zeros(shape, F) for name, shape in shapes.items()}
    p = shapes["ppi"]
    a["thv"][:] = (300.0 + rng.standard_normal(p)).astype(F)
    a["piadv"][:] = (1.0 - 0.01 * np.arange(nk + 2)[:, None, None] + 1e-4 * rng.standard_normal(p)).astype(F)
    a["ppterm"][:] = (0.4 + 0.01 * rng.standard_normal(p)).astype(F)
    a["ppten"][:] = (1e-5 * rng.standard_normal(p)).astype(F)
    for name in ("ppi", "pp3d", "ppx"):
        a[name][:] = (1e-4 * rng.standard_normal(p)).astype(F)
    for name, scale in (("ua", 5.0), ("u3d", 5.0), ("va", 5.0), ("v3d", 5.0), ("wa", 1.0), ("w3d", 1.0),
                        ("uten", 1e-3), ("vten", 1e-3), ("wten", 1e-3)):
        a[name][:] = (scale * rng.standard_normal(a[name].shape)).astype(F)
    for name in ("radbcw", "radbce", "radbcs", "radbcn"):
        a[name][:] = (-5.0 * np.abs(rng.standard_normal(a[name].shape))).astype(F)
    for name in ("uh", "vh", "ruf", "rvf"):
        a[name][:] = F(1.0)
    a["scalar"][:] = (1.0 + rng.standard_normal(shapes["scalar"])).astype(F)
    m = seeded_metrics(rng, ni, nj, nk, terrain=terrain)
    return a, m


def psolver3():
    """cpu_psolver3_terrain_open_stage_inplace, three stages (all-open family)."""
    _ref()
    from cm1core.cuda_psolver3_terrain_open import cpu_psolver3_terrain_open_stage_inplace

    out = {}
    cases = ((9, 8, 8, 31, 6, 2.0, 250.0, 200.0, 150.0, True), (10, 9, 9, 32, 8, 1.5, 500.0, 500.0, 250.0, False))
    for case, (ni, nj, nk, seed, ns, dt, dx, dy, dz, terrain) in enumerate(cases):
        rng = np.random.default_rng(seed)
        a, m = psolver_inputs(rng, ni, nj, nk, terrain=terrain)
        out.update({f"c{case}_in_{k}": v.copy() for k, v in a.items()})
        out.update({f"c{case}_metric_{k}": v for k, v in m.items()})
        for nrk in (1, 2, 3):
            cpu_psolver3_terrain_open_stage_inplace(a, m, ni=ni, nj=nj, nk=nk, dt=dt, nsound=ns, nrk=nrk,
                                                   dx=dx, dy=dy, dz=dz)
            for k in ("u3d", "v3d", "w3d", "pp3d", "ppd", "ppx", "rru", "rrv", "rrw",
                      "radbcw", "radbce", "radbcs", "radbcn", "scalar", "qk", "pk1", "pk2", "mm"):
                out[f"c{case}_nrk{nrk}_{k}"] = a[k].copy()
        out[f"c{case}_params"] = np.array([ni, nj, nk, ns, dt, dx, dy, dz], np.float64)
    _save("psolver3", **out)


def stage():
    """eqtset=2 cell terms, theta packing, A-anchored compose, held seed and bc2d."""
    _ref()
    from cm1core.cuda_moist_stage_bridge import cpu_eqtset2_cell_f32
    from cm1core.cuda_momentum_rk import cpu_bc2d_open_w_surface_inplace, cpu_seed_momentum_rk_tendencies_inplace
    from cm1core.cuda_scalar_moist_rk import cpu_compose_scalar_moist_rk_stage_inplace, cpu_theta_transport_scalar_inplace

    rng = np.random.default_rng(41)
    out = {}
    # eqtset=2 terms on single cells: profiles, perturbations and six species
    # (some negative, some exactly zero) drawn per cell
    n = 240
    cells = {
        "th0": rng.uniform(295.0, 350.0, n), "qv0": rng.uniform(0.0, 0.015, n), "pi0": rng.uniform(0.6, 1.0, n),
        "qc0": np.where(rng.random(n) < 0.3, rng.uniform(0.0, 1e-3, n), 0.0),
        "qi0": np.where(rng.random(n) < 0.3, rng.uniform(0.0, 1e-4, n), 0.0),
        "th3d": rng.normal(0.0, 2.0, n), "pp3d": rng.normal(0.0, 1e-3, n),
    }
    for q in ("qv", "qc", "qr", "qi", "qs", "qg"):
        v = rng.uniform(-2e-4, 3e-3 if q in ("qv", "qc", "qr") else 5e-4, n)
        v[rng.random(n) < 0.25] = 0.0
        cells[q] = v
    cells = {k: v.astype(F) for k, v in cells.items()}
    res = {k: np.zeros(n, F) for k in ("buoyancy", "theta_rho", "theta_term", "pressure_term")}
    for c in range(n):
        r = cpu_eqtset2_cell_f32(**{k: float(v[c]) for k, v in cells.items()})
        for k in res:
            res[k][c] = r[k]
    out.update({f"eq_in_{k}": v for k, v in cells.items()})
    out.update({f"eq_out_{k}": v for k, v in res.items()})
    # theta packing
    ni, nj, nk = 11, 9, 8
    th0 = rng.uniform(299.0, 345.0, nk).astype(F)
    th3d = rng.normal(0.0, 2.0, (nk, nj, ni)).astype(F)
    padded = np.zeros((nk + 2, nj + 2 * NG, ni + 2 * NG), F)
    cpu_theta_transport_scalar_inplace(th0, th3d, padded, ni=ni, nj=nj, nk=nk)
    out.update({"pack_th0": th0, "pack_th3d": th3d, "pack_padded": padded})
    # A-anchored compose, six species, three stages at dt 1.5 s
    shape = (nk, nj, ni)
    comp = {"theta_a": rng.normal(0.0, 2.0, shape), "theta_held": rng.normal(0.0, 1e-3, shape),
            "theta_divergence_term": rng.normal(0.0, 1e-3, shape), "theta_advection": rng.normal(0.0, 1e-2, shape),
            "moisture_a": np.abs(rng.normal(0.0, 1e-3, (6,) + shape)),
            "moisture_held": rng.normal(0.0, 1e-7, (6,) + shape),
            "moisture_advection": rng.normal(0.0, 1e-6, (6,) + shape)}
    comp = {k: v.astype(F) for k, v in comp.items()}
    comp["theta_a"].flat[::13] = F(1e-31)
    comp["theta_held"].flat[::13] = F(0.0)
    comp["theta_divergence_term"].flat[::13] = F(0.0)
    comp["theta_advection"].flat[::13] = F(0.0)
    comp["moisture_a"].flat[::11] = F(0.0)
    comp["moisture_held"].flat[::11] = F(0.0)
    comp["moisture_advection"].flat[::11] = F(0.0)
    out.update({f"comp_in_{k}": v for k, v in comp.items()})
    for nrk in (1, 2, 3):
        th = np.zeros(shape, F)
        q = np.zeros((6,) + shape, F)
        cpu_compose_scalar_moist_rk_stage_inplace(
            comp["theta_a"], comp["moisture_a"], comp["theta_held"], comp["moisture_held"],
            comp["theta_divergence_term"], comp["theta_advection"], comp["moisture_advection"], th, q,
            nrk=nrk, dt=1.5)
        out[f"comp_nrk{nrk}_theta"] = th
        out[f"comp_nrk{nrk}_moisture"] = q
    # held seed and bc2d on the padded momentum family
    sh = _shapes(ni, nj, nk)
    held = {n: rng.normal(0.0, 1.0, sh[f]).astype(F) for n, f in (("uten1", "u"), ("vten1", "v"), ("wten1", "w"))}
    tend = {n: rng.normal(0.0, 1.0, sh[f]).astype(F) for n, f in (("uten", "u"), ("vten", "v"), ("wten", "w"))}
    out.update({f"seed_in_{k}": v.copy() for k, v in {**held, **tend}.items()})
    cpu_seed_momentum_rk_tendencies_inplace(held, tend, ni=ni, nj=nj, nk=nk)
    out.update({f"seed_out_{k}": v for k, v in tend.items()})
    w = rng.normal(0.0, 1.0, sh["w"]).astype(F)
    out["bc2d_in_w"] = w.copy()
    cpu_bc2d_open_w_surface_inplace(w, ni=ni, nj=nj, nk=nk)
    out["bc2d_out_w"] = w
    out["dims"] = np.array([ni, nj, nk], np.int64)
    _save("stage", **out)


def pdef():
    """cpu_pdef_direction_inplace for x, y and z on seeded rows with activated negatives."""
    _ref()
    from cm1core.cuda_scalar_moist_rk import cpu_pdef_direction_inplace

    rng = np.random.default_rng(53)
    ni, nj, nk = 12, 10, 9
    sh = _shapes(ni, nj, nk)
    out = {"dims": np.array([ni, nj, nk], np.int64)}
    qa = np.abs(rng.normal(0.0, 1e-3, (nk, nj, ni))).astype(F)
    qa[rng.random(qa.shape) < 0.2] = F(0.0)
    s = np.abs(rng.normal(0.0, 1e-3, sh["s"])).astype(F)
    rho0 = (1.1 + 0.05 * rng.standard_normal((nk, nj, ni))).astype(F)
    gz = (1.0 + 0.02 * rng.standard_normal((nj, ni))).astype(F)
    rgz = (F(1.0) / gz).astype(F)
    out.update({"qa": qa, "s": s, "rho0": rho0, "gz": gz, "rgz": rgz})
    for axis, fshape, spacing in (("x", sh["u"], 250.0), ("y", sh["v"], 300.0), ("z", sh["w"], 200.0)):
        flux = (3.0 * rng.standard_normal(fshape)).astype(F)
        # directional tendencies large enough that about a fifth of the candidates go negative
        directional = (rng.normal(0.0, 1.0, (nk, nj, ni)) * 4e-4).astype(F)
        candidate = np.zeros((nk, nj, ni), F)
        out[f"{axis}_flux"] = flux
        out[f"{axis}_in_directional"] = directional.copy()
        cpu_pdef_direction_inplace(qa, s, rho0, gz, rgz, flux, directional, candidate, axis=axis, dt=1.5,
                                   spacing=spacing)
        out[f"{axis}_out_directional"] = directional
        out[f"{axis}_out_candidate"] = candidate
        out[f"{axis}_spacing"] = np.array(spacing)
    _save("pdef", **out)


def kessler_state(rng, ni, nj, nk, *, dz):
    """A seeded moist column field: some cells supersaturated, cloudy, raining or slightly negative."""
    z = (np.arange(nk) + 0.5) * dz
    th0 = (300.0 + 0.004 * z)[:, None, None] * np.ones((nk, nj, ni))
    pi0 = (1.0 - 9.81 / 1005.7 * z / 300.0)[:, None, None] * np.ones((nk, nj, ni))
    a = {"th0": th0.astype(F), "pi0": pi0.astype(F)}
    a["th3d"] = rng.normal(0.0, 1.0, (nk, nj, ni)).astype(F)
    a["pp3d"] = rng.normal(0.0, 1e-4, (nk, nj, ni)).astype(F)
    pi = (a["pi0"] + a["pp3d"]).astype(np.float64)
    prs = 1.0e5 * pi ** (1005.7 / 287.04)
    t = (a["th0"] + a["th3d"]).astype(np.float64) * pi
    es = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))
    qvs = 0.622 * es / (prs - es)
    a["qv"] = (qvs * rng.uniform(0.7, 1.06, t.shape)).astype(F)
    qc = np.where(rng.random(t.shape) < 0.5, rng.uniform(0.0, 2e-3, t.shape), 0.0)
    qc[rng.random(t.shape) < 0.05] = -1e-6
    qc[rng.random(t.shape) < 0.05] = 5e-15
    qr = np.where(rng.random(t.shape) < 0.5, rng.uniform(0.0, 3e-3, t.shape), 0.0)
    qr[rng.random(t.shape) < 0.05] = -2e-6
    qr[rng.random(t.shape) < 0.05] = 5e-9
    a["qc"], a["qr"] = qc.astype(F), qr.astype(F)
    a["prs"] = prs.astype(F)
    a["rho"] = (prs / (287.04 * (1.0 + a["qv"] / 0.622) * t)).astype(F)
    a["rr"] = a["rho"].copy()
    a["ruh"] = np.ones(ni, F)
    a["rvh"] = np.ones(nj, F)
    a["rmh"] = np.ones((nk, nj, ni), F)
    a["mh"] = np.ones((nk, nj, ni), F)
    a["fall_velocity"] = np.zeros((nk, nj, ni), F)
    a["rain"] = np.zeros((1, nj, ni), F)
    a["prate"] = np.zeros((nj, ni), F)
    a["asq"] = np.zeros(3, np.float64)
    a["qbudget"] = np.zeros(10, np.float64)
    return a


def kessler():
    """cpu_gm1_kessler_driver_inplace (eqtset 1 and 2, one and several fall substeps),
    cpu_gm1_kessler_large_step_inplace (eqtset 2 with held rates) and cpu_apmasscon_inplace."""
    _ref()
    from cm1core.cuda_apmasscon import cpu_apmasscon_inplace, cpu_initial_dry_mass
    from cm1core.cuda_gm1_kessler_driver import Gm1KesslerDriverSelection, cpu_gm1_kessler_driver_inplace
    from cm1core.cuda_gm1_kessler_large_step import cpu_gm1_kessler_large_step_inplace

    out = {}
    ni, nj, nk = 7, 5, 12
    out["dims"] = np.array([ni, nj, nk], np.int64)
    cases = ((0, 1.5, 250.0, 2), (1, 20.0, 100.0, 2), (2, 1.5, 250.0, 1))
    for case, dt, dz, eqt in cases:
        rng = np.random.default_rng(70 + case)
        a = kessler_state(rng, ni, nj, nk, dz=dz)
        out.update({f"d{case}_in_{k}": v.copy() for k, v in a.items()})
        res = cpu_gm1_kessler_driver_inplace(a, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=dz, dt=dt,
                                             selection=Gm1KesslerDriverSelection(equation_set=eqt))
        out.update({f"d{case}_out_{k}": v for k, v in a.items()})
        out[f"d{case}_iterations"] = res["satadj"]["iterations"]
        out[f"d{case}_nfall"] = np.array(res["max_fall_substeps"])
        out[f"d{case}_params"] = np.array([dt, dz, eqt], np.float64)
    # the large-step transition with held rates from a previous step
    rng = np.random.default_rng(80)
    a = kessler_state(rng, ni, nj, nk, dz=250.0)
    held = {n: rng.normal(0.0, s, (nk, nj, ni)).astype(F)
            for n, s in (("qpten", 1e-6), ("qtten", 1e-3), ("qvten", 1e-7), ("qcten", 1e-7))}
    out.update({f"t_in_{k}": v.copy() for k, v in a.items()})
    out.update({f"t_in_{k}": v.copy() for k, v in held.items()})
    cpu_gm1_kessler_large_step_inplace(a, held, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0, dt=1.5,
                                       selection=Gm1KesslerDriverSelection(equation_set=2))
    out.update({f"t_out_{k}": v for k, v in a.items()})
    out.update({f"t_out_{k}": v for k, v in held.items()})
    # dry-mass ledger and correction
    rng = np.random.default_rng(81)
    a = kessler_state(rng, ni, nj, nk, dz=250.0)
    m = {k: a[k] for k in ("rho", "ruh", "rvh", "rmh", "pi0", "th0", "th3d", "qv", "pp3d", "prs")}
    mass1 = cpu_initial_dry_mass(m, ni=ni, nj=nj, nk=nk) * (1.0 + 1e-5)
    out.update({f"m_in_{k}": v.copy() for k, v in m.items()})
    r = cpu_apmasscon_inplace(m, ni=ni, nj=nj, nk=nk, mass1=mass1)
    out.update({f"m_out_{k}": v for k, v in m.items()})
    out["m_mass1"] = np.array(mass1)
    out["m_result"] = np.array([r.mass2_before_f64, r.pressure_sum_before_f64, r.correction_f32, r.mass_after_f64])
    _save("kessler", **out)


def kessler_efall():
    """cpu_gm1_kessler_efall_inplace (the isolated getcvm/getefall/update, eqtset 1 and 2),
    cpu_gm1_kessler_driver_inplace with efall=1 (eqtset 1 and 2, one and several fall substeps) and
    cpu_gm1_kessler_large_step_inplace with efall=1 (eqtset 2 with held rates)."""
    _ref()
    from cm1core.cuda_gm1_kessler_driver import Gm1KesslerDriverSelection, cpu_gm1_kessler_driver_inplace
    from cm1core.cuda_gm1_kessler_efall import Gm1KesslerEfallSelection, cpu_gm1_kessler_efall_inplace
    from cm1core.cuda_gm1_kessler_large_step import cpu_gm1_kessler_large_step_inplace

    out = {}
    ni, nj, nk = 7, 5, 12
    out["dims"] = np.array([ni, nj, nk], np.int64)

    def with_mf(rng, a):
        a["mf"] = (1.0 + 0.1 * rng.uniform(-1.0, 1.0, (nk + 1, nj, ni))).astype(F)
        return a

    # the isolated operator on a post-geterain state (fall speed formed, temperature held)
    for case, (dt, dz, eqt) in enumerate(((1.5, 250.0, 2), (6.0, 100.0, 1))):
        rng = np.random.default_rng(90 + case)
        k = with_mf(rng, kessler_state(rng, ni, nj, nk, dz=dz))
        rho, qr = k["rho"].astype(np.float64), k["qr"].astype(np.float64)
        vt = np.where(qr > 1e-12, 14.34 * np.abs(rho * qr) ** 0.1346 * np.sqrt(1.15 / rho), 0.0)
        e = {n: k[n] for n in ("mf", "pi0", "th0", "rho", "qv", "qc", "qr", "pp3d", "th3d", "prs")}
        e["fall_velocity"] = vt.astype(F)
        e["temperature"] = ((k["th0"] + k["th3d"]) * (k["pi0"] + k["pp3d"])).astype(F)
        out.update({f"e{case}_in_{n}": v.copy() for n, v in e.items()})
        res = cpu_gm1_kessler_efall_inplace(e, ni=ni, nj=nj, nk=nk, dz=dz, dt=dt,
                                            selection=Gm1KesslerEfallSelection(equation_set=eqt))
        out.update({f"e{case}_out_{n}": v for n, v in e.items()})
        out[f"e{case}_cvm"] = res["mixture_cv"]
        out[f"e{case}_tendency"] = res["energy_tendency"]
        out[f"e{case}_params"] = np.array([dt, dz, eqt], np.float64)
    # the whole driver with efall=1
    cases = ((0, 1.5, 250.0, 2), (1, 20.0, 100.0, 2), (2, 6.0, 250.0, 1))
    for case, dt, dz, eqt in cases:
        rng = np.random.default_rng(95 + case)
        a = with_mf(rng, kessler_state(rng, ni, nj, nk, dz=dz))
        out.update({f"d{case}_in_{k}": v.copy() for k, v in a.items()})
        res = cpu_gm1_kessler_driver_inplace(a, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=dz, dt=dt,
                                             selection=Gm1KesslerDriverSelection(equation_set=eqt, efall=1))
        out.update({f"d{case}_out_{k}": v for k, v in a.items()})
        out[f"d{case}_iterations"] = res["satadj"]["iterations"]
        out[f"d{case}_nfall"] = np.array(res["max_fall_substeps"])
        out[f"d{case}_temperature"] = res["temperature_held"]
        out[f"d{case}_params"] = np.array([dt, dz, eqt], np.float64)
    # the large-step transition with efall=1 and held rates from a previous step
    rng = np.random.default_rng(99)
    a = with_mf(rng, kessler_state(rng, ni, nj, nk, dz=250.0))
    held = {n: rng.normal(0.0, s, (nk, nj, ni)).astype(F)
            for n, s in (("qpten", 1e-6), ("qtten", 1e-3), ("qvten", 1e-7), ("qcten", 1e-7))}
    out.update({f"t_in_{k}": v.copy() for k, v in a.items()})
    out.update({f"t_in_{k}": v.copy() for k, v in held.items()})
    cpu_gm1_kessler_large_step_inplace(a, held, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0, dt=1.5,
                                       selection=Gm1KesslerDriverSelection(equation_set=2, efall=1))
    out.update({f"t_out_{k}": v for k, v in a.items()})
    out.update({f"t_out_{k}": v for k, v in held.items()})
    _save("kessler_efall", **out)


def periodic_inputs(rng, ni, nj, nk, *, smooth):
    """Inputs of the periodic acoustic stage (our own construction), halos periodic."""
    _ref()
    from cm1core.cuda_sound3_periodic import (array_shapes, default_periodic_face_weights,
                                              fill_periodic_scalar_inplace, fill_periodic_uv_inplace)

    shapes = array_shapes(ni, nj, nk)
    a = {name: np.zeros(shape, F) for name, shape in shapes.items()}
    a["dtu0"], a["dtv0"] = default_periodic_face_weights(ni, nj)
    for name in ("uf", "vf", "uh", "vh"):
        a[name].fill(F(1.0))
    z = (np.arange(nk + 2, dtype=np.float64) / (nk + 1))[:, None, None]
    xs = 2.0 * np.pi * np.arange(ni) / ni
    ys = 2.0 * np.pi * np.arange(nj) / nj
    h = (0.7 * np.sin(xs)[None, :] + 0.4 * np.cos(ys)[:, None])[None] if smooth else rng.uniform(-1, 1, (1, nj, ni))
    inner = (slice(None), slice(1, nj + 1), slice(1, ni + 1))

    def put(name, value):
        a[name][inner] = np.broadcast_to(value, (nk + 2, nj, ni)).astype(F)

    put("mh", 1.0 if smooth else 1.0 + 0.25 * np.sin(3.0 * z) + 0.02 * h)
    put("mf", 1.0 if smooth else 1.0 + 0.2 * np.cos(2.0 * z) - 0.02 * h)
    put("rr0", 1.0 + 0.01 * z)
    put("rf0", 0.99 + 0.012 * z)
    put("c1", 0.45 if smooth else 0.45 + 0.05 * z)
    put("c2", 0.55 if smooth else 0.55 - 0.05 * z)
    put("piadv", 1.0 + 0.0003 * z + 0.00001 * h)
    put("ppterm", 0.014 + 0.0002 * z)
    put("thv", 299.0 + 1.8 * z + 0.03 * h)
    put("ppten", 2.0e-6 * h + 1.0e-6 * z)
    pressure = 2.0e-4 * h + 4.0e-5 * z
    put("ppi", pressure)
    put("ppx", 0.2 * pressure + 1.0e-6 * z)
    for name in ("mh", "rr0", "rf0", "c1", "c2", "mf", "piadv", "ppterm", "thv", "ppten", "ppi", "ppx"):
        fill_periodic_scalar_inplace(a[name], ni=ni, nj=nj, nk=nk)
    zk = (np.arange(1, nk + 1) / nk)[:, None, None]
    hw = h if smooth else rng.uniform(-1, 1, (1, nj, ni))
    a["wa"][1 : nk + 1] = (0.003 * hw * np.sin(np.pi * zk)).astype(F)
    a["wten"][1 : nk + 1] = np.broadcast_to(2.0e-5 * hw, (nk, nj, ni)).astype(F)
    xf = 2.0 * np.pi * np.arange(ni + 1) / ni
    yc = 2.0 * np.pi * np.arange(nj) / nj
    a["ua"][1 : nk + 1, 1 : nj + 1, :] = np.broadcast_to(4.0 + 0.15 * np.sin(xf)[None, :] * np.cos(yc)[:, None],
                                                         (nk, nj, ni + 1)).astype(F)
    a["uten"][1 : nk + 1, 1 : nj + 1, :] = np.broadcast_to(1.0e-4 * np.cos(xf)[None, :] * np.cos(yc)[:, None],
                                                           (nk, nj, ni + 1)).astype(F)
    yf = 2.0 * np.pi * np.arange(nj + 1) / nj
    xc = 2.0 * np.pi * np.arange(ni) / ni
    a["va"][1 : nk + 1, :, 1 : ni + 1] = np.broadcast_to(-1.5 + 0.12 * np.cos(xc)[None, :] * np.sin(yf)[:, None],
                                                         (nk, nj + 1, ni)).astype(F)
    a["vten"][1 : nk + 1, :, 1 : ni + 1] = np.broadcast_to(-8.0e-5 * np.sin(xc)[None, :] * np.cos(yf)[:, None],
                                                           (nk, nj + 1, ni)).astype(F)
    if not smooth:
        a["uten"][1 : nk + 1, 1 : nj + 1, :] += rng.normal(0, 2e-5, (nk, nj, ni + 1)).astype(F)
        a["uten"][1 : nk + 1, 1 : nj + 1, ni] = a["uten"][1 : nk + 1, 1 : nj + 1, 0]
        a["vten"][1 : nk + 1, :, 1 : ni + 1] += rng.normal(0, 2e-5, (nk, nj + 1, ni)).astype(F)
        a["vten"][1 : nk + 1, nj, 1 : ni + 1] = a["vten"][1 : nk + 1, 0, 1 : ni + 1]
    fill_periodic_uv_inplace(a["ua"], a["va"], ni=ni, nj=nj, nk=nk)
    for dst, src in (("u3d", "ua"), ("v3d", "va"), ("w3d", "wa"), ("pp3d", "ppi"), ("ppd", "ppx")):
        np.copyto(a[dst], a[src])
    return a


def sound3_periodic():
    _ref()
    from cm1core.cuda_sound3_periodic import (COMPARE_OUTPUT_NAMES, cpu_sound3_periodic_three_stage_inplace,
                                              validate_periodic_contract)

    out = {}
    cases = (("a", 5, 4, 7, True, dict(dt=2.0, nsound=6, rdx=1.0 / 1000.0, rdy=1.0 / 1000.0, rdz=1.0 / 500.0)),
             ("b", 9, 7, 10, False, dict(dt=3.0, nsound=6, rdx=1.0 / 500.0, rdy=1.0 / 700.0, rdz=1.0 / 250.0)),
             ("c", 6, 5, 8, False, dict(dt=1.5, nsound=8, rdx=1.0 / 250.0, rdy=1.0 / 250.0, rdz=1.0 / 250.0,
                                         alph=0.55, kdiv=0.05)))
    for tag, ni, nj, nk, smooth, params in cases:
        rng = np.random.default_rng(300 + ord(tag))
        a = periodic_inputs(rng, ni, nj, nk, smooth=smooth)
        validate_periodic_contract(a, ni=ni, nj=nj, nk=nk)
        out.update({f"{tag}_in_{k}": v.copy() for k, v in a.items()})
        result = cpu_sound3_periodic_three_stage_inplace(a, ni=ni, nj=nj, nk=nk, **params)
        out.update({f"{tag}_out_{k}": a[k] for k in COMPARE_OUTPUT_NAMES})
        out[f"{tag}_dims"] = np.array([ni, nj, nk])
        out[f"{tag}_params"] = np.array([params["dt"], params["nsound"], params["rdx"], params["rdy"], params["rdz"],
                                         params.get("alph", 0.60), params.get("kdiv", 0.10)])
        out[f"{tag}_cadence"] = np.array([[s.nrk, s.nloop, s.ppd_boundary_fills, s.stage_pressure_boundary_fills]
                                          for s in result.cadence])
    _save("sound3_periodic", **out)


GENERATORS = {"advection5": advection5, "momentum5": momentum5, "psolver3": psolver3, "stage": stage, "pdef": pdef,
              "kessler": kessler, "kessler_efall": kessler_efall, "sound3_periodic": sound3_periodic}


if __name__ == "__main__":
    names = sys.argv[1:] or list(GENERATORS)
    for name in names:
        GENERATORS[name]()
