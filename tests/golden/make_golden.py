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


GENERATORS = {"advection5": advection5, "momentum5": momentum5, "psolver3": psolver3}


if __name__ == "__main__":
    names = sys.argv[1:] or list(GENERATORS)
    for name in names:
        GENERATORS[name]()
