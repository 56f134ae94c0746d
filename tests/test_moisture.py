"""Moisture transport: the final-stage positive-definite chain and the moist RK stage.

CPU tests pin ``oracle/pdef.py`` to the reference's ``cpu_pdef_direction_inplace``
(cuda_scalar_moist_rk.py:352-520) on its exact interior and to closed forms;
GPU tests compare the gfx950 kernels with the oracle bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, NG, assert_bits_equal, golden, seeded_metrics, shapes, to_dev, to_host
from oracle import pdef as op


def _metrics_for_repair(g, ni, nj, nk):
    return {"rho0_3d": g["rho0"], "gz": g["gz"], "rgz": g["rgz"]}


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_repair_matches_reference_interior(axis):
    g = golden("pdef")
    ni, nj, nk = (int(v) for v in g["dims"])
    adv, cand = op.repair_direction(
        g["qa"], g["s"], g["rho0"], g["gz"], g["rgz"], g[f"{axis}_flux"], g[f"{axis}_in_directional"].copy(),
        axis=axis, ni=ni, nj=nj, nk=nk, dt=1.5, spacing=float(g[f"{axis}_spacing"]), sides=(1, 1, 1, 1))
    assert_bits_equal(cand, g[f"{axis}_out_candidate"], f"{axis} candidate")
    want = g[f"{axis}_out_directional"]
    if axis == "x":
        sl = (slice(None), slice(None), slice(2, ni - 2))
    elif axis == "y":
        sl = (slice(None), slice(2, nj - 2), slice(None))
    else:
        sl = (slice(None), slice(None), slice(None))
    assert_bits_equal(adv[sl], want[sl], f"{axis} corrected tendency (exact interior)")
    # the repair was exercised
    assert (adv[sl] != g[f"{axis}_in_directional"][sl]).mean() > 0.2


def test_repair_conserves_a_periodic_row_and_leaves_positive_rows_alone():
    rng = np.random.default_rng(3)
    n = 16
    vals = rng.normal(1e-3, 1e-3, (4, n)).astype(F)
    vals[0] = np.abs(vals[0])  # a row with no negatives is untouched
    mass, flag = op._row_mass(vals, op._family_index(n, 1, 1))
    assert not flag[0].any()
    repaired = np.where(flag, vals + mass, vals).astype(np.float64)
    np.testing.assert_allclose(repaired.sum(axis=1), vals.astype(np.float64).sum(axis=1), rtol=0, atol=1e-9)
    # every negative cell that had a positive donor is raised
    has_donor = (np.maximum(np.roll(vals, 1, 1), 0) + np.maximum(np.roll(vals, -1, 1), 0)) > 1e-30
    assert (repaired[(vals < 0) & has_donor] > vals[(vals < 0) & has_donor]).all()


def test_family_index_maps_halos_like_the_fills():
    n = 6
    np.testing.assert_array_equal(op._family_index(n, 1, 1), [4, 5, 0, 1, 2, 3, 4, 5, 0, 1])
    np.testing.assert_array_equal(op._family_index(n, 2, 2), [0, 0, 0, 1, 2, 3, 4, 5, 5, 5])
    np.testing.assert_array_equal(op._family_index(n, 3, 4), [1, 0, 0, 1, 2, 3, 4, 5, 5, 4])


# ---------------------------------------------------------------- GPU parity
def _moist_inputs(rng, ni, nj, nk):
    from oracle import boundary as obc
    from oracle.advection5 import cpu_terrain_mass_flux5_inplace

    sh = shapes(ni, nj, nk)
    m = seeded_metrics(rng, ni, nj, nk)
    a = {"u3d": (4 * rng.standard_normal(sh["u"])).astype(F), "v3d": (4 * rng.standard_normal(sh["v"])).astype(F),
         "w3d": (2 * rng.standard_normal(sh["w"])).astype(F), "rru": np.zeros(sh["u"], F),
         "rrv": np.zeros(sh["v"], F), "rrw": np.zeros(sh["w"], F)}
    divx = np.zeros((nk, nj, ni), F)
    cpu_terrain_mass_flux5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, dx=250.0, dy=300.0, dz=200.0)
    return a, m, divx


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1), (3, 4, 2, 1)])
@pytest.mark.parametrize("species", [0, 1])
def test_gpu_pdef_moisture_stage_bit_exact(device, sides, species):
    import torch

    from gm1_amd.scalar_moist_rk import launch_pdef_moisture
    from oracle import boundary as obc
    from oracle import stage as ost

    if sides == (3, 4, 2, 1):
        sides = (3, 4, 2, 2)
    ni, nj, nk = 37, 21, 12
    rng = np.random.default_rng(100 + species)
    a, m, divx = _moist_inputs(rng, ni, nj, nk)
    sh = shapes(ni, nj, nk)
    q = np.abs(rng.normal(0.0, 1e-3, (nk, nj, ni))).astype(F)
    q[rng.random(q.shape) < 0.3] = F(0.0)
    q[rng.random(q.shape) < 0.05] = F(1e-25)  # under the WENO gate
    qa = np.abs(q + rng.normal(0.0, 2e-4, q.shape)).astype(F)
    qa[rng.random(q.shape) < 0.3] = F(0.0)
    held = rng.normal(0.0, 1e-7, q.shape).astype(F)
    s = np.zeros(sh["s"], F)
    s[1 : nk + 1, NG : NG + nj, NG : NG + ni] = q
    obc.fill(s, obc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
    weps = float(F(F(1e-2) * F(1e-18)))
    dt = 1.5
    adv = op.pdef_advection(qa, s, a, m, divx, ni=ni, nj=nj, nk=nk, dt=F(dt), dx=250.0, dy=300.0, dz=200.0,
                            weps=weps, pdefweno=species != 0, sides=sides)
    want = ost.compose_moisture(qa, held, adv, F(dt))
    d = lambda x: to_dev(x, device)  # noqa: E731
    out = torch.zeros((nk, nj, ni), dtype=torch.float32, device=device)
    scratch = {n: torch.zeros((nk, nj, ni), dtype=torch.float32, device=device) for n in ("advx", "advy", "advz")}
    launch_pdef_moisture(d(s), {k: d(a[k]) for k in ("rru", "rrv", "rrw")}, d(m), d(divx), d(qa), d(held), out,
                         scratch, species=species, ni=ni, nj=nj, nk=nk, dt=dt, dx=250.0, dy=300.0, dz=200.0,
                         lateral_sides=sides)
    torch.cuda.synchronize()
    assert_bits_equal(to_host(out), want, f"pdef moisture species={species} sides={sides}")


@pytest.mark.gpu
@pytest.mark.parametrize("nrk", [1, 3])
def test_gpu_fused_scalar_compose_bit_exact(device, nrk):
    import torch

    from gm1_amd.scalar_moist_rk import launch_scalar_rk_compose
    from oracle import boundary as obc
    from oracle import stage as ost
    from oracle.advection5 import cpu_scalar_advection5_inplace

    ni, nj, nk, sides = 29, 18, 11, (2, 2, 1, 1)
    rng = np.random.default_rng(7 + nrk)
    a, m, divx = _moist_inputs(rng, ni, nj, nk)
    sh = shapes(ni, nj, nk)
    s = np.zeros(sh["s"], F)
    s[1 : nk + 1, NG : NG + nj, NG : NG + ni] = rng.normal(0.0, 2.0, (nk, nj, ni)).astype(F)
    obc.fill(s, obc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
    tha = rng.normal(0.0, 2.0, (nk, nj, ni)).astype(F)
    held = rng.normal(0.0, 1e-3, (nk, nj, ni)).astype(F)
    divterm = rng.normal(0.0, 1e-3, (nk, nj, ni)).astype(F)
    adv = np.zeros((nk, nj, ni), F)
    fl = {"x": np.zeros(sh["u"], F), "y": np.zeros(sh["v"], F), "z": np.zeros(sh["w"], F)}
    weps = float(F(F(10.0) * F(1e-18)))
    cpu_scalar_advection5_inplace(s, a, m, divx, adv, fl, ni=ni, nj=nj, nk=nk, nrk=nrk, dx=250.0, dy=300.0,
                                  dz=200.0, weps=weps, lateral_sides=sides)
    dttmp = F(F(1.5) / F(4 - nrk))
    want = ost.compose_theta(tha, held, divterm, adv, dttmp)
    d = lambda x: to_dev(x, device)  # noqa: E731
    out = torch.zeros((nk, nj, ni), dtype=torch.float32, device=device)
    launch_scalar_rk_compose(d(s), {k: d(a[k]) for k in ("rru", "rrv", "rrw")}, d(m), d(divx), d(tha), d(held),
                             d(divterm), out, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=1.5, dx=250.0, dy=300.0, dz=200.0,
                             weps=weps, lateral_sides=sides)
    torch.cuda.synchronize()
    assert_bits_equal(to_host(out), want, f"fused compose nrk={nrk}")


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(3, 4, 2, 2), (1, 1, 1, 1), (2, 2, 4, 3)])
@pytest.mark.parametrize("ns", [1, 2, 3, 4])
def test_gpu_batched_compose_equals_single_passes(device, sides, ns):
    """ns scalars composed in one pass (theta' with its pack and divergence term first, then species) write
    the words of their own single-scalar passes, at stages 1 and 2."""
    import torch

    from gm1_amd.scalar_moist_rk import launch_scalar_stage_from_carrier, launch_scalar_stages_from_carriers

    ni, nj, nk = 70, 21, 12
    rng = np.random.default_rng(500 + ns + sum(sides))
    a, m, divx = _moist_inputs(rng, ni, nj, nk)
    cell = (nk, nj, ni)
    d = lambda x: to_dev(x, device)  # noqa: E731
    mass = {k: d(a[k]) for k in ("rru", "rrv", "rrw")}
    grid = dict(ni=ni, nj=nj, nk=nk, dt=1.5, dx=250.0, dy=300.0, rdz=float(F(1.0) / F(200.0)), lateral_sides=sides)
    th0 = d((F(300.0) + np.linspace(0.0, 40.0, nk)).astype(F))
    rows = []
    for q in range(ns):
        theta = q == 0
        x = rng.normal(0.0, 2.0, cell) if theta else np.abs(rng.normal(0.0, 1e-3, cell))
        rows.append((d(x.astype(F)), th0 if theta else None, d(rng.normal(0.0, 1.0, cell).astype(F)),
                     d(rng.normal(0.0, 1e-3, cell).astype(F)),
                     d(rng.normal(0.0, 1e-3, cell).astype(F)) if theta else None))
    for nrk in (1, 2):
        print("batched compose", ns, sides, nrk, flush=True)
        want = []
        for src, t0, a0, held, divterm in rows:
            out = torch.zeros(cell, dtype=torch.float32, device=device)
            launch_scalar_stage_from_carrier(src, t0, mass, d(m), d(divx), a0, held, divterm, out, nrk=nrk,
                                             weps=1e-17, copy_back=False, **grid)
            want.append(out)
        got = [torch.full(cell, float("nan"), dtype=torch.float32, device=device) for _ in rows]
        cols = list(zip(*rows))
        launch_scalar_stages_from_carriers(list(cols[0]), list(cols[1]), mass, d(m), d(divx), list(cols[2]),
                                           list(cols[3]), list(cols[4]), got, nrk=nrk, **grid)
        torch.cuda.synchronize()
        for q in range(ns):
            assert_bits_equal(to_host(got[q]), to_host(want[q]), f"scalar {q} of {ns}, nrk={nrk}")


def test_batched_compose_refusals():
    """The batch launcher refuses the final stage, more than four scalars and outputs that alias."""
    from gm1_amd.scalar_moist_rk import launch_scalar_stages_from_carriers

    kw = dict(ni=8, nj=8, nk=4, dt=1.0, dx=1.0, dy=1.0, rdz=1.0, lateral_sides=(1, 1, 1, 1))
    with pytest.raises(ValueError):
        launch_scalar_stages_from_carriers([None] * 5, [None] * 5, {}, {}, None, [None] * 5, [None] * 5, [None] * 5,
                                           [None] * 5, nrk=1, **kw)
    with pytest.raises(ValueError):
        launch_scalar_stages_from_carriers([None], [None], {}, {}, None, [None], [None], [None], [None], nrk=3, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(3, 4, 2, 2), (1, 1, 1, 1), (2, 2, 4, 3)])
@pytest.mark.parametrize("row", ["theta-nrk1", "theta-nrk3", "qv-pdef", "qc-pdef"])
def test_gpu_stage_from_unpadded_carrier_bit_exact(device, sides, row):
    """The carrier entry points (pack and lateral fill formed in the tiles' loads) against the oracle's
    pack + fill + advection + compose, on rigid, periodic, open and mixed sides."""
    import torch

    from gm1_amd.scalar_moist_rk import launch_scalar_stage_from_carrier
    from oracle import boundary as obc
    from oracle import stage as ost
    from oracle.advection5 import cpu_scalar_advection5_inplace

    ni, nj, nk = 33, 19, 12
    rng = np.random.default_rng(300 + len(row) + sum(sides))
    a, m, divx = _moist_inputs(rng, ni, nj, nk)
    sh = shapes(ni, nj, nk)
    cell = (nk, nj, ni)
    dt, dx, dy, dz = 1.5, 250.0, 300.0, 200.0
    s = np.zeros(sh["s"], F)
    d = lambda x: to_dev(x, device)  # noqa: E731
    mass = {k: d(a[k]) for k in ("rru", "rrv", "rrw")}
    scratch = torch.zeros(cell, dtype=torch.float32, device=device)
    rdz = float(F(1.0) / F(dz))
    if row.startswith("theta"):
        nrk = 1 if row.endswith("1") else 3
        th0 = (F(300.0) + np.linspace(0.0, 40.0, nk)).astype(F)
        th = rng.normal(0.0, 2.0, cell).astype(F)
        ost.pack_theta(th0, th, s, ni=ni, nj=nj, nk=nk)
        obc.fill(s, obc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
        tha = rng.normal(0.0, 2.0, cell).astype(F)
        held = rng.normal(0.0, 1e-3, cell).astype(F)
        divterm = rng.normal(0.0, 1e-3, cell).astype(F)
        adv = np.zeros(cell, F)
        fl = {"x": np.zeros(sh["u"], F), "y": np.zeros(sh["v"], F), "z": np.zeros(sh["w"], F)}
        weps = float(F(F(10.0) * F(1e-18)))
        cpu_scalar_advection5_inplace(s, a, m, divx, adv, fl, ni=ni, nj=nj, nk=nk, nrk=nrk, dx=dx, dy=dy, dz=dz,
                                      weps=weps, lateral_sides=sides)
        want = ost.compose_theta(tha, held, divterm, adv, F(F(dt) / F(4 - nrk)))
        carrier = d(th)
        launch_scalar_stage_from_carrier(carrier, d(th0), mass, d(m), d(divx), d(tha), d(held), d(divterm), scratch,
                                         ni=ni, nj=nj, nk=nk, nrk=nrk, dt=dt, dx=dx, dy=dy, rdz=rdz,
                                         lateral_sides=sides, weps=weps)
    else:
        species = 0 if row.startswith("qv") else 1
        q = np.abs(rng.normal(0.0, 1e-3, cell)).astype(F)
        q[rng.random(cell) < 0.3] = F(0.0)
        qa = np.abs(q + rng.normal(0.0, 2e-4, cell)).astype(F)
        held = rng.normal(0.0, 1e-7, cell).astype(F)
        s[1 : nk + 1, NG : NG + nj, NG : NG + ni] = q
        obc.fill(s, obc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
        weps = float(F(F(1e-2) * F(1e-18)))
        adv = op.pdef_advection(qa, s, a, m, divx, ni=ni, nj=nj, nk=nk, dt=F(dt), dx=dx, dy=dy, dz=dz, weps=weps,
                                pdefweno=species != 0, sides=sides)
        want = ost.compose_moisture(qa, held, adv, F(dt))
        carrier = d(q)
        ps = {n: torch.zeros(cell, dtype=torch.float32, device=device) for n in ("advx", "advy", "advz")}
        launch_scalar_stage_from_carrier(carrier, None, mass, d(m), d(divx), d(qa), d(held), None, scratch,
                                         ni=ni, nj=nj, nk=nk, nrk=3, dt=dt, dx=dx, dy=dy, rdz=rdz,
                                         lateral_sides=sides, pdef_species=species, pdef_scratch=ps)
    torch.cuda.synchronize()
    assert_bits_equal(to_host(carrier), want, f"{row} sides={sides}")
