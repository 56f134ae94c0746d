"""Pins for the oracle functions whose reference form is CUDA kernel text only.

The reference ships no CPU transcription of the buoyancy interpolation, the
post-acoustic getdiv/getdivx, the positive-definite flux/directional/total
kernels, the held compose (``cuda_moist_stage_bridge.py:177-194``,
``cuda_scalar_moist_rk.py:592-664``, ``cuda_gm1_forecast.py:554-577``).  Each
test below ties one restatement to something outside its own text:

* a routine pinned to the reference's golden vectors that it must reduce to
  for a special class of inputs (getdivx -> ``cpu_terrain_mass_flux5_inplace``,
  getdiv -> getdivx at unit density, the positive-definite chain -> the WENO
  tendency of ``cpu_scalar_advection5_inplace``);
* a closed form of the mathematics (linear profiles interpolate exactly,
  a linear wind has a constant divergence, the Exner tendency is the chain
  rule of the equation of state);
* where a check is aimed at a plausible mistake (exchanged interpolation
  weights, a slope or level index off by one, the wrong vertical closure), a
  companion run plants that mistake in the oracle's inputs or, through
  ``monkeypatch``, in one of its helpers, and shows the check fails.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, NG, assert_bits_equal, seeded_metrics, shapes
from oracle import advection5 as oa
from oracle import boundary as obc
from oracle import pdef as opdef
from oracle import stage as ost
from oracle.common import terrain_static_shapes

f64 = np.float64


def _flat_metrics(ni, nj, nk):
    m = {name: np.ones(shape, F) for name, shape in terrain_static_shapes(ni, nj, nk).items()}
    for name in ("dzdx", "dzdy", "gxu", "gyv", "gx", "gy"):
        m[name][:] = F(0.0)
    return m


def _geometric_column(nk, dz0=100.0, ratio=1.12):
    """Faces zf[f] (f = 0..nk, w level kf = f+1) and centres zh[c] (c = 0..nk-1, cell k = c+1)."""
    dzs = dz0 * ratio ** np.arange(nk, dtype=f64)
    zf = np.concatenate([[0.0], np.cumsum(dzs)])
    zh = 0.5 * (zf[:-1] + zf[1:])
    return zf, zh


def _interpolation_weights(zf, zh, nj, ni):
    """CM1's c1 (lower cell) and c2 (upper cell) weights of w level f between cells f-1 and f."""
    nk = len(zh)
    c1 = np.zeros((nk + 1, nj, ni), F)
    c2 = np.zeros((nk + 1, nj, ni), F)
    for f in range(1, nk):
        lam = (zf[f] - zh[f - 1]) / (zh[f] - zh[f - 1])
        c2[f] = F(lam)
        c1[f] = F(1.0 - lam)
    return c1, c2


def _buoyancy_arrays(ni, nj, nk):
    sh = shapes(ni, nj, nk)
    return {"dum8": np.zeros(sh["s"], F), "wten": np.zeros(sh["w"], F), "uten": np.zeros(sh["u"], F),
            "vten": np.zeros(sh["v"], F)}


# ----------------------------------------------------------------------------- buoyancy


def _w_buoyancy_error(swap):
    ni, nj, nk = 6, 5, 12
    zf, zh = _geometric_column(nk)
    m = _flat_metrics(ni, nj, nk)
    c1, c2 = _interpolation_weights(zf, zh, nj, ni)
    m["c1"], m["c2"] = (c2, c1) if swap else (c1, c2)
    a = _buoyancy_arrays(ni, nj, nk)
    A, C, B = 0.3, 2.0e-4, 1.0e-4
    x = (np.arange(ni) + 0.5) * 1000.0
    a["dum8"][1 : nk + 1, NG : NG + nj, NG : NG + ni] = (A + C * zh[:, None, None] + B * x[None, None, :]).astype(F)
    ost.buoyancy(a, m, ni=ni, nj=nj, nk=nk)
    want = A + C * zf[1:nk, None, None] + B * x[None, None, :]
    got = a["wten"][2 : nk + 1, NG : NG + nj, NG : NG + ni].astype(f64)
    assert not np.any(a["wten"][1]) and not np.any(a["wten"][nk + 1])
    return float(np.max(np.abs(got - want)) / np.max(np.abs(want)))


def test_buoyancy_w_term_interpolates_a_linear_profile_on_a_stretched_column():
    """c1*b(k-1) + c2*b(k) at a w level is the linear profile's value there (cuda_moist_stage_bridge.py:180)."""
    assert _w_buoyancy_error(swap=False) < 4e-7
    # with the weights exchanged the same profile misses by the column's stretching
    assert _w_buoyancy_error(swap=True) > 1e-3


def _uv_buoyancy(mutate=None):
    ni, nj, nk = 7, 6, 9
    dx, dy, dz = 1000.0, 800.0, 200.0
    rng = np.random.default_rng(21)
    m = _flat_metrics(ni, nj, nk)
    m["c1"][1:nk] = F(0.5)
    m["c2"][1:nk] = F(0.5)
    gxu = (0.02 * rng.standard_normal(m["gxu"].shape)).astype(F)
    gyv = (0.02 * rng.standard_normal(m["gyv"].shape)).astype(F)
    m["gxu"][:] = gxu
    m["gyv"][:] = gyv
    if mutate == "level":  # the slope pair of the cell above: levels k, k+1 instead of k-1, k
        m["gxu"][:-1] = gxu[1:]
        m["gyv"][:-1] = gyv[1:]
    A, B, D, C = 0.3, 1.0e-4, -2.0e-4, 3.0e-4
    zh = (np.arange(nk) + 0.5) * dz
    x = (np.arange(ni) + 0.5) * dx
    y = (np.arange(nj) + 0.5) * dy
    d8 = A + B * x[None, None, :] + D * y[None, :, None] + C * zh[:, None, None]
    a = _buoyancy_arrays(ni, nj, nk)
    a["dum8"][1 : nk + 1, NG : NG + nj, NG : NG + ni] = d8.astype(F)
    ost.buoyancy(a, m, ni=ni, nj=nj, nk=nk)
    gxu, gyv = gxu.astype(f64), gyv.astype(f64)
    # u faces f = 1..ni-1 at x = f*dx; v faces g = 1..nj-1 at y = g*dy.  For a linear buoyancy
    # the four w-level values around a face average to the two cells' average at interior
    # levels, so the term vanishes; at k = 1 and nk the absent w level (b = 0) leaves
    # (-X + C*(zf_other - 2 zh))/4 times the sum of the face's two slope levels, X the
    # face's value of A + B x + D y.
    xf = np.arange(1, ni) * dx
    yf = np.arange(1, nj) * dy
    Xu = A + B * xf[None, :] + D * y[:, None]
    Xv = A + B * x[None, :] + D * yf[:, None]
    want_u = np.zeros((nk, nj, ni - 1))
    want_v = np.zeros((nk, nj - 1, ni))
    for k, ztail in ((1, C * (dz - 2 * zh[0])), (nk, C * ((nk - 1) * dz - 2 * zh[nk - 1]))):
        want_u[k - 1] = (-Xu + ztail) / 4.0 * (gxu[k - 1, :, 1:ni] + gxu[k, :, 1:ni])
        want_v[k - 1] = (-Xv + ztail) / 4.0 * (gyv[k - 1, 1:nj, :] + gyv[k, 1:nj, :])
    got_u = a["uten"][1 : nk + 1, NG : NG + nj, NG + 1 : NG + ni].astype(f64)
    got_v = a["vten"][1 : nk + 1, NG + 1 : NG + nj, NG : NG + ni].astype(f64)
    scale_u, scale_v = np.abs(want_u).max(), np.abs(want_v).max()
    return float(np.max(np.abs(got_u - want_u))) / scale_u, float(np.max(np.abs(got_v - want_v))) / scale_v


def test_buoyancy_terrain_terms_of_a_linear_buoyancy_have_their_closed_form():
    """u/v terrain buoyancy (cuda_moist_stage_bridge.py:181-182): zero at interior levels, closed form at the ends."""
    eu, ev = _uv_buoyancy()
    assert eu < 1e-5 and ev < 1e-5
    mu, mv = _uv_buoyancy(mutate="level")
    assert mu > 1e-2 and mv > 1e-2


# ----------------------------------------------------------------------------- getdivx / getdiv


def _winds(rng, ni, nj, nk):
    sh = shapes(ni, nj, nk)
    return ((3.0 * rng.standard_normal(sh["u"])).astype(F), (3.0 * rng.standard_normal(sh["v"])).astype(F),
            rng.standard_normal(sh["w"]).astype(F))


def _mass_flux_pair(unit_rgz):
    ni, nj, nk = 9, 8, 7
    dx, dy, dz = 250.0, 200.0, 150.0
    rng = np.random.default_rng(31)
    m = seeded_metrics(rng, ni, nj, nk, terrain=True)
    if unit_rgz:
        m["rgzu"][:] = F(1.0)
        m["rgzv"][:] = F(1.0)
    u, v, w = _winds(rng, ni, nj, nk)
    sh = shapes(ni, nj, nk)
    a1 = {"u3d": u, "v3d": v, "w3d": w, "rru": np.zeros(sh["u"], F), "rrv": np.zeros(sh["v"], F),
          "rrw": np.zeros(sh["w"], F)}
    d1 = np.zeros((nk, nj, ni), F)
    oa.cpu_terrain_mass_flux5_inplace(a1, m, d1, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    a2 = {"rru": u.copy(), "rrv": v.copy(), "rrw": w.copy()}
    d2 = np.zeros((nk, nj, ni), F)
    ost.getdivx(a2, m, d2, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    return (ni, nj, nk), a1, d1, a2, d2


def test_getdivx_reduces_to_the_golden_pinned_terrain_mass_flux():
    """With rgzu = rgzv = 1 the two associations coincide: every word the CPU form writes is getdivx's.

    getdivx rebuilds rru as ((0.5*(rho0_l+rho0_r))*avg)*rgzu (:190), the CPU form
    as (0.5*(rho0_l+rho0_r))*(u*rgzu) (cuda_advection5.py:328-341); the vertical
    mass flux and divx then follow the same statements.
    """
    (ni, nj, nk), a1, d1, a2, d2 = _mass_flux_pair(unit_rgz=True)
    ks = slice(1, nk + 1)
    assert_bits_equal(a2["rru"][ks, NG : NG + nj, NG + 1 : NG + ni], a1["rru"][ks, NG : NG + nj, NG + 1 : NG + ni], "rru")
    assert_bits_equal(a2["rrv"][ks, NG + 1 : NG + nj, NG : NG + ni], a1["rrv"][ks, NG + 1 : NG + nj, NG : NG + ni], "rrv")
    yi, xi = slice(NG + 1, NG + nj - 1), slice(NG + 1, NG + ni - 1)
    assert_bits_equal(a2["rrw"][1 : nk + 2, yi, xi], a1["rrw"][1 : nk + 2, yi, xi], "rrw")
    assert_bits_equal(d2[:, 1 : nj - 1, 1 : ni - 1], d1[:, 1 : nj - 1, 1 : ni - 1], "divx")
    assert np.abs(d1).max() > 1e-3


def test_getdivx_reduction_sees_a_misplaced_terrain_level(monkeypatch):
    """A w level's terrain term taken one level up changes rrw and divx, so the reduction above would fail."""
    good = ost._uniform_weights
    monkeypatch.setattr(ost, "_uniform_weights", lambda kf, dz: good(kf + 1, dz))
    (ni, nj, nk), a1, d1, a2, d2 = _mass_flux_pair(unit_rgz=True)
    yi, xi = slice(NG + 1, NG + nj - 1), slice(NG + 1, NG + ni - 1)
    assert not np.array_equal(a2["rrw"][2 : nk + 1, yi, xi], a1["rrw"][2 : nk + 1, yi, xi])
    assert not np.array_equal(d2[:, 1 : nj - 1, 1 : ni - 1], d1[:, 1 : nj - 1, 1 : ni - 1])


def test_getdivx_horizontal_fluxes_agree_within_rounding_for_any_metric():
    (ni, nj, nk), a1, _, a2, _ = _mass_flux_pair(unit_rgz=False)
    ks = slice(1, nk + 1)
    for name, sl in (("rru", (ks, slice(NG, NG + nj), slice(NG + 1, NG + ni))),
                     ("rrv", (ks, slice(NG + 1, NG + nj), slice(NG, NG + ni)))):
        x, y = a2[name][sl].astype(f64), a1[name][sl].astype(f64)
        assert np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-30)) <= 2.5 * 2.0 ** -24 * 2


def test_getdiv_is_getdivx_at_unit_density():
    """getdiv's transformed winds (:186-188) are getdivx's mass fluxes with rho0 = rf0 = 1, word for word."""
    ni, nj, nk = 8, 7, 6
    dx, dy, dz = 300.0, 250.0, 120.0
    rng = np.random.default_rng(32)
    m = seeded_metrics(rng, ni, nj, nk, terrain=True)
    u, v, w = _winds(rng, ni, nj, nk)
    thterm = (300.0 + rng.standard_normal((nk, nj, ni))).astype(F)
    a = {"rru": u, "rrv": v, "rrw": w}
    div = np.zeros((nk, nj, ni), F)
    tdiv = np.zeros((nk, nj, ni), F)
    ost.getdiv(a, m, div, tdiv, thterm, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    unit = dict(m)
    unit["rho0_3d"] = np.ones_like(m["rho0_3d"])
    unit["rf0"] = np.ones_like(m["rf0"])
    b = {"rru": u.copy(), "rrv": v.copy(), "rrw": w.copy()}
    divx = np.zeros((nk, nj, ni), F)
    ost.getdivx(b, unit, divx, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    assert_bits_equal(div, divx, "div")
    assert_bits_equal(tdiv, (div * thterm).astype(F), "theta_div")
    # getdiv leaves its inputs alone
    assert_bits_equal(a["rru"], u, "rru untouched")


@pytest.mark.parametrize("flat_column", [False, True])
def test_divergence_of_a_linear_wind_is_its_constant_gradient(flat_column):
    """u = U0 + a x, v = V0 + b y, w = c z: every cell's divergence is a + b + c."""
    ni, nj, nk = 8, 7, 9
    dx, dy, dz = 500.0, 400.0, 250.0
    ga, gb, gc = 2.0e-3, -1.5e-3, 4.0e-3
    sh = shapes(ni, nj, nk)
    m = _flat_metrics(ni, nj, nk)
    u = np.zeros(sh["u"], F)
    v = np.zeros(sh["v"], F)
    w = np.zeros(sh["w"], F)
    xf = (np.arange(sh["u"][2]) - NG) * dx
    yf = (np.arange(sh["v"][1]) - NG) * dy
    u[:] = (10.0 + ga * xf)[None, None, :].astype(F)
    v[:] = (-4.0 + gb * yf)[None, :, None].astype(F)
    zf = (np.arange(sh["w"][0]) - 1) * dz
    w[:] = (gc * zf)[:, None, None].astype(F)
    w[nk + 1] = F(0.0)
    a = {"rru": u, "rrv": v, "rrw": w}
    div = np.zeros((nk, nj, ni), F)
    tdiv = np.zeros((nk, nj, ni), F)
    thterm = np.ones((nk, nj, ni), F)
    if flat_column:
        from gm1_amd.grid import build_vertical_levels

        lv = build_vertical_levels(stretch_z=0, nk=nk, dz=dz)
        ost.getdiv_flat(a, div, tdiv, thterm, lv, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy)
    else:
        ost.getdiv(a, m, div, tdiv, thterm, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz)
    want = np.full((nk, nj, ni), ga + gb + gc)
    want[nk - 1] = ga + gb - gc * (nk - 1) * dz / dz  # the lid face carries w = 0
    assert np.max(np.abs(div.astype(f64) - want)) < 1e-4 * (abs(ga) + abs(gb) + abs(gc))


# ----------------------------------------------------------------------------- positive-definite chain


def _pdef_case(sides, seed=41, amplitude=1.0e-3, floor=0.7, dt=1.0):
    ni, nj, nk = 14, 12, 10
    dx, dy, dz = 250.0, 250.0, 200.0
    rng = np.random.default_rng(seed)
    m = seeded_metrics(rng, ni, nj, nk, terrain=True)
    sh = shapes(ni, nj, nk)
    a = {"rru": (2.0 * rng.standard_normal(sh["u"])).astype(F), "rrv": (2.0 * rng.standard_normal(sh["v"])).astype(F),
         "rrw": (0.5 * rng.standard_normal(sh["w"])).astype(F)}
    a["rrw"][1] = F(0.0)
    a["rrw"][nk + 1] = F(0.0)
    q = (amplitude * (floor + (1.0 - floor) * rng.random((nk, nj, ni)))).astype(F)
    padded = np.zeros(sh["s"], F)
    ost.pack_scalar(q, padded, ni=ni, nj=nj, nk=nk)
    obc.fill(padded, obc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
    divx = (1e-3 * rng.standard_normal((nk, nj, ni))).astype(F)
    return dict(ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz, dt=dt), m, a, q, padded, divx


@pytest.mark.parametrize("sides", [(1, 1, 1, 1), (2, 2, 2, 2), (2, 1, 3, 2)])
@pytest.mark.parametrize("pdefweno", [False, True])
def test_pdef_chain_without_negative_candidates_is_the_weno_tendency(sides, pdefweno):
    """Above the magnitude gate and with every candidate positive, the chain (:592-664) is the WENO5 tendency.

    ``cpu_scalar_advection5_inplace`` at nrk=3 is pinned to the reference's golden
    vectors; the positive-definite fluxes equal its WENO fluxes once every stencil
    value exceeds weps, the repair leaves positive candidates alone, and the total
    (((ax+ay)+az) + s*divx)*rr0*gz is its tendency.
    """
    c, m, a, q, padded, divx = _pdef_case(sides)
    ni, nj, nk = c["ni"], c["nj"], c["nk"]
    weps = float(F(F(1.0e-2) * F(1.0e-18)))
    got = opdef.pdef_advection(q, padded, a, m, divx, ni=ni, nj=nj, nk=nk, dt=c["dt"], dx=c["dx"], dy=c["dy"],
                               dz=c["dz"], weps=weps, pdefweno=pdefweno, sides=sides)
    want = np.zeros((nk, nj, ni), F)
    fl = {"x": np.zeros_like(a["rru"]), "y": np.zeros_like(a["rrv"]), "z": np.zeros_like(a["rrw"])}
    oa.cpu_scalar_advection5_inplace(padded, a, m, divx, want, fl, ni=ni, nj=nj, nk=nk, nrk=3, dx=c["dx"],
                                     dy=c["dy"], dz=c["dz"], weps=weps, lateral_sides=sides)
    assert_bits_equal(got, want, "pdef total")
    # no candidate was repaired: the check is about the unrepaired path
    fx, fy, fz = opdef.pdef_fluxes(padded, a["rru"], a["rrv"], a["rrw"], m["c1"], m["c2"], ni=ni, nj=nj, nk=nk,
                                   weps=weps, pdefweno=pdefweno)
    ax, _, _ = opdef.directional(fx, fy, fz, padded, a["rru"], a["rrv"], ni=ni, nj=nj, nk=nk, dx=c["dx"],
                                 dy=c["dy"], dz=c["dz"], sides=sides)
    _, cand = opdef.repair_direction(q, padded, m["rho0_3d"], m["gz"], m["rgz"], a["rru"], ax, axis="x", ni=ni,
                                     nj=nj, nk=nk, dt=c["dt"], spacing=c["dx"], sides=sides)
    assert (cand > 0).all()


def test_pdef_reduction_sees_an_unlimited_vertical_closure(monkeypatch):
    """With the lower-order closures left unlimited (flx3 for upstrpd) the chain leaves the WENO tendency."""
    from oracle.common import flx3

    sides = (2, 2, 2, 2)
    c, m, a, q, padded, divx = _pdef_case(sides)
    ni, nj, nk = c["ni"], c["nj"], c["nk"]
    weps = float(F(F(1.0e-2) * F(1.0e-18)))
    want = np.zeros((nk, nj, ni), F)
    fl = {"x": np.zeros_like(a["rru"]), "y": np.zeros_like(a["rrv"]), "z": np.zeros_like(a["rrw"])}
    oa.cpu_scalar_advection5_inplace(padded, a, m, divx, want, fl, ni=ni, nj=nj, nk=nk, nrk=3, dx=c["dx"],
                                     dy=c["dy"], dz=c["dz"], weps=weps, lateral_sides=sides)
    monkeypatch.setattr(opdef, "upstrpd", lambda s1, s2, s3, w: flx3(s1, s2, s3))
    got = opdef.pdef_advection(q, padded, a, m, divx, ni=ni, nj=nj, nk=nk, dt=c["dt"], dx=c["dx"], dy=c["dy"],
                               dz=c["dz"], weps=weps, pdefweno=True, sides=sides)
    assert not np.array_equal(got, want)


def test_pdef_gate_zeroes_fluxes_of_sub_weps_stencils():
    """Fluxes whose five stencil values are all <= weps vanish (the magnitude gate); the vertical closures stay."""
    sides = (1, 1, 1, 1)
    weps = float(F(F(1.0e-2) * F(1.0e-18)))
    c, m, a, q, padded, divx = _pdef_case(sides, amplitude=weps, floor=0.2)
    ni, nj, nk = c["ni"], c["nj"], c["nk"]
    assert np.abs(padded).max() <= weps
    fx, fy, fz = opdef.pdef_fluxes(padded, a["rru"], a["rrv"], a["rrw"], m["c1"], m["c2"], ni=ni, nj=nj, nk=nk,
                                   weps=weps, pdefweno=True)
    assert not np.any(fx) and not np.any(fy)
    assert not np.any(fz[4 : nk - 1])  # WENO faces 4..nk-2 are gated
    assert np.any(fz[2]) and np.any(fz[nk])  # the lower-order closures are not
    # the ungated WENO flux of the same data does not vanish
    flx = {"x": np.zeros_like(a["rru"]), "y": np.zeros_like(a["rrv"]), "z": np.zeros_like(a["rrw"])}
    t = np.zeros((nk, nj, ni), F)
    oa.cpu_scalar_advection5_inplace(padded, a, m, divx, t, flx, ni=ni, nj=nj, nk=nk, nrk=3, dx=c["dx"], dy=c["dy"],
                                     dz=c["dz"], weps=weps, lateral_sides=sides)
    assert np.any(flx["x"])


def test_pdef_directional_tendencies_of_a_periodic_row_sum_to_zero_through_the_repair():
    """Flux differences telescope along a periodic row; the donor repair moves mass without creating it."""
    sides = (1, 1, 1, 1)
    c, m, a, q, padded, divx = _pdef_case(sides, seed=43, amplitude=1.0e-3, floor=-0.6, dt=40.0)
    ni, nj, nk = c["ni"], c["nj"], c["nk"]
    for name in ("gz", "rgz"):
        m[name][:] = F(1.0)
    m["rho0_3d"][:] = m["rho0_3d"][:, :1, :1]
    # periodic duplicate faces: face ni is face 0, face nj is face 0
    a["rru"][:, :, NG + ni] = a["rru"][:, :, NG]
    a["rrv"][:, NG + nj, :] = a["rrv"][:, NG, :]
    weps = float(F(F(1.0e-2) * F(1.0e-18)))
    fx, fy, fz = opdef.pdef_fluxes(padded, a["rru"], a["rrv"], a["rrw"], m["c1"], m["c2"], ni=ni, nj=nj, nk=nk,
                                   weps=weps, pdefweno=True)
    ax, ay, az = opdef.directional(fx, fy, fz, padded, a["rru"], a["rrv"], ni=ni, nj=nj, nk=nk, dx=c["dx"],
                                   dy=c["dy"], dz=c["dz"], sides=sides)
    scale = float(np.abs(ax).max())
    assert np.abs(ax.astype(f64).sum(axis=2)).max() < 1e-5 * scale * ni
    rx, cand = opdef.repair_direction(q, padded, m["rho0_3d"], m["gz"], m["rgz"], a["rru"], ax, axis="x", ni=ni,
                                      nj=nj, nk=nk, dt=c["dt"], spacing=c["dx"], sides=sides)
    assert (cand < 0).any(), "the case must exercise the repair"
    assert not np.array_equal(rx, ax)
    assert np.abs(rx.astype(f64).sum(axis=2)).max() < 1e-4 * scale * ni


# ----------------------------------------------------------------------------- held compose


def _eos_exner(rho, theta, qv):
    rd, rv, cp = 287.04, 461.5, 1005.7
    return (rho * theta * (rd + rv * qv) / 1.0e5) ** (rd / (cp - rd))


def _held_case(ht, qvt, seed=51):
    ni, nj, nk = 6, 5, 8
    rng = np.random.default_rng(seed)
    th0 = np.linspace(300.0, 330.0, nk)
    pi0 = np.linspace(1.0, 0.8, nk)
    base = {"th0": th0.astype(F), "pi0": pi0.astype(F)}
    rho = (1.1 - 0.05 * np.arange(nk)[:, None, None] + 0.01 * rng.random((nk, nj, ni))).astype(F)
    tha = (0.5 * rng.standard_normal((nk, nj, ni))).astype(F)
    qa = np.zeros((3, nk, nj, ni), F)
    qa[0] = (0.01 + 0.002 * rng.random((nk, nj, ni))).astype(F)
    pi = _eos_exner(rho.astype(f64), th0[:, None, None] + tha.astype(f64), qa[0].astype(f64))
    ppi = (pi - pi0[:, None, None]).astype(F)
    thheld = np.full((nk, nj, ni), ht, F)
    qheld = np.zeros((3, nk, nj, ni), F)
    qheld[0] = F(qvt)
    rates = {n: np.zeros((nk, nj, ni), F) for n in ("qpten", "qtten", "qvten", "qcten")}
    return (ni, nj, nk), base, rho, ppi, tha, qa, thheld, qheld, rates, pi


@pytest.mark.parametrize("ht,qvt", [(0.1, 0.0), (0.0, 1.0e-4), (-0.05, 5.0e-5)])
def test_held_compose_exner_tendency_is_the_chain_rule_of_the_equation_of_state(ht, qvt):
    """ppten = d(pi)/dt of pi = (rho theta (rd + rv qv)/p00)^(rd/cv) along the held theta and qv tendencies."""
    (ni, nj, nk), base, rho, ppi, tha, qa, thheld, qheld, rates, pi = _held_case(ht, qvt)
    ppten = np.zeros((nk + 2, nj + 2, ni + 2), F)
    dt = 1.5
    ost.predict_and_compose_held(base, rho, ppi, tha, qa, thheld, qheld, ppten, rates, dt, ni=ni, nj=nj, nk=nk)
    rd, rv, cp = 287.04, 461.5, 1005.7
    theta = base["th0"].astype(f64)[:, None, None] + tha.astype(f64)
    qv = qa[0].astype(f64)
    want = rd / (cp - rd) * pi * (ht / theta + rv * qvt / (rd + rv * qv))
    got = ppten[1 : nk + 1, 1 : nj + 1, 1 : ni + 1].astype(f64)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 2e-2


def test_held_compose_without_held_theta_or_qv_tendency_passes_the_rates_through():
    (ni, nj, nk), base, rho, ppi, tha, qa, thheld, qheld, rates, _ = _held_case(0.0, 0.0)
    rng = np.random.default_rng(52)
    for n in rates:
        rates[n][:] = (1e-4 * rng.standard_normal(rates[n].shape)).astype(F)
    qc_before = qheld[1].copy()
    ppten = np.zeros((nk + 2, nj + 2, ni + 2), F)
    ost.predict_and_compose_held(base, rho, ppi, tha, qa, thheld, qheld, ppten, rates, 1.5, ni=ni, nj=nj, nk=nk)
    assert_bits_equal(ppten[1 : nk + 1, 1 : nj + 1, 1 : ni + 1], rates["qpten"], "ppten")
    assert_bits_equal(thheld, rates["qtten"], "thten1")
    assert_bits_equal(qheld[0], rates["qvten"], "qvten")
    assert_bits_equal(qheld[1], (qc_before + rates["qcten"]).astype(F), "qcten")
