"""The flat stretched column (CM1 stretch_z=1 on a flat grid).

The reference has no CPU transcription of these operators (they are kernel
texts derived from the terrain texts, cuda_gm1_child_stretched_*.py and
cuda_gm1_child_psolver3_open.py:340-535); the oracle restates those texts.
Pins on the CPU:

* on a uniform column (mh = mf = 1) the flat scalar advection, pdef chain,
  momentum advection (periodic sides) and getdiv/getdivx give the same
  values as the terrain family, whose oracle is pinned to reference golden
  vectors;
* on a stretched column a constant scalar in any flow has no advective
  tendency (the vertical factor of the scalar tendency is the one divx
  carries), and the flat divx obeys the discrete divergence theorem
  (sum over a periodic domain of divx * dz/mh(k) vanishes);
* a stretched state at rest stays exactly at rest through a moist large step;
* the large step is translation-equivariant on a periodic stretched domain.

GPU: every flat operator and the composed moist large step on a stretched
column reproduce the oracle bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, NG, assert_bits_equal, to_host
from gm1_amd.fields import CaseConfig
from gm1_amd.large_step import host_initial_state
from oracle import advection5 as oadv
from oracle import large_step as ol
from oracle import momentum5 as omom
from oracle import pdef as opdef
from oracle import sound as osound
from oracle import stage as ost

PROGNOSTIC = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ppd")
MOIST_FIELDS = PROGNOSTIC + ("q3d", "rho", "prs", "rr", "rf", "rain", "prate", "qpten", "qtten", "qvten", "qcten",
                             "fall_velocity")


def stretched_case(ni, nj, nk, sides, *, seed=7, amp=3.0, noise=0.1, wind="hodograph", uniform=False,
                   saturated_bubble=True):
    """A moist warm bubble on a stretch_z=1 column (20 m at the ground, 180 m aloft, 1.6 km top)."""
    stretch = None if uniform else (1600.0, 100.0, 700.0, 20.0, 180.0)
    dz = 100.0
    return CaseConfig(f"stretched-{ni}x{nj}x{nk}", ni, nj, nk, 100.0, 100.0, dz, 0.6, 10, tuple(sides), True, wind,
                      ((amp, ni * 50.0, nj * 50.0, 600.0),), bubble_radius=(ni * 30.0, 500.0), noise=noise,
                      seed=seed, extra={"bubble_rh": 1.03} if saturated_bubble else {}, stretched=stretch)


def cfg_dict(cfg):
    return {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
            "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}


def oracle_steps(host, n, *, mass_correction=None):
    cfg = host["cfg"]
    levels = host["levels"]
    a = {k: v.copy() for k, v in host["arrays"].items()}
    for _ in range(n):
        ol.large_step(a, host["metrics"], host["base"], host["prof"], cfg_dict(cfg), zt=F(levels.ztop),
                      mass_correction=mass_correction, flat=levels)
    return a


def uniform_flat(nk, dz):
    from gm1_amd.grid import build_vertical_levels

    return build_vertical_levels(stretch_z=0, nk=nk, dz=dz)


def evolved(sides=(2, 2, 2, 2), ni=16, nj=12, nk=16, steps=1, uniform=False):
    host = host_initial_state(stretched_case(ni, nj, nk, sides, uniform=uniform))
    if uniform:
        host["levels"] = uniform_flat(nk, 100.0)
    a = oracle_steps(host, steps)
    return host, a


def same_values(x, y, msg):
    x, y = np.asarray(x), np.asarray(y)
    bad = ~((x == y) | (np.isnan(x) & np.isnan(y)))
    assert not bad.any(), f"{msg}: {int(bad.sum())} values differ"


# ----------------------------------------------------------------------------- CPU


def test_uniform_column_scalar_advection_equals_the_terrain_family():
    host, a = evolved(uniform=True)
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    padded = np.zeros((nk + 2, nj + 2 * NG, ni + 2 * NG), np.float32)
    ost.pack_theta(np.asarray(host["base"]["th0"]), a["th3d"], padded, ni=ni, nj=nj, nk=nk)
    from oracle import boundary as bc

    bc.fill(padded, bc.KIND_SCALAR, cfg.lateral_sides, ni=ni, nj=nj, nk=nk)
    out = {}
    for flat in (None, lv):
        t = np.zeros((nk, nj, ni), np.float32)
        fl = {"x": np.zeros((nk + 2, nj + 2 * NG, ni + 1 + 2 * NG), np.float32),
              "y": np.zeros((nk + 2, nj + 1 + 2 * NG, ni + 2 * NG), np.float32),
              "z": np.zeros((nk + 3, nj + 2 * NG, ni + 2 * NG), np.float32)}
        oadv.cpu_scalar_advection5_inplace(padded, a, m, a["divx"], t, fl, ni=ni, nj=nj, nk=nk, nrk=3, dx=cfg.dx,
                                           dy=cfg.dy, dz=cfg.dz, weps=ol.THETA_WEPS, lateral_sides=cfg.lateral_sides,
                                           flat=flat)
        out[flat is None] = t
    assert np.abs(out[True]).max() > 0
    same_values(out[True], out[False], "scalar tendency")


def test_uniform_column_pdef_chain_equals_the_terrain_family():
    host, a = evolved(uniform=True)
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    padded = np.zeros((nk + 2, nj + 2 * NG, ni + 2 * NG), np.float32)
    ost.pack_scalar(a["q3d"][1], padded, ni=ni, nj=nj, nk=nk)
    from oracle import boundary as bc

    bc.fill(padded, bc.KIND_SCALAR, cfg.lateral_sides, ni=ni, nj=nj, nk=nk)
    res = [opdef.pdef_advection(a["qa"][1], padded, a, m, a["divx"], ni=ni, nj=nj, nk=nk, dt=F(0.6), dx=cfg.dx,
                                dy=cfg.dy, dz=cfg.dz, weps=ol.MOISTURE_WEPS, pdefweno=True, sides=cfg.lateral_sides,
                                flat=flat) for flat in (None, lv)]
    assert np.abs(res[0]).max() > 0
    same_values(res[0], res[1], "pdef advection")


def test_uniform_column_momentum_equals_the_terrain_family_on_periodic_sides():
    host, a = evolved(sides=(1, 1, 1, 1), uniform=True)
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    outs = []
    for flat in (None, lv):
        b = {k: v.copy() for k, v in a.items()}
        for n in ("uten", "vten", "wten"):
            b[n][...] = F(0.0)
        divx = np.zeros((nk, nj, ni), np.float32)
        omom.cpu_terrain_open_momentum_advection5_inplace(b, m, divx, ni=ni, nj=nj, nk=nk, nrk=2, dx=cfg.dx,
                                                          dy=cfg.dy, dz=cfg.dz, lateral_sides=(1, 1, 1, 1), flat=flat)
        outs.append((b, divx))
    for n in ("uten", "vten", "wten", "rru", "rrv", "rrw"):
        same_values(outs[0][0][n], outs[1][0][n], n)
    assert np.abs(outs[0][0]["wten"]).max() > 0
    big = np.abs(outs[0][1]) >= 1e-20
    same_values(outs[0][1][big], outs[1][1][big], "divx")


def test_stretched_constant_scalar_has_no_advective_tendency():
    host, a = evolved()
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    # getdivx's mass fluxes and divx of the evolved (stretched, divergent) flow
    b = {k: v.copy() for k, v in a.items()}
    ost.getdivx_flat(b, m, b["divx"], lv, ni=ni, nj=nj, nk=nk, dx=cfg.dx, dy=cfg.dy)
    padded = np.full((nk + 2, nj + 2 * NG, ni + 2 * NG), F(2.5), np.float32)
    t = np.zeros((nk, nj, ni), np.float32)
    fl = {"x": np.zeros((nk + 2, nj + 2 * NG, ni + 1 + 2 * NG), np.float32),
          "y": np.zeros((nk + 2, nj + 1 + 2 * NG, ni + 2 * NG), np.float32),
          "z": np.zeros((nk + 3, nj + 2 * NG, ni + 2 * NG), np.float32)}
    oadv.cpu_scalar_advection5_inplace(padded, b, m, b["divx"], t, fl, ni=ni, nj=nj, nk=nk, nrk=1, dx=cfg.dx,
                                       dy=cfg.dy, dz=cfg.dz, weps=ol.THETA_WEPS, lateral_sides=(1, 1, 1, 1), flat=lv)
    # scale: the largest single face-flux term the tendency sums (roundoff is relative to it)
    vz = np.abs(fl["z"][1 : nk + 2, NG : NG + nj, NG : NG + ni] * F(lv.rdz)).max() * float(lv.mh[1 : nk + 1].max())
    vx = np.abs(fl["x"][1 : nk + 1, NG : NG + nj, NG : NG + ni + 1]).max() / cfg.dx
    assert vz > 0 and vx > 0
    assert np.abs(t).max() <= 1e-5 * max(vx, vz) * float(np.max(m["rr0"]))
    # with mf in place of mh the tendency would not vanish
    wrong = type("W", (), {"mh": lv.mf[: nk + 2], "rdz": lv.rdz})()
    t2 = np.zeros_like(t)
    oadv.cpu_scalar_advection5_inplace(padded, b, m, b["divx"], t2, fl, ni=ni, nj=nj, nk=nk, nrk=1, dx=cfg.dx,
                                       dy=cfg.dy, dz=cfg.dz, weps=ol.THETA_WEPS, lateral_sides=(1, 1, 1, 1),
                                       flat=wrong)
    assert np.abs(t2).max() > 100 * np.abs(t).max()


def test_stretched_divx_obeys_the_divergence_theorem():
    host, a = evolved(sides=(1, 1, 1, 1))
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    b = {k: v.copy() for k, v in a.items()}
    ost.getdivx_flat(b, m, b["divx"], lv, ni=ni, nj=nj, nk=nk, dx=cfg.dx, dy=cfg.dy)
    depth = (np.float64(lv.dz) / lv.mh[1 : nk + 1].astype(np.float64))[:, None, None]
    total = float(np.sum(b["divx"].astype(np.float64) * depth))
    scale = float(np.sum(np.abs(b["divx"].astype(np.float64)) * depth))
    assert scale > 0 and abs(total) < 1e-5 * scale


def test_stretched_state_at_rest_stays_exactly_at_rest():
    cfg = stretched_case(12, 10, 16, (2, 2, 2, 2), amp=0.0, noise=0.0, saturated_bubble=False)
    host = host_initial_state(cfg)
    for name in ("u3d", "ua", "v3d", "va"):
        host["arrays"][name][...] = F(0.0)
    a = oracle_steps(host, 1)
    for name in PROGNOSTIC:
        assert not np.any(a[name]), name
    assert np.array_equal(a["q3d"], host["arrays"]["q3d"])


def test_periodic_stretched_step_is_translation_equivariant():
    ni, nj, nk = 16, 12, 16
    h0 = host_initial_state(stretched_case(ni, nj, nk, (1, 1, 1, 1)))
    h1 = host_initial_state(stretched_case(ni, nj, nk, (1, 1, 1, 1)))
    di, dj = 5, 3
    for name in ("th3d", "tha"):
        h1["arrays"][name][...] = np.roll(h0["arrays"][name], (dj, di), axis=(1, 2))
    for name in ("q3d", "qa"):
        h1["arrays"][name][...] = np.roll(h0["arrays"][name], (dj, di), axis=(2, 3))
    for name in ("prs", "rho", "rr"):
        h1["arrays"][name][...] = np.roll(h0["arrays"][name], (dj, di), axis=(1, 2))
    a0, a1 = oracle_steps(h0, 1), oracle_steps(h1, 1)
    for name in ("th3d", "pp3d"):
        src = a0[name] if name == "th3d" else a0[name][1 : nk + 1, 1 : nj + 1, 1 : ni + 1]
        dst = a1[name] if name == "th3d" else a1[name][1 : nk + 1, 1 : nj + 1, 1 : ni + 1]
        assert np.array_equal(np.roll(src, (dj, di), axis=(1, 2)).view(np.uint32), dst.view(np.uint32)), name
    w0 = a0["w3d"][:, NG : NG + nj, NG : NG + ni]
    w1 = a1["w3d"][:, NG : NG + nj, NG : NG + ni]
    assert np.array_equal(np.roll(w0, (dj, di), axis=(1, 2)).view(np.uint32), w1.view(np.uint32))


ACOUSTIC_OUTPUTS = ("u3d", "v3d", "w3d", "pp3d", "ppd", "rru", "rrv", "rrw")


def _flat_and_terrain_stages(sides, uniform, *, flat_levels=None, flat_metrics=None):
    """Three acoustic stages of the flat column and of the terrain-coordinate stage on the same flat grid.

    The terrain-coordinate stage (``cpu_psolver3_stage_inplace``, pinned to the
    reference's golden three-stage vectors) runs on sigma = zh, sigmaf = zf with
    gz = 1 and zero slopes; there rds = rdz*mf and rdsf = rdz*mh, so the two
    stages are the same equations in different association.  Returns the
    worst relative max-norm difference per output over the three stages.
    """
    host, a = evolved(sides=sides, uniform=uniform)
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    assert not np.any(m["dzdx"]) and not np.any(m["gxu"]) and np.all(m["gz"] == 1)
    ost.stage_terms(a, host["base"], a["th3d"], ol._species6(a["q3d"]), ni=ni, nj=nj, nk=nk, sides=sides)
    ost.buoyancy_flat(a, m, ni=ni, nj=nj, nk=nk)
    b = {k: v.copy() for k, v in a.items()}
    prof = osound.terrain_coordinate_profiles(nk, cfg.dz, sigma=lv.sigma, sigmaf=lv.sigmaf)
    worst = dict.fromkeys(ACOUSTIC_OUTPUTS, 0.0)
    for nrk in (1, 2, 3):
        osound.cpu_flat_stage_inplace(a, flat_metrics or m, flat_levels or lv, ni=ni, nj=nj, nk=nk, dt=cfg.dt,
                                      nsound=cfg.nsound, nrk=nrk, dx=cfg.dx, dy=cfg.dy, lateral_sides=sides)
        osound.cpu_psolver3_stage_inplace(b, m, ni=ni, nj=nj, nk=nk, dt=cfg.dt, nsound=cfg.nsound, nrk=nrk,
                                          dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, lateral_sides=sides, prof=prof,
                                          zt=lv.ztop)
        for n in ACOUSTIC_OUTPUTS:
            x, y = a[n].astype(np.float64), b[n].astype(np.float64)
            worst[n] = max(worst[n], float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-30)))
    return worst


@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_flat_acoustic_stage_equals_the_terrain_coordinate_stage_on_a_flat_grid(sides, uniform):
    """The flat stage (cuda_gm1_child_psolver3_open.py:340-535) against the golden-pinned terrain stage.

    Same equations, different association: within the 1e-5 relative max-norm of
    the task's tolerance for every output of three stages (measured <= 1.5e-6),
    on a uniform and on a stretch_z=1 column.
    """
    worst = _flat_and_terrain_stages(sides, uniform)
    assert max(worst.values()) < 1e-5, worst


def test_flat_acoustic_stage_comparison_sees_exchanged_column_factors():
    """mh and mf exchanged, or c1 and c2, move the stretched flat stage far from the terrain stage."""
    from dataclasses import replace

    lv = host_initial_state(stretched_case(16, 12, 16, (2, 2, 2, 2)))["levels"]
    swapped = replace(lv, mh=lv.mf[: lv.nk + 2].copy(), mf=np.concatenate([lv.mh, lv.mh[-1:]]).astype(F))
    worst = _flat_and_terrain_stages((2, 2, 2, 2), False, flat_levels=swapped)
    assert max(worst.values()) > 1e-3, worst
    host = host_initial_state(stretched_case(16, 12, 16, (2, 2, 2, 2)))
    m = dict(host["metrics"])
    m["c1"], m["c2"] = m["c2"], m["c1"]
    worst = _flat_and_terrain_stages((2, 2, 2, 2), False, flat_metrics=m)
    assert max(worst.values()) > 1e-3, worst


def test_flat_tridiagonal_rows_solve_like_a_dense_binary64_solve():
    """On the stretched column the flat stage's Thomas sweep gives a dense binary64 solve of its rows."""
    host, a = evolved(sides=(2, 2, 2, 2))
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    ost.stage_terms(a, host["base"], a["th3d"], ol._species6(a["q3d"]), ni=ni, nj=nj, nk=nk, sides=cfg.lateral_sides)
    nloop, dts = osound.sound3_stage_schedule_f32(cfg.dt, cfg.nsound, 3)
    dts = F(dts)
    osound.flat_prepare_coefficients(a, m, lv, ni=ni, nj=nj, nk=nk, dts=dts)
    a["ppd"][1 : nk + 1] = a["ppx"][1 : nk + 1]
    rec = {}
    rdx, rdy = F(1.0) / F(cfg.dx), F(1.0) / F(cfg.dy)
    osound.flat_small_step(a, m, lv, cfg.lateral_sides, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy, nrk=3, n=1,
                           nloop=nloop, alph=F(0.6), beta=F(0.4), kdiv=F(0.1), tridiagonal=rec)
    rows = {k: (aa, bb, cc, dd) for k, aa, bb, cc, dd in rec["rows"]}
    spacing = lv.spacings
    assert spacing.max() / spacing.min() > 5  # a genuinely stretched column
    n = nk - 1
    worst = 0.0
    for j in range(nj):
        for i in range(ni):
            A = np.zeros((n, n))
            rhs = np.zeros(n)
            for r, k in enumerate(range(2, nk + 1)):
                aa, bb, cc, dd = (np.asarray(x, np.float64)[j, i] for x in rows[k])
                A[r, r] = bb
                if r > 0:
                    A[r, r - 1] = aa
                if r < n - 1:
                    A[r, r + 1] = cc
                rhs[r] = dd
            exact = np.linalg.solve(A, rhs)
            got = a["w3d"][2 : nk + 1, j + NG, i + NG].astype(np.float64)
            worst = max(worst, float(np.max(np.abs(got - exact)) / (np.max(np.abs(exact)) + 1e-30)))
            assert np.all(np.diag(A) > 0.9)
    assert worst < 1e-5


# ----------------------------------------------------------------------------- GPU


def _device_state(host, a, device):
    import torch

    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in a.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_gpu_flat_momentum_matches_oracle(device, sides):
    import torch

    from gm1_amd.stretched import flat_stretched_column, launch_cuda_gm1_child_stretched_momentum5

    host, a = evolved(sides=sides)
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    print(f"flat momentum {ni}x{nj}x{nk} sides {sides}", flush=True)
    d = _device_state(host, a, device)
    dm = _device_state(host, m, device)
    col = flat_stretched_column(lv, device)
    launch_cuda_gm1_child_stretched_momentum5(d, dm, col, ni=ni, nj=nj, nk=nk, nrk=2, dx=cfg.dx, dy=cfg.dy,
                                              divx=d["divx"], lateral_sides=sides)
    torch.cuda.synchronize()
    omom.cpu_terrain_open_momentum_advection5_inplace(a, m, a["divx"], ni=ni, nj=nj, nk=nk, nrk=2, dx=cfg.dx,
                                                      dy=cfg.dy, dz=cfg.dz, lateral_sides=sides, flat=lv)
    for n in ("rru", "rrv", "rrw", "divx", "uten", "vten", "wten"):
        assert_bits_equal(to_host(d[n]), a[n], n)


@pytest.mark.gpu
def test_gpu_flat_buoyancy_post_acoustic_and_acoustic_stage_match_oracle(device):
    import torch

    from gm1_amd.stretched import (flat_stretched_column, launch_cuda_gm1_child_stretched_post_acoustic,
                                   launch_cuda_gm1_flat_acoustic_stage, launch_cuda_gm1_flat_stretched_buoyancy)

    host, a = evolved()
    cfg, m, lv = host["cfg"], host["metrics"], host["levels"]
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    sides = cfg.lateral_sides
    ost.stage_terms(a, host["base"], a["th3d"], ol._species6(a["q3d"]), ni=ni, nj=nj, nk=nk, sides=sides)
    print(f"flat buoyancy/acoustic/post-acoustic {ni}x{nj}x{nk}", flush=True)
    d = _device_state(host, a, device)
    dm = _device_state(host, m, device)
    col = flat_stretched_column(lv, device)
    launch_cuda_gm1_flat_stretched_buoyancy(d, dm, ni=ni, nj=nj, nk=nk)
    ost.buoyancy_flat(a, m, ni=ni, nj=nj, nk=nk)
    torch.cuda.synchronize()
    assert_bits_equal(to_host(d["wten"]), a["wten"], "wten after buoyancy")
    for nrk in (1, 2, 3):
        launch_cuda_gm1_flat_acoustic_stage(d, dm, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=cfg.dt, nsound=cfg.nsound,
                                            dx=cfg.dx, dy=cfg.dy, vertical=col, lateral_sides=sides)
        osound.cpu_flat_stage_inplace(a, m, lv, ni=ni, nj=nj, nk=nk, dt=cfg.dt, nsound=cfg.nsound, nrk=nrk,
                                      dx=cfg.dx, dy=cfg.dy, lateral_sides=sides)
        torch.cuda.synchronize()
        for n in ("u3d", "v3d", "w3d", "pp3d", "ppd", "ppx", "rru", "rrv", "rrw", "radbcw", "radbcs"):
            assert_bits_equal(to_host(d[n]), a[n], f"stage {nrk} {n}")
    launch_cuda_gm1_child_stretched_post_acoustic(d, dm, col, ni=ni, nj=nj, nk=nk, dx=cfg.dx, dy=cfg.dy,
                                                  thterm=d["thterm"], divergence=d["divergence"],
                                                  theta_divergence_term=d["theta_div"], divx=d["divx"])
    ost.getdiv_flat(a, a["divergence"], a["theta_div"], a["thterm"], lv, ni=ni, nj=nj, nk=nk, dx=cfg.dx, dy=cfg.dy)
    ost.getdivx_flat(a, m, a["divx"], lv, ni=ni, nj=nj, nk=nk, dx=cfg.dx, dy=cfg.dy)
    torch.cuda.synchronize()
    for n in ("divergence", "theta_div", "rru", "rrv", "rrw", "divx"):
        assert_bits_equal(to_host(d[n]), a[n], n)


@pytest.mark.gpu
@pytest.mark.parametrize("sides,column_pass,ni", [((2, 2, 2, 2), "staged", 24), ((1, 1, 1, 1), "staged", 24),
                                                  ((2, 2, 2, 2), "pipelined", 24), ((2, 2, 2, 2), "pipelined16", 24),
                                                  ((1, 1, 1, 1), "staged16", 24), ((2, 2, 2, 2), "fused", 70),
                                                  ((1, 1, 1, 1), "fused", 70), ((2, 2, 2, 2), "fused16", 24),
                                                  ((2, 2, 2, 2), "resident", 70), ((1, 1, 1, 1), "resident16", 24)])
def test_gpu_stretched_moist_large_step_bit_exact(device, sides, column_pass, ni):
    """Two moist large steps on a stretched column against the oracle (device dry-mass correction injected).

    At ni = 70 the 32-column fused pass has two full tiles and a ragged one, so
    a tile's last lane and the lane west of the east column both form their own
    east u face; the 16-column one does at ni = 24.
    """
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step

    cfg = stretched_case(ni, 20, 16, sides)
    host = host_initial_state(cfg)
    st = create_gm1_large_step_state(host, device)
    st.column_pass = column_pass
    assert st.flat is not None
    a = {k: v.copy() for k, v in host["arrays"].items()}
    lv = host["levels"]
    print(f"stretched moist large step {cfg.ni}x{cfg.nj}x{cfg.nk} sides {sides} {column_pass}", flush=True)
    for step in range(2):
        launch_cuda_gm1_large_step(st)
        torch.cuda.synchronize()
        corr = float(st.dry_mass.correction.cpu().numpy()[0])
        ol.large_step(a, host["metrics"], host["base"], host["prof"], cfg_dict(cfg), zt=F(lv.ztop),
                      mass_correction=corr, flat=lv)
        for name in MOIST_FIELDS:
            assert_bits_equal(to_host(st.a[name]), a[name], f"step {step + 1} {name}")
        np.testing.assert_allclose(to_host(st.a["qbudget"]), a["qbudget"], rtol=1e-11, atol=1e-25)
    assert np.abs(a["w3d"]).max() > 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_gpu_fused_column_pass_equals_the_two_pass_stage(device, sides):
    """The fused and resident passes against the pipelined stage, word for word, on the 80-level stretched
    column at 130x90 (about 5x91 and 9x91 workgroups a pass): every output of the three RK stages (3, 5 and
    10 small steps: both parities of the alternation), including ppd's halo ring and the radiation phase
    speeds."""
    from dataclasses import replace

    import torch

    from gm1_amd.fields import large_stretched_c3
    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step
    from gm1_amd.stretched import launch_cuda_gm1_flat_acoustic_stage

    cfg = replace(large_stretched_c3(130, 90, 80), lateral_sides=tuple(sides))
    print(f"fused column pass {cfg.ni}x{cfg.nj}x{cfg.nk} sides {sides}", flush=True)
    st = create_gm1_large_step_state(host_initial_state(cfg), device)
    launch_cuda_gm1_large_step(st)  # a stirred state: winds, pressure and tendencies all nonzero
    torch.cuda.synchronize()
    names = ("u3d", "v3d", "w3d", "pp3d", "ppd", "ppx", "ppi", "rru", "rrv", "rrw", "radbcw", "radbce", "radbcs",
             "radbcn")
    runs = {}
    for column_pass in ("pipelined", "fused", "fused16", "resident", "resident16"):
        a = {k: v.clone() for k, v in st.a.items()}
        alt = {k: torch.full_like(st.alt[k], float("nan")) for k in ("u3d", "v3d", "ppd")}
        out = []
        for nrk in (1, 2, 3):
            launch_cuda_gm1_flat_acoustic_stage(a, st.metrics, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, nrk=nrk, dt=cfg.dt,
                                                nsound=cfg.nsound, dx=cfg.dx, dy=cfg.dy, vertical=st.flat,
                                                lateral_sides=sides, column_pass=column_pass, alternates=alt)
            torch.cuda.synchronize()
            out.append({n: to_host(a[n]) for n in names})
        runs[column_pass] = out
    for column_pass in ("fused", "fused16", "resident", "resident16"):
        for nrk in (1, 2, 3):
            for n in names:
                assert_bits_equal(runs[column_pass][nrk - 1][n], runs["pipelined"][nrk - 1][n],
                                  f"{column_pass} stage {nrk} {n}")
    assert np.abs(runs["pipelined"][2]["w3d"]).max() > 0.0


def test_fused_column_pass_flags():
    """Only the fused forms carry the fused-pass flag (the flag that makes the stage take alternates)."""
    from gm1_amd.psolver3 import FLAT_COLUMN_PASSES, FUSED_PASS_FLAG

    assert FLAT_COLUMN_PASSES["fused"] & FUSED_PASS_FLAG and FLAT_COLUMN_PASSES["fused16"] & FUSED_PASS_FLAG
    assert not any(v & FUSED_PASS_FLAG for k, v in FLAT_COLUMN_PASSES.items() if not k.startswith("fused"))
