"""The reference's stateful entry points, mirrored: stage bridge, scalar and momentum RK stages, Kessler transition.

Each mirror keeps the reference's names, keywords, state lifecycle and
cadence guards (cuda_moist_stage_bridge.py:303-505, cuda_scalar_moist_rk.py:
823-1499, cuda_momentum_rk.py:572-850, cuda_gm1_kessler_large_step.py:
295-640, cuda_gm1_forecast.py:1569-1627).  GPU tests drive the mirrors in
the reference composer's order and compare with the oracle bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, assert_bits_equal, to_host
from gm1_amd.errors import CM1Refusal
from gm1_amd.fields import CaseConfig
from gm1_amd.large_step import host_initial_state
from oracle import boundary as obc
from oracle import large_step as ol
from oracle import sound as osound
from oracle import stage as ost
from oracle.momentum5 import cpu_terrain_open_momentum_advection5_inplace


def dry_case(ni, nj, nk, sides):
    return CaseConfig(f"dry-{ni}x{nj}x{nk}", ni, nj, nk, 1000.0, 1000.0, 500.0, 6.0, 6, tuple(sides), False,
                      "linear15", ((2.0, ni * 450.0, nj * 520.0, 1400.0),), bubble_radius=(ni * 300.0, 1400.0),
                      noise=0.05, seed=3)


def moist_case(ni, nj, nk, sides):
    return CaseConfig(f"moist-{ni}x{nj}x{nk}", ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 8, tuple(sides), True,
                      "hodograph", ((3.0, ni * 120.0, nj * 130.0, 1000.0),), bubble_radius=(ni * 80.0, 1000.0),
                      noise=0.1, seed=5, extra={"bubble_rh": 1.03})


def cfgd(cfg):
    return {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
            "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}


# ----------------------------------------------------------------------------- CPU: the boundary mirror's checks
def test_boundary_mirror_shapes_and_refusals():
    """The family extents (cuda_boundary_families.py:844-853 for ngxy = 3) and the grids, kinds and side words
    the fills refuse before anything is enqueued."""
    from gm1_amd import boundary

    assert boundary.family_array_shape(0, ni=10, nj=8, nk=5) == (7, 14, 16)
    assert boundary.family_array_shape(1, ni=10, nj=8, nk=5) == (7, 14, 17)
    assert boundary.family_array_shape(2, ni=10, nj=8, nk=5) == (7, 15, 16)
    assert boundary.family_array_shape(3, ni=10, nj=8, nk=5) == (8, 14, 16)
    assert boundary.pressure_array_shape(ni=10, nj=8, nk=5) == (7, 10, 12)
    for bad in ({"ni": 3, "nj": 8, "nk": 5}, {"ni": 10, "nj": 8, "nk": 2}, {"ni": True, "nj": 8, "nk": 5}):
        with pytest.raises(ValueError):
            boundary.family_array_shape(0, **bad)
    with pytest.raises(ValueError):
        boundary.family_array_shape(4, ni=10, nj=8, nk=5)
    with pytest.raises(ValueError):  # a side word outside 1..4
        boundary.fill_pressure(None, (2, 2, 5, 2), ni=10, nj=8, nk=5)
    with pytest.raises(ValueError):  # the tile-local word has no family to fill
        boundary.fill_family(None, 0, None, ni=10, nj=8, nk=5)
    with pytest.raises(TypeError):  # not a device tensor: refused before any launch
        boundary.fill_family(np.zeros((7, 14, 16), np.float32), 0, (1, 1, 2, 3), ni=10, nj=8, nk=5)


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(1, 1, 1, 1), (2, 2, 2, 2), (3, 4, 2, 2), (2, 2, 4, 3)])
def test_gpu_boundary_mirror_fills_match_the_oracle(device, sides):
    """Every fill through the mirror on random carriers equals the oracle's fill word for word (mixed families
    included: periodic, open-radiative, rigid free-slip and no-slip on different sides)."""
    import torch

    from gm1_amd import boundary

    ni, nj, nk = 13, 10, 6
    rng = np.random.default_rng(11)
    print("boundary mirror fills", sides, flush=True)
    for kind in (0, 1, 2, 3):
        h = rng.standard_normal(boundary.family_array_shape(kind, ni=ni, nj=nj, nk=nk)).astype(np.float32)
        d = torch.from_numpy(h.copy()).to(device)
        boundary.fill_family(d, kind, sides, ni=ni, nj=nj, nk=nk)
        obc.fill(h, kind, sides, ni=ni, nj=nj, nk=nk)
        torch.cuda.synchronize()
        assert_bits_equal(to_host(d), h, f"kind {kind}")
    p = rng.standard_normal(boundary.pressure_array_shape(ni=ni, nj=nj, nk=nk)).astype(np.float32)
    dp = torch.from_numpy(p.copy()).to(device)
    boundary.fill_pressure(dp, sides, ni=ni, nj=nj, nk=nk)
    obc.bcp(p, sides, ni=ni, nj=nj, nk=nk)
    w = rng.standard_normal(boundary.family_array_shape(3, ni=ni, nj=nj, nk=nk)).astype(np.float32)
    dw = torch.from_numpy(w.copy()).to(device)
    boundary.w_rigid(dw, ni=ni, nj=nj, nk=nk)
    obc.bcw_rigid(w, ni=ni, nj=nj, nk=nk)
    ha = {n: rng.standard_normal(boundary.family_array_shape(kd, ni=ni, nj=nj, nk=nk)).astype(np.float32)
          for n, kd in (("u3d", 1), ("v3d", 2))}
    ha["w3d"] = w
    hm = {n: rng.standard_normal((nj, ni)).astype(np.float32) for n in ("dzdx", "dzdy", "gz")}
    da = {n: torch.from_numpy(x.copy()).to(device) for n, x in ha.items() if n != "w3d"}
    da["w3d"] = dw
    dm = {n: torch.from_numpy(x).to(device) for n, x in hm.items()}
    boundary.bcwsfc(da, dm, ni=ni, nj=nj)
    obc.bcwsfc(ha, hm, ni=ni, nj=nj, nk=nk)
    boundary.bc2d_open(dw, ni=ni, nj=nj)
    obc.bc2d_open_w_surface(w, ni=ni, nj=nj)
    torch.cuda.synchronize()
    assert_bits_equal(to_host(dp), p, "pressure")
    assert_bits_equal(to_host(dw), w, "w rigid + bcwsfc + bc2d")
    su, sw = tuple(da["u3d"].shape), tuple(dw.shape)
    shared = dw.new_zeros(max(int(np.prod(su)), int(np.prod(sw))))
    with pytest.raises(ValueError):  # w aliasing u is refused before the launch
        boundary.bcwsfc({"u3d": shared[: int(np.prod(su))].view(su), "v3d": da["v3d"],
                         "w3d": shared[: int(np.prod(sw))].view(sw)}, dm, ni=ni, nj=nj)


# ----------------------------------------------------------------------------- CPU: refusals before any launch


def test_large_step_refuses_lateral_forcing_and_reflectivity_before_any_launch():
    from gm1_amd.large_step import launch_cuda_gm1_large_step

    with pytest.raises(CM1Refusal, match="GM1AMD-LATERAL-FORCING"):
        launch_cuda_gm1_large_step(None, lateral_forcing=object())
    with pytest.raises(CM1Refusal, match="REFLECTIVITY"):
        launch_cuda_gm1_large_step(None, reflectivity_due=True)
    with pytest.raises(TypeError):
        launch_cuda_gm1_large_step(None, solve1_hook=3)


def test_bridge_refuses_a_3d_base_state():
    from gm1_amd.moist_stage_bridge import create_cuda_moist_stage_bridge_state

    with pytest.raises(CM1Refusal, match="BASE3D"):
        create_cuda_moist_stage_bridge_state(ni=10, nj=10, nk=10, base3d=True)
    with pytest.raises(TypeError):
        create_cuda_moist_stage_bridge_state(ni=10, nj=10, nk=10, base3d=1)


# ----------------------------------------------------------------------------- GPU


def _dev(host, device):
    import torch

    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in host.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_gpu_momentum_rk_state_three_stages(device, sides):
    """create_/launch_/begin_cuda_momentum_rk_stage against cpu_momentum_rk_stage_inplace's order, bit for bit."""
    import torch

    from gm1_amd.momentum_rk import (begin_cuda_momentum_rk_step, create_cuda_momentum_rk_state,
                                     launch_cuda_momentum_rk_stage)

    cfg = dry_case(24, 20, 12, sides)
    host = host_initial_state(cfg)
    ha = {k: v.copy() for k, v in host["arrays"].items()}
    ha.setdefault("scalar", np.zeros((cfg.nk + 2, cfg.nj + 6, cfg.ni + 6), F))
    rng = np.random.default_rng(4)
    for n in ("uten1", "vten1", "wten1"):
        ha[n][...] = (1e-3 * rng.standard_normal(ha[n].shape)).astype(F)
    m = host["metrics"]
    # the stage terms the acoustic stage reads (thv, ppterm, piadv), held fixed over the three stages
    ost.stage_terms(ha, host["base"], ha["th3d"], None, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, sides=sides)
    print(f"momentum RK mirror {cfg.name} sides={sides}", flush=True)
    da, dm = _dev(ha, device), _dev(m, device)
    st = create_cuda_momentum_rk_state(da, dm, {n: da[n] for n in ("uten1", "vten1", "wten1")}, ni=cfg.ni,
                                       nj=cfg.nj, nk=cfg.nk, stream=torch.cuda.current_stream(), lateral_sides=sides)
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    divx = np.zeros((cfg.nk, cfg.nj, cfg.ni), F)
    for nrk in (1, 2, 3):
        launch_cuda_momentum_rk_stage(st, nrk=nrk, dt=cfg.dt, nsound=cfg.nsound, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz)
        ost.seed_held((ha["uten1"], ha["vten1"], ha["wten1"]), (ha["uten"], ha["vten"], ha["wten"]), ni=cfg.ni,
                      nj=cfg.nj, nk=cfg.nk)
        if nrk >= 2:
            obc.bcwsfc(ha, m, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk)
            obc.bc2d_open_w_surface(ha["w3d"], ni=cfg.ni, nj=cfg.nj)
        cpu_terrain_open_momentum_advection5_inplace(ha, m, divx, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, nrk=nrk,
                                                     dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, lateral_sides=sides)
        osound.cpu_psolver3_stage_inplace(ha, m, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, dt=cfg.dt, nsound=cfg.nsound,
                                          nrk=nrk, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, lateral_sides=sides, prof=prof)
        torch.cuda.synchronize()
        for name in ("u3d", "v3d", "w3d", "pp3d", "ppx", "rru", "rrv", "rrw", "uten", "wten"):
            assert_bits_equal(to_host(da[name]), ha[name], f"stage {nrk} {name}")
        assert_bits_equal(to_host(st.divx), divx, f"stage {nrk} divx")
    with pytest.raises(ValueError, match="cadence"):
        launch_cuda_momentum_rk_stage(st, nrk=1, dt=cfg.dt, nsound=cfg.nsound, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz)
    begin_cuda_momentum_rk_step(st)
    with pytest.raises(ValueError, match="stage 3"):
        begin_cuda_momentum_rk_step(st)


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_gpu_bridge_and_scalar_rk_mirrors_compose_three_moist_stages(device, sides):
    """The moist bridge, the momentum advection, the acoustic stage and the scalar RK state driven in the
    CartesianDynamics order (cuda_cartesian_dynamics.py:1047-1262) equal the oracle's RK stages bit for bit."""
    import torch

    from gm1_amd import _native
    from gm1_amd._device import ptr, stream_handle
    from gm1_amd.moist_stage_bridge import (begin_cuda_moist_stage, create_cuda_moist_stage_bridge_state,
                                            launch_cuda_eqtset2_stage_terms, launch_cuda_moist_buoyancy,
                                            launch_cuda_post_acoustic_flux_bridge)
    from gm1_amd.momentum_advection5 import launch_cuda_terrain_open_momentum_advection5
    from gm1_amd.psolver3 import launch_cuda_psolver3_terrain_open_stage
    from gm1_amd.scalar_moist_rk import (begin_cuda_scalar_moist_rk_step, create_cuda_scalar_moist_rk_state,
                                         launch_cuda_scalar_moist_rk_stage)

    cfg = moist_case(24, 20, 16, sides)
    host = host_initial_state(cfg)
    ha = {k: v.copy() for k, v in host["arrays"].items()}
    m, base = host["metrics"], host["base"]
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    print(f"bridge + scalar RK mirrors {cfg.name} sides={sides}", flush=True)
    da, dm, db = _dev(ha, device), _dev(m, device), _dev(base, device)
    stream = torch.cuda.current_stream()
    bridge = create_cuda_moist_stage_bridge_state(ni=ni, nj=nj, nk=nk, base3d=False, stream=stream, device=device,
                                                  lateral_sides=sides)
    srk = create_cuda_scalar_moist_rk_state(da, da, db["th0"], ni=ni, nj=nj, nk=nk, stream=stream,
                                            qmag=(1.0e-2,) * 3, names=("qv", "qc", "qr"), lateral_sides=sides)
    for nrk in (1, 2, 3):
        if nrk > 1:
            begin_cuda_moist_stage(bridge)
        launch_cuda_eqtset2_stage_terms(bridge, da, dm, th0=db["th0"], qv0=db["qv0"], pi0=db["pi0"], th3d=da["th3d"],
                                        q3d=da["q3d"])
        h = stream_handle(stream)
        _native.call("gm1hip_seed_held_momentum", ptr(da["uten1"]), ptr(da["vten1"]), ptr(da["wten1"]),
                     ptr(da["uten"]), ptr(da["vten"]), ptr(da["wten"]), ni, nj, nk, h)
        if nrk >= 2:
            _native.call("gm1hip_bcwsfc", ptr(da["u3d"]), ptr(da["v3d"]), ptr(dm["dzdx"]), ptr(dm["dzdy"]),
                         ptr(dm["gz"]), ptr(da["w3d"]), ni, nj, h)
            _native.call("gm1hip_bc2d_open", ptr(da["w3d"]), ni, nj, h)
        launch_cuda_terrain_open_momentum_advection5(da, dm, ni=ni, nj=nj, nk=nk, nrk=nrk, dx=cfg.dx, dy=cfg.dy,
                                                     dz=cfg.dz, divx=bridge.divx, lateral_sides=sides)
        launch_cuda_moist_buoyancy(bridge, da, dm)
        launch_cuda_psolver3_terrain_open_stage(da, dm, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=cfg.dt, nsound=cfg.nsound,
                                                dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, lateral_sides=sides)
        launch_cuda_post_acoustic_flux_bridge(bridge, da, dm, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz)
        launch_cuda_scalar_moist_rk_stage(srk, {"rru": da["rru"], "rrv": da["rrv"], "rrw": da["rrw"]}, dm,
                                          bridge.divx, bridge.theta_divergence_term, nrk=nrk, dt=cfg.dt, dx=cfg.dx,
                                          dy=cfg.dy, dz=cfg.dz)
        ol.rk_stage(ha, m, base, prof, cfgd(cfg), nrk, zt=F(nk * cfg.dz))
        torch.cuda.synchronize()
        for name in ("u3d", "v3d", "w3d", "pp3d", "th3d", "q3d", "rru", "rrv", "rrw"):
            assert_bits_equal(to_host(da[name]), ha[name], f"stage {nrk} {name}")
        assert_bits_equal(to_host(bridge.divx), ha["divx"], f"stage {nrk} divx")
        assert_bits_equal(to_host(bridge.theta_divergence_term), ha["theta_div"], f"stage {nrk} theta_div")
    with pytest.raises(ValueError, match="begin the next moist stage"):
        launch_cuda_eqtset2_stage_terms(bridge, da, dm, th0=db["th0"], qv0=db["qv0"], pi0=db["pi0"],
                                        th3d=da["th3d"], q3d=da["q3d"])
    with pytest.raises(ValueError, match="cadence requires stage 4"):
        launch_cuda_scalar_moist_rk_stage(srk, {"rru": da["rru"], "rrv": da["rrv"], "rrw": da["rrw"]}, dm,
                                          bridge.divx, bridge.theta_divergence_term, nrk=1, dt=cfg.dt, dx=cfg.dx,
                                          dy=cfg.dy, dz=cfg.dz)
    begin_cuda_scalar_moist_rk_step(srk)


@pytest.mark.gpu
def test_gpu_kessler_transition_object_matches_the_transition_launcher(device):
    """CudaGm1KesslerLargeStepTransition.apply with the dry-mass hook equals launch_cuda_gm1_kessler_transition
    over two large steps; the held rates and persistent accumulators are its restart inventory."""
    import torch

    from gm1_amd.kessler import (Gm1KesslerDriverSelection, create_cuda_gm1_kessler_driver_workspace,
                                 create_cuda_gm1_kessler_large_step_transition, create_dry_mass_workspace,
                                 launch_cuda_apmasscon, launch_cuda_gm1_kessler_transition)
    from helpers import golden

    g = golden("kessler")
    ni, nj, nk = (int(v) for v in g["dims"])
    a = {k[len("t_in_"):]: g[k].copy() for k in g.files if k.startswith("t_in_")}
    held0 = {k: a.pop(k) for k in ("qpten", "qtten", "qvten", "qcten")}
    from oracle import kessler as ok

    mass1 = ok.dry_mass(a["rho"], a["ruh"], a["rvh"], a["rmh"]) * (1.0 + 3e-6)
    sel = Gm1KesslerDriverSelection(equation_set=2)
    ref, got = _dev(a, device), _dev(a, device)
    ref_held = _dev(held0, device)
    tr = create_cuda_gm1_kessler_large_step_transition(got, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0,
                                                       dt=1.5, selection=sel, stream=torch.cuda.current_stream())
    for c, src in zip((tr.carriers.qpten, tr.carriers.qtten, tr.carriers.qvten, tr.carriers.qcten),
                      (ref_held["qpten"], ref_held["qtten"], ref_held["qvten"], ref_held["qcten"])):
        c.copy_(src)
    ws = create_cuda_gm1_kessler_driver_workspace(ni, nj, nk, device)
    dm_ref, dm_got = create_dry_mass_workspace(ni, nj, nk, device), create_dry_mass_workspace(ni, nj, nk, device)

    def hook(carriers):
        launch_cuda_apmasscon(got, ni=ni, nj=nj, nk=nk, mass1=mass1, workspace=dm_got)
        return "apmasscon"

    for _ in range(2):
        launch_cuda_gm1_kessler_transition(ref, ref_held, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0, dt=1.5,
                                           selection=sel, workspace=ws, mass1=mass1, dry_mass=dm_ref)
        res = tr.apply(pre_component_hook=hook)
    torch.cuda.synchronize()
    for name in ("rho", "rr", "pp3d", "th3d", "prs", "qv", "qc", "qr", "rain", "prate"):
        assert_bits_equal(to_host(got[name]), to_host(ref[name]), name)
    for name, c in (("qpten", tr.carriers.qpten), ("qtten", tr.carriers.qtten), ("qvten", tr.carriers.qvten),
                    ("qcten", tr.carriers.qcten)):
        assert_bits_equal(to_host(c), to_host(ref_held[name]), name)
    assert res.completed_transitions == 2 and list(res.restart.arrays) == [
        "held.qpten", "held.qtten", "held.qvten", "held.qcten", "persistent.rain", "persistent.prate",
        "persistent.asq", "persistent.qbudget"]
    assert tr.receipt_dict()["source_order"][2] == "dry-mass correction hook"
    with pytest.raises(CM1Refusal, match="CADENCE"):
        tr.apply(final_rk=False)


@pytest.mark.gpu
def test_gpu_large_step_hooks(device):
    """post_stage_hook runs after stages 1, 2, 3; a solve1_hook that adds nothing leaves the step word for word."""
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step

    cfg = moist_case(24, 20, 16, (2, 2, 2, 2))
    host = host_initial_state(cfg)
    plain = create_gm1_large_step_state(host, device)
    hooked = create_gm1_large_step_state(host, device)
    seen = []
    launch_cuda_gm1_large_step(plain)
    r = launch_cuda_gm1_large_step(hooked, solve1_hook=lambda st, c: c.thten1.add_(0.0),
                                   post_stage_hook=lambda st, nrk: seen.append(nrk))
    torch.cuda.synchronize()
    assert seen == [1, 2, 3] and r["solve1_hook"] and r["post_stage_hook"]
    for name in ("u3d", "v3d", "w3d", "th3d", "q3d", "pp3d", "rho"):
        assert_bits_equal(to_host(hooked.a[name]), to_host(plain.a[name]), name)
    # a held theta tendency added by the hook reaches theta
    r = launch_cuda_gm1_large_step(hooked, solve1_hook=lambda st, c: c.thten1.add_(1e-3))
    launch_cuda_gm1_large_step(plain)
    torch.cuda.synchronize()
    assert not torch.equal(hooked.a["th3d"], plain.a["th3d"])
