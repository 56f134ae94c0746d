"""The composed dry large step: oracle invariants on the CPU, bit parity on the GPU.

The composition (cuda_gm1_forecast.py:1241-1627) has no CPU form in the
reference; each operator is pinned in its own test file, and the
composition is pinned here by a state at rest staying exactly at rest,
periodic translation equivariance, and the stage operators' closed forms.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, NG, assert_bits_equal, seeded_metrics, shapes, to_dev, to_host
from gm1_amd.fields import CaseConfig
from gm1_amd.large_step import host_initial_state
from oracle import large_step as ol
from oracle import sound as osound
from oracle import stage as ost

PROGNOSTIC = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ppd")


def dry_case(ni, nj, nk, sides, *, amp=2.0, noise=0.05, wind="linear15", seed=3):
    return CaseConfig(f"dry-{ni}x{nj}x{nk}", ni, nj, nk, 1000.0, 1000.0, 500.0, 6.0, 6, tuple(sides), False, wind,
                      ((amp, ni * 450.0, nj * 520.0, 1400.0),), bubble_radius=(ni * 300.0, 1400.0), noise=noise,
                      seed=seed)


def oracle_run(host, nsteps):
    cfg = host["cfg"]
    a = {k: v.copy() for k, v in host["arrays"].items()}
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    c = {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}
    out = []
    for _ in range(nsteps):
        ol.large_step(a, host["metrics"], host["base"], prof, c, zt=F(cfg.nk * cfg.dz))
        out.append({k: a[k].copy() for k in PROGNOSTIC})
    return out


def test_state_at_rest_stays_exactly_at_rest():
    cfg = dry_case(12, 10, 10, (1, 1, 1, 1), amp=0.0, noise=0.0)
    host = host_initial_state(cfg)
    for name in ("u3d", "ua", "v3d", "va"):
        host["arrays"][name][...] = F(0.0)
    (s,) = oracle_run(host, 1)
    for name in PROGNOSTIC:
        assert not np.any(s[name]), name


def test_periodic_large_step_is_translation_equivariant():
    ni, nj, nk = 16, 12, 10
    cfg = dry_case(ni, nj, nk, (1, 1, 1, 1))
    h0 = host_initial_state(cfg)
    h1 = host_initial_state(cfg)
    di, dj = 5, 3
    for name in ("th3d", "tha"):
        h1["arrays"][name][...] = np.roll(h0["arrays"][name], (dj, di), axis=(1, 2))
    (s0,) = oracle_run(h0, 1)
    (s1,) = oracle_run(h1, 1)
    interior = {"u3d": (NG, NG + nj, NG, NG + ni), "v3d": (NG, NG + nj, NG, NG + ni), "w3d": (NG, NG + nj, NG, NG + ni),
                "pp3d": (1, nj + 1, 1, ni + 1)}
    for name, (y0, y1, x0, x1) in interior.items():
        want = np.roll(s0[name][:, y0:y1, x0:x1], (dj, di), axis=(1, 2))
        assert_bits_equal(s1[name][:, y0:y1, x0:x1], want, name)
    assert_bits_equal(s1["th3d"], np.roll(s0["th3d"], (dj, di), axis=(1, 2)), "th3d")


def test_bubble_rises_and_fields_stay_bounded():
    cfg = dry_case(16, 12, 12, (1, 1, 1, 1), noise=0.0)
    host = host_initial_state(cfg)
    s = oracle_run(host, 2)[-1]
    w = s["w3d"][2 : cfg.nk + 1, NG : NG + cfg.nj, NG : NG + cfg.ni]
    assert np.isfinite(w).all() and w.max() > 1e-3 and np.abs(w).max() < 5.0
    assert np.abs(s["pp3d"]).max() < 1e-3


def test_compose_clamps_only_after_the_update():
    """A-anchored compose (cuda_scalar_moist_rk.py:336-341); tests/test_cuda_scalar_moist_rk.py:119-183 shape."""
    tha = np.array([1.0, 1e-31, 0.0, 2.0], F)
    held = np.array([0.0, 0.0, 1e-31, -1.0], F)
    div = np.array([0.0, 0.0, 0.0, 0.0], F)
    adv = np.array([0.0, 0.0, 0.0, 0.5], F)
    out = ost.compose_theta(tha, held, div, adv, F(2.0))
    np.testing.assert_array_equal(out, np.array([1.0, 0.0, 0.0, 1.0], F))


def test_density_finish_closed_forms():
    rho = np.linspace(1.2, 0.4, 8, dtype=F)[:, None, None] * np.ones((8, 2, 3), F)
    c2 = np.full((9, 2, 3), F(0.5))
    c1 = (F(1.0) - c2).astype(F)
    rr, rf = ost.solve3_density(rho, c1, c2)
    np.testing.assert_array_equal(rr, F(1.0) / rho)
    # three-point extrapolation is exact for a linear column
    np.testing.assert_allclose(rf[0], 1.2 + 0.5 * (1.2 - (1.2 - 0.8 / 7)), rtol=2e-6)
    np.testing.assert_allclose(rf[1:8], 0.5 * (rho[:-1] + rho[1:]), rtol=1e-7)


def _reference_schedule(moist, terrain, numq=3):
    """The operator order of one large step as the reference's composer runs it.

    Phases and their order: ``gm1_runtime.large_step_schedule`` (gm1_runtime.py:1969-2007);
    inside them ``CartesianDynamics.finalize_held_tendencies`` (cuda_cartesian_dynamics.py:1014),
    ``pre_acoustic`` (:1047-1110: stage terms, held momentum, bcwsfc + bc2d from nrk 2 on a
    terrain column, momentum advection, buoyancy), ``acoustic`` (:1127), ``post_acoustic``
    (:1204: the flux bridge's getdiv then getdivx, cuda_moist_stage_bridge.py:456-520),
    ``scalar_moist`` (:1235: theta, then each moisture row; the final stage's rows through the
    positive-definite chain, cuda_scalar_moist_rk.py:1041-1499), then Kessler's large-step
    transition and solve3 (cuda_gm1_forecast.py:1414-1457).
    """
    out = []
    if moist:
        out.append(("held", None))
    for nrk in (1, 2, 3):
        out += [("stage_terms", nrk), ("seed_held", nrk)]
        if nrk >= 2 and terrain:
            out += [("bcwsfc", nrk), ("bc2d", nrk)]
        out += [("momentum", nrk), ("buoyancy", nrk), ("acoustic", nrk)]
        out += [("getdiv", nrk), ("getdivx", nrk)]
        out += [("pack_theta", nrk), ("scalar_advection", nrk), ("compose_theta", nrk)]
        if moist:
            for _ in range(numq):
                out += [("pack_scalar", nrk), ("pdef" if nrk == 3 else "scalar_advection", nrk),
                        ("compose_moisture", nrk)]
    if moist:
        out.append(("kessler_transition", None))
    out.append(("solve3_pressure", None))
    if moist:
        out.append(("solve3_density", None))
    return out


@pytest.mark.parametrize("moist", [False, True])
def test_composition_follows_the_reference_schedule(monkeypatch, moist):
    """The oracle's large step calls its pinned operators in the reference composer's order (outermost calls)."""
    from oracle import boundary as obc
    from oracle import kessler as okes

    trace = []
    depth = [0]
    stage = [None]

    def wrap(owner, attr, label):
        fn = getattr(owner, attr)

        def recorder(*args, **kwargs):
            if depth[0] == 0:
                trace.append((label, stage[0]))
            depth[0] += 1
            try:
                return fn(*args, **kwargs)
            finally:
                depth[0] -= 1

        monkeypatch.setattr(owner, attr, recorder)

    rk_stage = ol.rk_stage

    def staged(a, m, base, prof, cfg, nrk, **kw):
        stage[0] = nrk
        try:
            return rk_stage(a, m, base, prof, cfg, nrk, **kw)
        finally:
            stage[0] = None

    monkeypatch.setattr(ol, "rk_stage", staged)
    for owner, attr, label in (
            (ost, "predict_and_compose_held", "held"), (ost, "stage_terms", "stage_terms"),
            (ost, "seed_held", "seed_held"), (obc, "bcwsfc", "bcwsfc"), (obc, "bc2d_open_w_surface", "bc2d"),
            (ol, "cpu_terrain_open_momentum_advection5_inplace", "momentum"), (ost, "buoyancy", "buoyancy"),
            (ol, "cpu_psolver3_stage_inplace", "acoustic"), (ost, "getdiv", "getdiv"), (ost, "getdivx", "getdivx"),
            (ost, "pack_theta", "pack_theta"), (ol, "cpu_scalar_advection5_inplace", "scalar_advection"),
            (ost, "compose_theta", "compose_theta"), (ost, "pack_scalar", "pack_scalar"),
            (ol, "pdef_advection", "pdef"), (ost, "compose_moisture", "compose_moisture"),
            (okes, "large_step_transition", "kessler_transition"), (ost, "solve3_pressure", "solve3_pressure"),
            (ost, "solve3_density", "solve3_density")):
        wrap(owner, attr, label)
    cfg = moist_case(12, 10, 10, (2, 2, 2, 2)) if moist else dry_case(12, 10, 10, (2, 2, 2, 2))
    host = host_initial_state(cfg)
    a = {k: v.copy() for k, v in host["arrays"].items()}
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    c = {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}
    ol.large_step(a, host["metrics"], host["base"], prof, c, zt=F(cfg.nk * cfg.dz),
                  mass_correction=0.0 if moist else None, budgets=False)
    # bcwsfc also runs inside the acoustic stage's fills; only outermost calls are recorded
    assert trace == _reference_schedule(moist, terrain=True)


def _stage_inputs(rng, ni, nj, nk):
    sh = shapes(ni, nj, nk)
    m = seeded_metrics(rng, ni, nj, nk)
    a = {}
    for n in ("u3d", "uten", "rru", "uten1"):
        a[n] = (2 * rng.standard_normal(sh["u"])).astype(F)
    for n in ("v3d", "vten", "rrv", "vten1"):
        a[n] = (2 * rng.standard_normal(sh["v"])).astype(F)
    for n in ("w3d", "wten", "rrw", "wten1"):
        a[n] = rng.standard_normal(sh["w"]).astype(F)
    for n in ("thv", "ppterm", "piadv", "pp3d", "ppx", "ppi"):
        a[n] = (1e-3 * rng.standard_normal(sh["p"])).astype(F)
    a["dum8"] = np.zeros(sh["s"], F)
    a["thterm"] = np.zeros((nk, nj, ni), F)
    base = {"th0": np.linspace(300, 340, nk).astype(F), "qv0": np.linspace(0.014, 0.001, nk).astype(F),
            "pi0": np.linspace(1.0, 0.7, nk).astype(F)}
    th3d = rng.standard_normal((nk, nj, ni)).astype(F)
    q3d = np.abs(1e-3 * rng.standard_normal((3, nk, nj, ni))).astype(F)
    return a, m, base, th3d, q3d


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(1, 1, 1, 1), (2, 2, 2, 2), (3, 4, 2, 2)])
@pytest.mark.parametrize("moist", [False, True])
def test_gpu_stage_terms_and_buoyancy(device, sides, moist):
    import torch

    from gm1_amd import _native
    from gm1_amd._device import lateral_code, ptr, stream_handle

    ni, nj, nk = 19, 13, 9
    rng = np.random.default_rng(11 + moist)
    a, m, base, th3d, q3d = _stage_inputs(rng, ni, nj, nk)
    q = q3d if moist else None
    ha = {k: v.copy() for k, v in a.items()}
    qfull = None
    if moist:
        qfull = np.zeros((6, nk, nj, ni), F)
        qfull[:3] = q3d
    ost.stage_terms(ha, base, th3d, qfull, ni=ni, nj=nj, nk=nk, sides=sides)
    ost.buoyancy(ha, m, ni=ni, nj=nj, nk=nk)
    da, dm, db = to_dev(a, device), to_dev(m, device), to_dev(base, device)
    dth, dq = to_dev(th3d, device), (to_dev(q, device) if moist else None)
    h = stream_handle(None)
    _native.call("gm1hip_stage_terms", ptr(db["th0"]), ptr(db["qv0"]), ptr(db["pi0"]), None, None, ptr(dth),
                 ptr(dq), 3 if moist else 0, ptr(da["pp3d"]), ptr(da["thv"]), ptr(da["ppterm"]), ptr(da["piadv"]),
                 ptr(da["dum8"]), ptr(da["thterm"]), ni, nj, nk, lateral_code(sides), h)
    _native.call("gm1hip_buoyancy", ptr(da["dum8"]), ptr(dm["c1"]), ptr(dm["c2"]), ptr(dm["gxu"]), ptr(dm["gyv"]),
                 ptr(da["uten"]), ptr(da["vten"]), ptr(da["wten"]), ni, nj, nk, h)
    torch.cuda.synchronize()
    got = to_host(da)
    for name in ("thv", "ppterm", "thterm", "dum8", "uten", "vten", "wten"):
        assert_bits_equal(got[name], ha[name], name)
    ks = slice(1, nk + 1)
    assert_bits_equal(got["piadv"][ks], ha["piadv"][ks], "piadv")


@pytest.mark.gpu
def test_gpu_post_acoustic_and_compose(device):
    import torch

    from gm1_amd import _native
    from gm1_amd._device import ptr, stream_handle

    ni, nj, nk = 21, 17, 10
    rng = np.random.default_rng(5)
    a, m, base, th3d, q3d = _stage_inputs(rng, ni, nj, nk)
    a["thterm"] = rng.standard_normal((nk, nj, ni)).astype(F)
    ha = {k: v.copy() for k, v in a.items()}
    div, tdiv, divx = (np.zeros((nk, nj, ni), F) for _ in range(3))
    ost.getdiv(ha, m, div, tdiv, ha["thterm"], ni=ni, nj=nj, nk=nk, dx=250.0, dy=300.0, dz=200.0)
    ost.getdivx(ha, m, divx, ni=ni, nj=nj, nk=nk, dx=250.0, dy=300.0, dz=200.0)
    da, dm = to_dev(a, device), to_dev(m, device)
    dd = {n: torch.zeros((nk, nj, ni), dtype=torch.float32, device=device) for n in ("div", "tdiv", "divx")}
    h = stream_handle(None)
    _native.call("gm1hip_post_acoustic", ptr(da["rru"]), ptr(da["rrv"]), ptr(da["rrw"]), ptr(dm["rho0_3d"]),
                 ptr(dm["rf0"]), ptr(dm["rgzu"]), ptr(dm["rgzv"]), ptr(dm["dzdx"]), ptr(dm["dzdy"]), ptr(dm["gz"]),
                 ptr(da["thterm"]), ptr(dd["div"]), ptr(dd["tdiv"]), ptr(dd["divx"]), ni, nj, nk, 250.0, 300.0,
                 200.0, h)
    # compose: A + dttmp*((held - theta_div) + adv), |x| < 1e-30 -> 0
    tha = rng.standard_normal((nk, nj, ni)).astype(F)
    tha.flat[::7] = F(1e-31)
    held = (1e-3 * rng.standard_normal((nk, nj, ni))).astype(F)
    adv = (1e-3 * rng.standard_normal((nk, nj, ni))).astype(F)
    held.flat[::7] = F(0.0)
    adv.flat[::7] = F(0.0)
    tdiv_zeroed = tdiv.copy()
    tdiv_zeroed.flat[::7] = F(0.0)
    want = ost.compose_theta(tha, held, tdiv_zeroed, adv, F(2.0))
    dtha, dheld, dadv, dtd = (to_dev(x, device) for x in (tha, held, adv, tdiv_zeroed))
    out = torch.empty_like(dtha)
    _native.call("gm1hip_rk_compose", ptr(dtha), ptr(dheld), ptr(dtd), ptr(dadv), ptr(out), tha.size, 2.0, h)
    torch.cuda.synchronize()
    assert_bits_equal(to_host(dd["div"]), div, "divergence")
    assert_bits_equal(to_host(dd["tdiv"]), tdiv, "theta_div")
    assert_bits_equal(to_host(dd["divx"]), divx, "divx")
    got = to_host(da)
    for name in ("rru", "rrv", "rrw"):
        assert_bits_equal(got[name], ha[name], name)
    assert_bits_equal(to_host(out), want, "compose")


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(1, 1, 1, 1), (2, 2, 2, 2), (1, 1, 2, 2)])
def test_gpu_dry_large_step_bit_exact(device, sides):
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step

    cfg = dry_case(24, 20, 12, sides)
    host = host_initial_state(cfg)
    want = oracle_run(host, 2)
    st = create_gm1_large_step_state(host, device)
    for step in range(2):
        launch_cuda_gm1_large_step(st)
        torch.cuda.synchronize()
        for name in PROGNOSTIC:
            assert_bits_equal(to_host(st.a[name]), want[step][name], f"step {step + 1} {name}")


# ---------------------------------------------------------------- golden vectors
def test_eqtset2_terms_match_reference_cells():
    """cpu_eqtset2_cell_f32 (cuda_moist_stage_bridge.py:62-128) on 240 seeded cells."""
    from helpers import golden

    g = golden("stage")
    n = g["eq_in_th0"].size
    base = {"th0": g["eq_in_th0"], "qv0": g["eq_in_qv0"], "pi0": g["eq_in_pi0"], "qc0": g["eq_in_qc0"],
            "qi0": g["eq_in_qi0"]}
    a = {"thv": np.zeros((n + 2, 3, 3), F), "ppterm": np.zeros((n + 2, 3, 3), F),
         "piadv": np.zeros((n + 2, 3, 3), F), "pp3d": np.zeros((n + 2, 3, 3), F),
         "dum8": np.zeros((n + 2, 1 + 2 * NG, 1 + 2 * NG), F), "thterm": np.zeros((n, 1, 1), F)}
    a["pp3d"][1 : n + 1, 1, 1] = g["eq_in_pp3d"]
    q = np.stack([g[f"eq_in_{s}"] for s in ("qv", "qc", "qr", "qi", "qs", "qg")])[:, :, None, None]
    ost.stage_terms(a, base, g["eq_in_th3d"][:, None, None], q, ni=1, nj=1, nk=n, sides=(2, 2, 2, 2))
    assert_bits_equal(a["dum8"][1 : n + 1, NG, NG], g["eq_out_buoyancy"], "buoyancy")
    assert_bits_equal(a["thv"][1 : n + 1, 1, 1], g["eq_out_theta_rho"], "theta_rho")
    assert_bits_equal(a["thterm"][:, 0, 0], g["eq_out_theta_term"], "theta_term")
    assert_bits_equal(a["ppterm"][1 : n + 1, 1, 1], g["eq_out_pressure_term"], "pressure_term")


def test_theta_packing_and_compose_match_reference():
    """cpu_theta_transport_scalar_inplace and cpu_compose_scalar_moist_rk_stage_inplace
    (cuda_scalar_moist_rk.py:255-349) on seeded carriers, all three stages."""
    from helpers import golden

    g = golden("stage")
    ni, nj, nk = (int(x) for x in g["dims"])
    padded = np.zeros((nk + 2, nj + 2 * NG, ni + 2 * NG), F)
    ost.pack_theta(g["pack_th0"], g["pack_th3d"], padded, ni=ni, nj=nj, nk=nk)
    assert_bits_equal(padded, g["pack_padded"], "theta packing")
    for nrk in (1, 2, 3):
        dttmp = F(F(1.5) / F(4 - nrk))
        th = ost.compose_theta(g["comp_in_theta_a"], g["comp_in_theta_held"], g["comp_in_theta_divergence_term"],
                               g["comp_in_theta_advection"], dttmp)
        q = ost.compose_moisture(g["comp_in_moisture_a"], g["comp_in_moisture_held"],
                                 g["comp_in_moisture_advection"], dttmp)
        assert_bits_equal(th, g[f"comp_nrk{nrk}_theta"], f"theta nrk={nrk}")
        assert_bits_equal(q, g[f"comp_nrk{nrk}_moisture"], f"moisture nrk={nrk}")
    # the reference zeroes only after the update: the seeded 1e-31 anchors clamp
    assert not np.any(g["comp_nrk3_theta"].flat[::13])


def test_held_seed_and_bc2d_match_reference():
    """cpu_seed_momentum_rk_tendencies_inplace and cpu_bc2d_open_w_surface_inplace
    (cuda_momentum_rk.py:242-290)."""
    from oracle import boundary as obc
    from helpers import golden

    g = golden("stage")
    ni, nj, nk = (int(x) for x in g["dims"])
    held = [g[f"seed_in_{n}"] for n in ("uten1", "vten1", "wten1")]
    tend = [g[f"seed_in_{n}"].copy() for n in ("uten", "vten", "wten")]
    ost.seed_held(held, tend, ni=ni, nj=nj, nk=nk)
    for name, t in zip(("uten", "vten", "wten"), tend):
        assert_bits_equal(t, g[f"seed_out_{name}"], name)
    w = g["bc2d_in_w"].copy()
    obc.bc2d_open_w_surface(w, ni=ni, nj=nj)
    assert_bits_equal(w, g["bc2d_out_w"], "bc2d")


def moist_case(ni, nj, nk, sides, *, seed=5):
    return CaseConfig(f"moist-{ni}x{nj}x{nk}", ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 8, tuple(sides), True,
                      "hodograph", ((3.0, ni * 120.0, nj * 130.0, 1000.0),), bubble_radius=(ni * 80.0, 1000.0),
                      noise=0.1, seed=seed, extra={"bubble_rh": 1.03})


MOIST_FIELDS = PROGNOSTIC + ("q3d", "rho", "prs", "rr", "rf", "rain", "prate", "qpten", "qtten", "qvten", "qcten",
                             "fall_velocity")


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_gpu_moist_large_step_bit_exact(device, sides):
    """Two moist large steps (Kessler, pdef moisture, dry-mass correction) against the oracle.

    The dry-mass correction is the one reduction whose order differs (a fixed
    tree on the device); the oracle applies the device's binary32 correction
    after checking its ledger agrees to 1e-13.
    """
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, kessler_status, launch_cuda_gm1_large_step

    cfg = moist_case(24, 20, 16, sides)
    host = host_initial_state(cfg)
    st = create_gm1_large_step_state(host, device)
    a = {k: v.copy() for k, v in host["arrays"].items()}
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    c = {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}
    for step in range(2):
        launch_cuda_gm1_large_step(st)
        torch.cuda.synchronize()
        corr = float(st.dry_mass.correction.cpu().numpy()[0])
        ol.large_step(a, host["metrics"], host["base"], prof, c, zt=F(cfg.nk * cfg.dz), mass_correction=corr)
        for name in MOIST_FIELDS:
            assert_bits_equal(to_host(st.a[name]), a[name], f"step {step + 1} {name}")
        np.testing.assert_allclose(to_host(st.a["qbudget"]), a["qbudget"], rtol=1e-11, atol=1e-25)
        np.testing.assert_allclose(to_host(st.a["asq"]), a["asq"], rtol=1e-11, atol=1e-25)
    # the bubble condensed
    assert a["q3d"][1].max() > 1e-5
    assert kessler_status(st)["non_converged_cells"] == 0


@pytest.mark.gpu
def test_gpu_graph_replays_match_eager_steps(device):
    """A captured moist large step replayed three times equals three eager steps, replay by replay.

    Eager work and host synchronisations run between the replays.
    """
    import torch

    from gm1_amd.graph import Gm1LargeStepGraph
    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step

    cfg = moist_case(24, 20, 16, (2, 2, 2, 2), seed=9)
    host = host_initial_state(cfg)
    eager = create_gm1_large_step_state(host, device)
    graphed = create_gm1_large_step_state(host, device, stream=torch.cuda.Stream())
    g = Gm1LargeStepGraph(graphed, steps=1)
    for step in range(3):
        launch_cuda_gm1_large_step(eager)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        scratch = torch.ones(1 << 16, device=device).cumsum(0)  # eager work between replays
        torch.cuda.synchronize()
        assert float(scratch[-1]) == float(1 << 16)
        for name in MOIST_FIELDS:
            assert_bits_equal(to_host(graphed.a[name]), to_host(eager.a[name]), f"replay {step + 1} {name}")
    assert graphed.step_index == eager.step_index == 3
