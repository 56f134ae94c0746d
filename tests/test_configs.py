"""The benchmark configurations at their own sizes and lengths.

* 64x64x40 dry periodic (dx 1 km, dz 500 m, dt 6 s, 6 substeps): ten large
  steps on the device equal the oracle's word for word, step by step.
* The 320x320x80 moist case's physics on a 48x48x40 grid, 600 large steps:
  the device's run equals the oracle's word for word at every 50th step
  (``tests/golden/moist_long_run.npz``, written by
  ``tests/golden/make_long_run.py``).  The oracle replays the device's
  binary32 dry-mass corrections (``moist_long_run_device.npz``) after
  checking its own ledger against the device's every step; the one
  reduction whose order differs.  Beside it, the oracle's run with its own
  corrections (``moist_long_run_independent.npz``) shows how far max w,
  min pi' and the condensate move under those last-bit differences.
* The seeded 24-member ensemble at its own sizes (128-320 x 128-320 x 64-80,
  86.5 M cells): two batched steps equal every member's eager steps, and
  every member stays finite and bounded; one member's physics on a reduced
  grid equals the oracle word for word.
"""

from __future__ import annotations

import hashlib
from dataclasses import replace

import numpy as np
import pytest

from helpers import GOLDEN, assert_bits_equal, to_host
from gm1_amd.fields import dry_c1, ensemble_members, moist_c2
from gm1_amd.large_step import host_initial_state
from oracle import large_step as ol
from oracle import sound as osound

PROGNOSTIC = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ppd")
MOIST_FIELDS = PROGNOSTIC + ("q3d", "rho", "prs", "rr", "rf", "rain", "prate", "qpten", "qtten", "qvten", "qcten")
NG = 3


def _oracle_cfg(cfg):
    return {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
            "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}


def long_run_case():
    """The moist case's physics on 48x48x40 with the bubble radius scaled to the domain (2 km)."""
    return replace(moist_c2(48, 48, 40), name="moist-kessler-48x48x40", bubble_radius=(2000.0, 1400.0))


def long_run_diagnostics(w3d, pp3d, q3d, rho, ni, nj, nk, dx, dy, dz) -> np.ndarray:
    """(max w, min pi', total condensate kg), binary64 from the binary32 words."""
    w = w3d[1 : nk + 2, NG : NG + nj, NG : NG + ni].astype(np.float64)
    pp = pp3d[1 : nk + 1, 1 : nj + 1, 1 : ni + 1].astype(np.float64)
    q = q3d.astype(np.float64)
    condensate = float(np.sum(rho.astype(np.float64) * (q[1] + q[2]))) * dx * dy * dz
    return np.array([w.max(), pp.min(), condensate], np.float64)


LONG_RUN_CARRIERS = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppi", "ppx", "q3d", "rho", "prs", "rain")


def carrier_digest(arrays) -> str:
    h = hashlib.sha256()
    for name in LONG_RUN_CARRIERS:
        h.update(np.ascontiguousarray(arrays[name]).tobytes())
    return h.hexdigest()


def test_long_run_fixture_matches_its_case():
    """The committed long-run records are the 600-step, 50-step-cadence records of this case: the device's
    ledger, the oracle's replay of it (every ledger within 1e-13) and the oracle's own run; the bubble rose
    and condensed, and the two oracle runs start identical."""
    replay = np.load(GOLDEN / "moist_long_run.npz")
    indep = np.load(GOLDEN / "moist_long_run_independent.npz")
    device = np.load(GOLDEN / "moist_long_run_device.npz")
    for fx in (replay, indep):
        assert list(fx["steps"]) == list(range(50, 601, 50))
        assert fx["corrections"].shape == (600,) and np.all(np.isfinite(fx["corrections"]))
        assert tuple(fx["carriers"]) == LONG_RUN_CARRIERS
        d = fx["diagnostics"]
        assert d.shape == (12, 3) and np.all(np.isfinite(d))
        assert d[:, 2].max() > 1e6 and d[:, 0].max() > 5.0  # the bubble rose and condensed
    assert replay["corrections"].tobytes() == device["corrections"].tobytes()
    assert np.all(replay["ledger_rel"] <= 1e-13)
    first = int(np.argmax(replay["corrections"].view(np.uint32) != indep["corrections"].view(np.uint32)))
    assert first > 0 and replay["corrections"][:first].tobytes() == indep["corrections"][:first].tobytes()
    # the records start from this case's initial words (the same initializer)
    host = host_initial_state(long_run_case())
    digest = carrier_digest(host["arrays"])
    assert digest == str(replay["initial_digest"]) == str(indep["initial_digest"]) == str(device["initial_digest"])


@pytest.mark.gpu
def test_gpu_dry_64x64x40_ten_steps_bit_exact(device):
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step

    cfg = dry_c1()
    print("case", cfg.name, flush=True)
    host = host_initial_state(cfg)
    st = create_gm1_large_step_state(host, device)
    a = {k: v.copy() for k, v in host["arrays"].items()}
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    c = _oracle_cfg(cfg)
    for step in range(10):
        launch_cuda_gm1_large_step(st)
        ol.large_step(a, host["metrics"], host["base"], prof, c, zt=np.float32(cfg.nk * cfg.dz))
        torch.cuda.synchronize()
        for name in PROGNOSTIC:
            assert_bits_equal(to_host(st.a[name]), a[name], f"step {step + 1} {name}")
    assert np.abs(a["w3d"]).max() > 0.01


@pytest.mark.gpu
def test_gpu_moist_600_steps_against_the_oracle_record(device):
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, kessler_status, launch_cuda_gm1_large_step

    fx = np.load(GOLDEN / "moist_long_run.npz")
    indep = np.load(GOLDEN / "moist_long_run_independent.npz")
    cfg = long_run_case()
    print("case", cfg.name, "600 steps", flush=True)
    host = host_initial_state(cfg)
    assert carrier_digest(host["arrays"]) == str(fx["initial_digest"]), "initial words differ from the record's"
    st = create_gm1_large_step_state(host, device)
    del host
    checkpoints = {int(s): i for i, s in enumerate(fx["steps"])}
    for step in range(1, 601):
        launch_cuda_gm1_large_step(st)
        corr = np.float32(st.dry_mass.correction[0].item())
        assert corr.tobytes() == fx["corrections"][step - 1].tobytes(), (step, corr, fx["corrections"][step - 1])
        if step in checkpoints:
            torch.cuda.synchronize()
            h = {n: to_host(st.a[n]) for n in LONG_RUN_CARRIERS}
            got = long_run_diagnostics(h["w3d"], h["pp3d"], h["q3d"], h["rho"], cfg.ni, cfg.nj, cfg.nk,
                                       cfg.dx, cfg.dy, cfg.dz)
            c = checkpoints[step]
            other = indep["diagnostics"][c]
            sens = np.abs(got - other) / np.maximum(np.abs(other), 1e-30)
            print(f"step {step}: max w {got[0]:.5f} min pp {got[1]:.6e} condensate {got[2]:.6e}; against the "
                  f"oracle's own-correction run {sens[0]:.1e} {sens[1]:.1e} {sens[2]:.1e}", flush=True)
            assert carrier_digest(h) == str(fx["digests"][c]), f"step {step}: carriers differ from the oracle's"
            assert got.tobytes() == fx["diagnostics"][c].tobytes()
    assert kessler_status(st)["non_converged_cells"] == 0


def _same_words(x, y) -> bool:
    import torch

    return x.shape == y.shape and bool(torch.equal(x.contiguous().view(torch.int32), y.contiguous().view(torch.int32)))


@pytest.mark.gpu
def test_gpu_seeded_ensemble_at_its_sizes(device):
    import torch

    from gm1_amd.ensemble import Gm1Ensemble, create_ensemble_states, sequential_steps

    members = ensemble_members()
    print("ensemble", len(members), "members,", sum(m.cells for m in members), "cells", flush=True)
    eager = create_ensemble_states(members, device)
    batched = create_ensemble_states(members, device)
    ens = Gm1Ensemble(batched, streams=4)
    names = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ppd", "q3d", "rho", "prs", "rain", "qpten")
    for step in range(2):
        sequential_steps(eager, 1)
        ens.step()
        torch.cuda.synchronize()
        for e, b in zip(eager, batched):
            for name in names:
                assert _same_words(b.a[name], e.a[name]), f"{b.cfg.name} step {step + 1} {name}"
        print("step", step + 1, "batched == eager for every member", flush=True)
    for st in batched:
        for name in names:
            assert bool(torch.isfinite(st.a[name]).all()), f"{st.cfg.name} {name} not finite"
        w = float(st.a["w3d"].abs().max())
        u = float(st.a["u3d"].abs().max())
        th = float(st.a["th3d"].abs().max())
        q = st.a["q3d"]
        assert w < 30.0 and u < 60.0 and th < 10.0, (st.cfg.name, w, u, th)
        assert float(q.min()) >= 0.0 and float(q[0].max()) < 0.03, st.cfg.name
        assert st.step_index == 2


@pytest.mark.gpu
def test_gpu_ensemble_member_physics_on_a_reduced_grid_bit_exact(device):
    """Member 0's draws (amplitude, moisture scale, noise seed, centre offset scaled to the grid) at 40x36x24."""
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step

    m = ensemble_members()[0]
    amp, x0, y0, z0 = m.bubbles[0]
    ni, nj, nk = 40, 36, 24
    cfg = replace(m, name=f"{m.name}-reduced-{ni}x{nj}x{nk}", ni=ni, nj=nj, nk=nk,
                  bubbles=((amp, x0 * ni / m.ni, y0 * nj / m.nj, z0),), bubble_radius=(2500.0, 1400.0))
    print("case", cfg.name, flush=True)
    host = host_initial_state(cfg)
    st = create_gm1_large_step_state(host, device)
    a = {k: v.copy() for k, v in host["arrays"].items()}
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    c = _oracle_cfg(cfg)
    for step in range(3):
        launch_cuda_gm1_large_step(st)
        torch.cuda.synchronize()
        corr = float(st.dry_mass.correction.cpu().numpy()[0])
        ol.large_step(a, host["metrics"], host["base"], prof, c, zt=np.float32(cfg.nk * cfg.dz),
                      mass_correction=corr)
        for name in MOIST_FIELDS:
            assert_bits_equal(to_host(st.a[name]), a[name], f"step {step + 1} {name}")
