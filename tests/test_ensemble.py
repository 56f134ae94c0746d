"""Ensembles of perturbed storms on uneven grids.

CPU: the member draws are deterministic in the seed and stay inside the
declared ranges (sizes, levels, amplitude, centre offset, moisture scale).
GPU: a batched ensemble (each member's large step a HIP graph, replays spread
over streams) reproduces every member's own eager steps bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_bits_equal, to_host
from gm1_amd.fields import ENSEMBLE_LEVELS, ENSEMBLE_SIZES, CaseConfig, ensemble_members

FIELDS = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ppd", "q3d", "rho", "prs", "rain", "qpten")


def test_members_are_deterministic_and_in_range():
    a, b = ensemble_members(7, 24), ensemble_members(7, 24)
    assert [m.name for m in a] == [m.name for m in b]
    assert ensemble_members(8, 24)[0].name != a[0].name or ensemble_members(8, 24)[0].bubbles != a[0].bubbles
    for m in a:
        assert m.ni in ENSEMBLE_SIZES and m.nj in ENSEMBLE_SIZES and m.nk in ENSEMBLE_LEVELS
        amp, x0, y0, _ = m.bubbles[0]
        assert 1.5 <= amp <= 3.5
        assert abs(x0 - m.ni * 125.0) <= 5000.0 and abs(y0 - m.nj * 125.0) <= 5000.0
        assert 0.9 <= m.extra["qv_scale"] <= 1.1
        assert (m.dx, m.dz, m.dt, m.nsound, m.lateral_sides) == (250.0, 250.0, 1.5, 8, (2, 2, 2, 2))
    total = sum(m.cells for m in ensemble_members())
    assert 60e6 < total < 120e6


def test_moisture_scale_scales_the_base_profile():
    from gm1_amd.fields import base_state, profiles_for

    base = dict(ni=16, nj=16, nk=20, dx=250.0, dy=250.0, dz=250.0, dt=1.5, nsound=8, lateral_sides=(2, 2, 2, 2),
                moist=True, wind="hodograph", bubbles=((2.0, 2000.0, 2000.0, 1400.0),))
    c1 = CaseConfig("a", **base)
    c2 = CaseConfig("b", **base, extra={"qv_scale": 1.1})
    q1 = base_state(c1, profiles_for(c1))["qv0"].astype(np.float64)
    q2 = base_state(c2, profiles_for(c2))["qv0"].astype(np.float64)
    assert np.all(q2 > q1)
    np.testing.assert_allclose(q2[0] / q1[0], 1.1, rtol=0.02)


def small_members():
    out = []
    for m, (ni, nj, nk, amp) in enumerate(((24, 20, 16, 3.0), (20, 28, 16, 2.0), (32, 24, 20, 2.5))):
        out.append(CaseConfig(f"small-{m}", ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 8, (2, 2, 2, 2), True, "hodograph",
                              ((amp, ni * 120.0, nj * 130.0, 1000.0),), bubble_radius=(ni * 80.0, 1000.0),
                              noise=0.1, seed=40 + m, extra={"bubble_rh": 1.03, "qv_scale": 0.95 + 0.05 * m}))
    return out


@pytest.mark.gpu
def test_gpu_batched_ensemble_equals_each_members_eager_steps(device):
    import torch

    from gm1_amd.ensemble import Gm1Ensemble, create_ensemble_states, sequential_steps

    members = small_members()
    print("ensemble of", [m.name for m in members], flush=True)
    eager = create_ensemble_states(members, device)
    batched = create_ensemble_states(members, device)
    ens = Gm1Ensemble(batched, streams=2)
    for step in range(2):
        sequential_steps(eager, 1)
        ens.step()
        torch.cuda.synchronize()
        for e, b in zip(eager, batched):
            for name in FIELDS:
                assert_bits_equal(to_host(b.a[name]), to_host(e.a[name]), f"{b.cfg.name} step {step + 1} {name}")
    assert all(st.step_index == 2 for st in batched)


@pytest.mark.gpu
def test_gpu_autotuned_members_equal_untuned_members(device):
    import torch

    from gm1_amd import autotune
    from gm1_amd.ensemble import create_ensemble_states, sequential_steps

    members = small_members()[:2]
    print("autotune of", [m.name for m in members], flush=True)
    plain = create_ensemble_states(members, device)
    tuned = create_ensemble_states(members, device, autotune=True)
    for st in tuned:
        rec = autotune.tune_column_pass(st)
        assert st.column_pass == rec["choice"] in autotune.CANDIDATES
        assert set(rec["us"]) == set(autotune.CANDIDATES) and all(v > 0 for v in rec["us"].values())
    sequential_steps(plain, 2)
    sequential_steps(tuned, 2)
    torch.cuda.synchronize()
    for p, t in zip(plain, tuned):
        for name in FIELDS:
            assert_bits_equal(to_host(t.a[name]), to_host(p.a[name]), f"{t.cfg.name} tuned {t.column_pass} {name}")


def test_autotune_choice_is_read_from_the_cache_file(tmp_path, monkeypatch):
    """A cached choice for a grid shape is returned without timing anything."""
    import json
    from types import SimpleNamespace

    from gm1_amd import autotune

    cfg = SimpleNamespace(ni=48, nj=40, nk=20, lateral_sides=(2, 2, 2, 2))
    st = SimpleNamespace(cfg=cfg, flat=None, column_pass="pipelined")
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({autotune._key(st): {"choice": "staged16", "us": {"staged16": 1.0}}}))
    monkeypatch.setenv("GM1_AUTOTUNE_CACHE", str(path))
    monkeypatch.setattr(autotune, "_memo", {})
    assert autotune.autotune_state(st) == "staged16" and st.column_pass == "staged16"
    flat = SimpleNamespace(cfg=cfg, flat=object(), column_pass="pipelined")
    assert autotune._key(flat) != autotune._key(st)  # the stretched family is tuned on its own
    # a cached entry naming no known pass is not trusted; every candidate runs on both families
    assert not autotune._valid({"choice": "bogus"}, False) and not autotune._valid({"choice": "register"}, True)
    assert autotune._valid({"choice": "staged16"}, False) and autotune._valid({"choice": "pipelined16"}, True)
    # the key names the built library, so a cache written by another build is not reused
    assert autotune._library_tag() in autotune._key(st)
