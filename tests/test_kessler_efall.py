"""Kessler efall=1 (energy of falling rain): the oracle against the reference's CPU functions, GPU parity.

Golden vectors (``tests/golden/kessler_efall.npz``, ``make_golden.py kessler_efall``) hold
the outputs of ``cpu_gm1_kessler_efall_inplace`` (the isolated getcvm / getefall /
guarded update, cuda_gm1_kessler_efall.py:201-256), of ``cpu_gm1_kessler_driver_inplace``
with efall=1 (cuda_gm1_kessler_driver.py:482-551) and of
``cpu_gm1_kessler_large_step_inplace`` with efall=1.  The oracle reproduces them word
for word with NumPy's float32 transcendentals; the device matches the oracle word for
word with binary64-rounded ones.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, assert_bits_equal, golden
from oracle import kessler as ok

FIELDS = ("rho", "rr", "pp3d", "th3d", "prs", "qv", "qc", "qr", "fall_velocity", "rain", "prate")


def _inputs(g, prefix):
    return {k[len(prefix):]: g[k].copy() for k in g.files if k.startswith(prefix)}


@pytest.mark.parametrize("case", [0, 1])
def test_efall_operator_matches_reference(case):
    g = golden("kessler_efall")
    a = _inputs(g, f"e{case}_in_")
    dt, dz, eqt = (float(v) for v in g[f"e{case}_params"])
    t = a.pop("temperature")
    cvm, tend = ok.efall(a, t, dz=dz, dt=dt, equation_set=int(eqt))
    assert_bits_equal(cvm, g[f"e{case}_cvm"], "cvm")
    assert_bits_equal(tend, g[f"e{case}_tendency"], "tendency")
    assert_bits_equal(t, g[f"e{case}_out_temperature"], "temperature")
    for name in ("pp3d", "th3d", "prs"):
        assert_bits_equal(a[name], g[f"e{case}_out_{name}"], name)


def test_efall_cases_take_both_sides_of_the_update_guard():
    g = golden("kessler_efall")
    for case in (0, 1):
        changed = g[f"e{case}_out_temperature"] != g[f"e{case}_in_temperature"]
        assert changed.any() and (~changed[:-1]).any()
        assert not changed[-1].any()  # the top level takes no energy fallout
        assert (g[f"e{case}_tendency"][-1] == 0).all()


def test_efall_closed_forms():
    """The tendency is qr*vt*(cpl*dT/dz*mf + g)/cvm: zero without rain, g*qr*vt/cvm in an isothermal column,
    and the updated state satisfies the equation of state it was formed from."""
    g = golden("kessler_efall")
    a = _inputs(g, "e0_in_")
    t = a.pop("temperature")
    t[...] = F(280.0)
    cvm, tend = ok.efall({k: v.copy() for k, v in a.items()}, t.copy(), dz=250.0, dt=1.5, equation_set=2)
    want = (a["qr"][:-1].astype(np.float64) * a["fall_velocity"][:-1] * 9.81) / cvm[:-1]
    np.testing.assert_allclose(tend[:-1], want, rtol=2e-6, atol=1e-30)
    dry = {k: v.copy() for k, v in a.items()}
    dry["qr"][...] = F(0.0)
    _, tend0 = ok.efall(dry, t.copy(), dz=250.0, dt=1.5, equation_set=2)
    assert (tend0 == 0).all()
    # eqtset 1: cvm is the dry cv alone
    cvm1, _ = ok.efall({k: v.copy() for k, v in a.items()}, t.copy(), dz=250.0, dt=1.5, equation_set=1)
    assert (cvm1 == ok.CV).all()


@pytest.mark.parametrize("case", [0, 1, 2])
def test_driver_with_efall_matches_reference(case):
    g = golden("kessler_efall")
    a = _inputs(g, f"d{case}_in_")
    dt, dz, eqt = (float(v) for v in g[f"d{case}_params"])
    res = ok.kessler_driver(a, dx=250.0, dy=250.0, dz=dz, dt=dt, equation_set=int(eqt), efall_on=1)
    for name in FIELDS:
        assert_bits_equal(a[name], g[f"d{case}_out_{name}"], f"case {case} {name}")
    assert_bits_equal(res["temperature_held"], g[f"d{case}_temperature"], "held temperature")
    np.testing.assert_array_equal(res["iterations"], g[f"d{case}_iterations"])
    assert res["max_fall_substeps"] == int(g[f"d{case}_nfall"])
    np.testing.assert_array_equal(a["asq"], g[f"d{case}_out_asq"])
    np.testing.assert_array_equal(a["qbudget"], g[f"d{case}_out_qbudget"])


def test_driver_with_efall_differs_from_efall_zero():
    """A plausible mistake -- efall skipped, or the local conversion reading a fresh temperature -- fails."""
    g = golden("kessler_efall")
    a0 = _inputs(g, "d0_in_")
    a1 = {k: v.copy() for k, v in a0.items()}
    ok.kessler_driver(a0, dx=250.0, dy=250.0, dz=250.0, dt=1.5, equation_set=2, efall_on=0)
    ok.kessler_driver(a1, dx=250.0, dy=250.0, dz=250.0, dt=1.5, equation_set=2, efall_on=1)
    assert (a0["th3d"] != a1["th3d"]).any() and (a0["pp3d"] != a1["pp3d"]).any()
    assert (a1["th3d"] == g["d0_out_th3d"]).all() and not (a0["th3d"] == g["d0_out_th3d"]).all()


def test_large_step_transition_with_efall_matches_reference():
    g = golden("kessler_efall")
    a = _inputs(g, "t_in_")
    held = {k: a.pop(k) for k in ("qpten", "qtten", "qvten", "qcten")}
    ok.large_step_transition(a, held, dx=250.0, dy=250.0, dz=250.0, dt=1.5, equation_set=2, efall_on=1)
    for name in FIELDS:
        assert_bits_equal(a[name], g[f"t_out_{name}"], name)
    for name in held:
        assert_bits_equal(held[name], g[f"t_out_{name}"], name)


def test_no_allocation_driver_refuses_efall_before_touching_anything():
    from gm1_amd.errors import CM1Refusal
    from gm1_amd.kessler import Gm1KesslerDriverSelection, launch_cuda_gm1_kessler_driver_noalloc

    with pytest.raises(CM1Refusal):
        launch_cuda_gm1_kessler_driver_noalloc({}, None, ni=4, nj=4, nk=4, dx=1.0, dy=1.0, dz=1.0, dt=1.0,
                                               selection=Gm1KesslerDriverSelection(equation_set=2, efall=1))


def test_driver_selection_admits_efall_zero_and_one_only():
    from gm1_amd.errors import CM1Refusal
    from gm1_amd.kessler import Gm1KesslerDriverSelection, validate_selection

    validate_selection(Gm1KesslerDriverSelection(efall=1))
    with pytest.raises(CM1Refusal):
        validate_selection(Gm1KesslerDriverSelection(efall=2))


# ---------------------------------------------------------------- GPU parity
def _device_arrays(a, device):
    import torch

    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in a.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("case", [0, 1, 2])
def test_gpu_driver_with_efall_bit_exact_against_rounded_oracle(device, case):
    """Fields, the held temperature and the branch words bit for bit against the oracle with binary64-rounded
    transcendentals; budgets to 1e-12; the reference's own words to 1e-5."""
    import torch

    from gm1_amd.kessler import Gm1KesslerDriverSelection, launch_cuda_gm1_kessler_driver

    g = golden("kessler_efall")
    ni, nj, nk = (int(v) for v in g["dims"])
    a = _inputs(g, f"d{case}_in_")
    dt, dz, eqt = (float(v) for v in g[f"d{case}_params"])
    print("efall driver case", case, (dt, dz, eqt), flush=True)
    da = _device_arrays(a, device)
    words = torch.zeros((nk, nj, ni), dtype=torch.int32, device=device)
    ws = launch_cuda_gm1_kessler_driver(da, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=dz, dt=dt,
                                        selection=Gm1KesslerDriverSelection(equation_set=int(eqt), efall=1),
                                        decisions=words)
    torch.cuda.synchronize()
    with ok.transcendentals("rounded"):
        res = ok.kessler_driver(a, dx=250.0, dy=250.0, dz=dz, dt=dt, equation_set=int(eqt), efall_on=1)
    for name in FIELDS:
        assert_bits_equal(da[name].cpu().numpy(), a[name], f"case {case} {name}")
    assert_bits_equal(ws.temperature.cpu().numpy(), res["temperature_held"], "held temperature")
    np.testing.assert_array_equal(words.cpu().numpy().view(np.uint32), res["decisions"])
    np.testing.assert_allclose(da["asq"].cpu().numpy(), a["asq"], rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(da["qbudget"].cpu().numpy(), a["qbudget"], rtol=1e-12, atol=1e-30)
    assert ws.status.cpu().numpy()[1] == 0
    totals = {"pp3d": g[f"d{case}_in_pi0"], "th3d": g[f"d{case}_in_th0"]}
    for name in FIELDS:
        off = totals.get(name, 0.0)
        want = g[f"d{case}_out_{name}"].astype(np.float64) + off
        got = da[name].cpu().numpy().astype(np.float64) + off
        assert np.abs(got - want).max() / max(np.abs(want).max(), 1e-30) < 1e-5, name


@pytest.mark.gpu
def test_gpu_transition_with_efall_matches_oracle(device):
    import torch

    from gm1_amd.kessler import (Gm1KesslerDriverSelection, create_cuda_gm1_kessler_driver_workspace,
                                 launch_cuda_gm1_kessler_transition)

    g = golden("kessler_efall")
    ni, nj, nk = (int(v) for v in g["dims"])
    a = _inputs(g, "t_in_")
    held = {k: a.pop(k) for k in ("qpten", "qtten", "qvten", "qcten")}
    da, dh = _device_arrays(a, device), _device_arrays(held, device)
    ws = create_cuda_gm1_kessler_driver_workspace(ni, nj, nk, device)
    launch_cuda_gm1_kessler_transition(da, dh, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0, dt=1.5,
                                       selection=Gm1KesslerDriverSelection(equation_set=2, efall=1), workspace=ws)
    torch.cuda.synchronize()
    with ok.transcendentals("rounded"):
        ok.large_step_transition(a, held, dx=250.0, dy=250.0, dz=250.0, dt=1.5, equation_set=2, efall_on=1)
    for name in FIELDS:
        assert_bits_equal(da[name].cpu().numpy(), a[name], name)
    for name in held:
        assert_bits_equal(dh[name].cpu().numpy(), held[name], name)


@pytest.mark.gpu
def test_gpu_driver_with_efall_on_c2_like_fields(device):
    """327 680 cloudy cells of the 320x320x80 case's column with a seeded mf: device equals the rounded oracle."""
    import torch

    from gm1_amd.kessler import Gm1KesslerDriverSelection, launch_cuda_gm1_kessler_driver
    from test_kessler import c2_like_kessler_fields

    cfg, k = c2_like_kessler_fields()
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    rng = np.random.default_rng(17)
    k["mf"] = (1.0 + 0.05 * rng.uniform(-1.0, 1.0, (nk + 1, nj, ni))).astype(F)
    arrays = {n: v for n, v in k.items() if isinstance(v, np.ndarray)}
    print("efall c2-like", ni, nj, nk, flush=True)
    da = _device_arrays(arrays, device)
    ws = launch_cuda_gm1_kessler_driver(da, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0, dt=1.5,
                                        selection=Gm1KesslerDriverSelection(equation_set=2, efall=1))
    torch.cuda.synchronize()
    b = {n: (v.copy() if isinstance(v, np.ndarray) else v) for n, v in k.items()}
    with ok.transcendentals("rounded"):
        res = ok.kessler_driver(b, dx=250.0, dy=250.0, dz=250.0, dt=1.5, equation_set=2, budgets=False, efall_on=1)
    for name in FIELDS:
        assert_bits_equal(da[name].cpu().numpy(), b[name], name)
    assert_bits_equal(ws.temperature.cpu().numpy(), res["temperature_held"], "held temperature")
    t0 = (b["th0"] + arrays["th3d"]) * (b["pi0"] + arrays["pp3d"])
    assert (res["temperature_held"] != t0).any()  # the energy fallout moved some temperatures
