"""Halo-free dry periodic step on the device (gm1_amd.adjoint_step) and its adjoint.

CPU: the workspace slot list matches the kernels' pointer table.
GPU: the device forward step equals the binary32 restatement
(oracle.large_step) word for word on a C5-like state; the device adjoint of
one large step against the binary64 adjoint (oracle/adjoint_large_step.py)
at the same trajectory.
"""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

from gm1_amd.adjoint_step import SLOTS
from gm1_amd.large_step import host_initial_state
from test_adjoint_sound import c5_like

ROOT = Path(__file__).resolve().parents[1]
FIELDS = ("u", "v", "w", "th", "pp", "ppx")


def test_slot_table_matches_the_kernels():
    text = (ROOT / "gm1_amd" / "csrc" / "adjoint_step.hip").read_text()
    body = re.search(r"enum Slot \{(.*?)NSLOT", text, flags=re.S).group(1)
    body = re.sub(r"//[^\n]*", "", body)
    names = [n.strip() for n in body.split(",") if n.strip()]
    assert tuple(names) == SLOTS


def state32(seed=3, ni=14, nj=12, nk=10):
    """A binary32 C5-like padded state with winds, theta' and pressure everywhere (periodic-consistent)."""
    from oracle import boundary as bc
    from oracle import sound as osd

    cfg = c5_like(ni, nj, nk, seed=seed)
    host = host_initial_state(cfg)
    a = {k: v.copy() for k, v in host["arrays"].items()}
    rng = np.random.default_rng(seed)
    ng = 3
    f32 = np.float32
    a["th3d"] = (a["th3d"] + 0.3 * rng.standard_normal(a["th3d"].shape)).astype(f32)
    for name in ("pp3d", "ppx"):
        a[name][1 : nk + 1, 1 : nj + 1, 1 : ni + 1] += (1e-4 * rng.standard_normal((nk, nj, ni))).astype(f32)
    a["w3d"][2 : nk + 1, ng : ng + nj, ng : ng + ni] += (0.5 * rng.standard_normal((nk - 1, nj, ni))).astype(f32)
    for name in ("u3d", "v3d"):
        a[name][1 : nk + 1, ng : ng + nj, ng : ng + ni] += (3.0 * rng.standard_normal((nk, nj, ni))).astype(f32)
    a["u3d"][:, :, ng + ni] = a["u3d"][:, :, ng]
    a["v3d"][:, ng + nj, :] = a["v3d"][:, ng, :]
    sides = (1, 1, 1, 1)
    for name, kind in (("u3d", bc.KIND_U), ("v3d", bc.KIND_V), ("w3d", bc.KIND_W)):
        bc.fill(a[name], kind, sides, ni=ni, nj=nj, nk=nk)
    for nm, src in (("ua", "u3d"), ("va", "v3d"), ("wa", "w3d"), ("tha", "th3d")):
        a[nm] = a[src].copy()
    for name in ("pp3d", "ppx"):
        bc.bcp(a[name], sides, ni=ni, nj=nj, nk=nk)
    a["ppi"] = a["pp3d"].copy()
    prof = osd.terrain_coordinate_profiles(nk, cfg.dz)
    c = {"ni": ni, "nj": nj, "nk": nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": sides}
    return cfg, host, a, prof, c


def unpad_np(a, ni, nj, nk):
    from gm1_amd.adjoint_step import from_padded

    return {n: np.ascontiguousarray(v) for n, v in from_padded(a, ni, nj, nk).items()}


@pytest.mark.gpu
def test_gpu_halo_free_step_equals_the_restatement(device):
    import torch

    from gm1_amd.adjoint_step import DryPeriodicStep
    from oracle import large_step as ol

    cfg, host, a, prof, c = state32()
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    S = unpad_np(a, ni, nj, nk)
    a1 = {k: v.copy() for k, v in a.items()}
    ol.large_step(a1, host["metrics"], host["base"], prof, c, zt=np.float32(nk * cfg.dz))
    ref = unpad_np(a1, ni, nj, nk)
    step = DryPeriodicStep(cfg, host["metrics"], host["base"], device)
    out = step.forward({n: torch.from_numpy(v).to(device) for n, v in S.items()})
    torch.cuda.synchronize()
    for n in FIELDS:
        got = out[n].cpu().numpy()
        diff = np.abs(got.astype(np.float64) - ref[n]).max() / max(np.abs(ref[n]).max(), 1e-30)
        print(n, f"{diff:.2e}", int(np.sum(got != ref[n])))
    for n in FIELDS:
        assert np.array_equal(out[n].cpu().numpy(), ref[n]), n


def _oracle64_trajectory(a, host, cfg, prof, c):
    from test_binary64 import oracle_binary64, promote

    o64 = oracle_binary64()
    als = o64.adjoint_large_step
    p = als.Params(promote(host["metrics"]), promote(host["base"]), prof, c)
    return o64, als, p


def _rel(g, r):
    return float(np.abs(g.astype(np.float64) - r).max() / max(np.abs(r).max(), 1e-300))


@pytest.mark.gpu
@pytest.mark.parametrize("nrk,weno", [(1, "exact"), (2, "exact"), (3, "exact"), (3, "frozen")])
def test_gpu_stage_adjoint_matches_the_binary64_stage_adjoint(device, nrk, weno):
    """One RK stage's device adjoint against oracle rk_stage_T at the same (binary32) stage input."""
    import torch

    from gm1_amd.adjoint_step import DryPeriodicStep

    cfg, host, a, prof, c = state32()
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    o64, als, p = _oracle64_trajectory(a, host, cfg, prof, c)
    S = {n: v.astype(np.float64) for n, v in unpad_np(a, ni, nj, nk).items()}
    _, T = als.Step(p).forward(S)
    # the stage input of this stage, rounded to binary32 and used by both sides
    X = {n: T.traj[f"s{nrk}"]["X"][n].astype(np.float32) for n in ("u", "v", "w", "th", "pp")}
    ppx = T.traj[f"s{nrk}"]["ppx"].astype(np.float32)
    A = {n: S[n].astype(np.float32) for n in ("u", "v", "w", "th", "pp")}
    T2 = als.Tape(exact=weno == "exact")
    X64 = {n: v.astype(np.float64) for n, v in X.items()}
    A64 = {n: v.astype(np.float64) for n, v in A.items()}
    T2.traj["s1"] = {"X": A64}
    Xn, ppxn = als.rk_stage(p, T2, nrk, X64, A64, ppx.astype(np.float64))
    rng = np.random.default_rng(20 + nrk)
    Xb = {n: rng.standard_normal(Xn[n].shape) for n in Xn}
    ppxb = rng.standard_normal(ppx.shape)
    # the oracle's stage 3 leaves tend in ppx; solve3 (ppx = pp + tend) is folded into the device stage
    Ab = {n: np.zeros_like(A64[n]) for n in A64}
    Xb_o = dict(Xb)
    if nrk == 3:
        Xb_o["pp"] = Xb["pp"] + ppxb
    ref_X, ref_ppx = als.rk_stage_T(p, T2, nrk, A64, Xb_o, ppxb, Ab)
    step = DryPeriodicStep(cfg, host["metrics"], host["base"], device, weno=weno)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)  # noqa: E731
    step.zero_start_adjoint()
    got_X, got_ppx = step.adjoint_stage(nrk, {n: dev(v) for n, v in X.items()}, {n: dev(v) for n, v in A.items()},
                                        dev(ppx), {n: dev(v) for n, v in Xb.items()}, dev(ppxb))
    got_A = step.start_adjoint()
    torch.cuda.synchronize()
    errs = {f"X.{n}": _rel(got_X[n].cpu().numpy(), ref_X[n]) for n in ref_X}
    errs["ppx"] = _rel(got_ppx.cpu().numpy(), ref_ppx)
    for n in Ab:
        if np.abs(Ab[n]).max() > 0:
            errs[f"A.{n}"] = _rel(got_A[n].cpu().numpy(), Ab[n])
    print(f"nrk={nrk} weno={weno}", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v <= 1e-4, (k, v)


@pytest.mark.gpu
def test_gpu_step_adjoint_matches_the_binary64_adjoint(device):
    import torch

    from gm1_amd.adjoint_step import DryPeriodicStep

    cfg, host, a, prof, c = state32()
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    o64, als, p = _oracle64_trajectory(a, host, cfg, prof, c)
    S32 = unpad_np(a, ni, nj, nk)
    S = {n: v.astype(np.float64) for n, v in S32.items()}
    out, T = als.Step(p).forward(S)
    rng = np.random.default_rng(9)
    Sb = {n: rng.standard_normal(out[n].shape) for n in FIELDS}
    ref = als.Step(p).adjoint(T, Sb)
    step = DryPeriodicStep(cfg, host["metrics"], host["base"], device)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)  # noqa: E731
    got = step.adjoint({n: dev(v) for n, v in S32.items()}, {n: dev(v) for n, v in Sb.items()})
    torch.cuda.synchronize()
    errs = {n: _rel(got[n].cpu().numpy(), ref[n]) for n in FIELDS}
    print({k: f"{v:.1e}" for k, v in errs.items()})
    for n, e in errs.items():
        assert e <= 1e-4, (n, e)


def test_revolve_schedule_reaches_the_binomial_bound():
    """The split table's forward-step count equals Griewank's minimum t*n - C(s+t, s+1) (s free slots + the
    held start = s+1 snapshots, t the smallest repetition number with C(s+1+t, t) >= n)."""
    from math import comb

    from gm1_amd.adjoint_step import revolve_schedule

    for n, s in ((120, 8), (120, 12), (37, 3), (10, 2), (64, 5)):
        snaps = s + 1
        t = 0
        while comb(snaps + t, t) < n:
            t += 1
        bound = t * n - comb(snaps + t, t - 1) if t > 0 else 0
        _, cost = revolve_schedule(n, s)
        assert cost == bound, (n, s, cost, bound)


def _window_case(device, nsteps, snapshots):
    import torch

    from gm1_amd.adjoint_step import DryPeriodicStep, WindowGradient

    cfg, host, a, prof, c = state32()
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    S32 = unpad_np(a, ni, nj, nk)
    step = DryPeriodicStep(cfg, host["metrics"], host["base"], device)
    win = WindowGradient(step, nsteps, snapshots, zlo=500.0, zhi=2000.0, metrics=host["metrics"])
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)  # noqa: E731
    return cfg, host, a, prof, c, S32, step, win, {n: dev(v) for n, v in S32.items()}


def _fd_directions(rng, S32, g, device, count=8, pool=16):
    """``gm1_amd.adjoint_step.fd_directions``, the replaced draws printed."""
    from gm1_amd.adjoint_step import fd_directions

    kept, replaced = fd_directions(rng, S32, g, device, count, pool)
    for k, ad in replaced:
        print(f"draw {k + 1}: g.d = {ad:.3e} nearly orthogonal to the gradient, replaced")
    return kept


@pytest.mark.gpu
def test_gpu_window_gradient_against_the_binary64_gradient(device):
    """dJ/d(theta', u, v) after 3 steps: device (binary32, checkpointed) against the binary64 adjoint."""
    import torch

    nsteps = 3
    cfg, host, a, prof, c, S32, step, win, S0 = _window_case(device, nsteps, 1)
    J, g = win.gradient(S0)
    torch.cuda.synchronize()
    o64, als, p = _oracle64_trajectory(a, host, cfg, prof, c)
    S = {n: v.astype(np.float64) for n, v in S32.items()}
    tapes, states = [], [S]
    for _ in range(nsteps):
        out, T = als.Step(p).forward(states[-1])
        tapes.append(T)
        states.append(out)
    J64, wb = als.storm_intensity(p, states[-1]["w"], zlo=500.0, zhi=2000.0)
    Sb = {n: np.zeros_like(states[-1][n]) for n in FIELDS}
    Sb["w"] = wb
    for T in reversed(tapes):
        Sb = als.Step(p).adjoint(T, Sb)
    print("J", float(J), J64, "recomputed steps", win.forward_steps)
    assert abs(float(J) - J64) <= 1e-5 * abs(J64)
    for n in ("th", "u", "v", "w", "pp"):
        e = _rel(g[n].cpu().numpy(), Sb[n])
        print(n, f"{e:.2e}")
        assert e <= 1e-4, (n, e)


@pytest.mark.gpu
def test_gpu_checkpoint_schedules_agree(device):
    import torch

    cfg, host, a, prof, c, S32, step, win0, S0 = _window_case(device, 5, 0)
    J0, g0 = win0.gradient(S0)
    from gm1_amd.adjoint_step import WindowGradient

    win3 = WindowGradient(step, 5, 3, zlo=500.0, zhi=2000.0, metrics=host["metrics"])
    J3, g3 = win3.gradient(S0)
    torch.cuda.synchronize()
    print("forward steps", win0.forward_steps, win3.forward_steps)
    assert win0.forward_steps > win3.forward_steps
    assert float(J0) == float(J3)
    for n in ("th", "u", "v"):
        e = _rel(g3[n].cpu().numpy(), g0[n].cpu().numpy().astype(np.float64))
        assert e <= 1e-5, (n, e)


@pytest.mark.gpu
def test_gpu_gradient_central_differences_and_autograd(device):
    """Central differences (eps = 1e-3 of each field's scale) along 8 smooth random directions, and autograd.

    This small state carries white noise of 0.5 m/s in every wind, so many
    faces sit within eps of an upwind switch of the 5th-order fluxes, where
    J is not differentiable; the binary32 central differences agree to about
    3e-3 here, held to 5e-3 (the 192x192x64 window below is held to 1e-3).  Directions nearly
    orthogonal to the gradient are replaced (``_fd_directions``).
    """
    import torch

    from gm1_amd.adjoint_step import StormIntensity

    cfg, host, a, prof, c, S32, step, win, S0 = _window_case(device, 3, 2)
    J, g = win.gradient(S0)
    worst = 0.0
    for k, (d, ad) in enumerate(_fd_directions(np.random.default_rng(31), S32, g, device)):
        eps = 1e-3
        Sp = {**S0, **{n: S0[n] + eps * d[n] for n in d}}
        Sm = {**S0, **{n: S0[n] - eps * d[n] for n in d}}
        Jp = torch.sum(win.weights.double() * win.advance(Sp, 3)["w"].double() ** 2)
        Jm = torch.sum(win.weights.double() * win.advance(Sm, 3)["w"].double() ** 2)
        fd = float((Jp - Jm) / (2 * eps))
        rel = abs(fd - ad) / abs(fd)
        worst = max(worst, rel)
        print(f"direction {k + 1}: fd={fd:.8e} adjoint={ad:.8e} rel={rel:.2e}")
    assert worst <= 5e-3, worst
    base = {n: S0[n] for n in ("w", "pp", "ppx")}
    th, u, v = (S0[n].clone().requires_grad_(True) for n in ("th", "u", "v"))
    Jt = StormIntensity.apply(win, base, th, u, v)
    gth, gu, gv = torch.autograd.grad(Jt, (th, u, v))
    assert float(Jt.detach()) == float(J)
    for got, ref in ((gth, g["th"]), (gu, g["u"]), (gv, g["v"])):
        assert _rel(got.cpu().numpy(), ref.cpu().numpy().astype(np.float64)) <= 1e-5


@pytest.mark.gpu
def test_gpu_storm_intensity_gradient_over_120_steps_central_differences(device):
    """The 192x192x64 dry periodic case: the 120-step window's gradient (16 free checkpoint slots) against
    central differences of the device forward, eps = 1e-3 of each field's scale, along 8 smooth random
    directions of (theta', u, v), each within 1e-3 (directions nearly orthogonal to the gradient replaced
    as ``_fd_directions`` states)."""
    import torch

    from gm1_amd.adjoint_step import DryPeriodicStep, WindowGradient, from_padded
    from gm1_amd.fields import adjoint_c5

    cfg = adjoint_c5()
    print("case", cfg.name, "window 120", flush=True)
    host = host_initial_state(cfg)
    S0 = {n: v.contiguous().clone() for n, v in from_padded(
        {k: torch.from_numpy(np.ascontiguousarray(host["arrays"][k])).to(device)
         for k in ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx")}, cfg.ni, cfg.nj, cfg.nk).items()}
    step = DryPeriodicStep(cfg, host["metrics"], host["base"], device)
    win = WindowGradient(step, 120, 16, metrics=host["metrics"])
    J, g = win.gradient(S0)
    torch.cuda.synchronize()
    print("J", float(J), "forward steps", win.forward_steps, flush=True)
    S32 = {n: S0[n].cpu().numpy() for n in ("th", "u", "v")}
    worst = 0.0
    for k, (d, ad) in enumerate(_fd_directions(np.random.default_rng(11), S32, g, device)):
        eps = 1e-3
        Jp = torch.sum(win.weights.double() * win.advance({**S0, **{n: S0[n] + eps * d[n] for n in d}},
                                                          120)["w"].double() ** 2)
        Jm = torch.sum(win.weights.double() * win.advance({**S0, **{n: S0[n] - eps * d[n] for n in d}},
                                                          120)["w"].double() ** 2)
        fd = float((Jp - Jm) / (2 * eps))
        rel = abs(fd - ad) / abs(fd)
        worst = max(worst, rel)
        print(f"direction {k + 1}: fd={fd:.8e} adjoint={ad:.8e} rel={rel:.2e}", flush=True)
    assert worst <= 1e-3, worst
