"""Adjoint of the acoustic RK stage on the C5 geometry (dry, flat, uniform, periodic).

CPU: the binary64 adjoint (oracle/adjoint_sound.py) against the binary64
restatement of the forward stage by the dot-product identity
<ybar, F(x + d) - F(x)> = <F^T ybar, d>: for fixed coefficients the stage is
linear in its state and forcing, so the identity holds to binary64 rounding
for any d.  Each RK stage (nrk = 1, 2, 3) is checked with random output
adjoints over every word of every output (halos included) and random
perturbations of the stage's inputs.

GPU: the binary32 device adjoint (gm1hip_psolver3_stage_adjoint) against the
binary64 adjoint (relative max-norm error per input adjoint <= 1e-5), by the
dot-product identity against the device forward stage, and through
``PeriodicAcousticStage`` (torch.autograd) against the direct launch.
"""

from __future__ import annotations

import numpy as np
import pytest

from gm1_amd.adjoint_sound import INPUTS, STATE, PeriodicAcousticStage, launch_cuda_psolver3_stage_adjoint
from gm1_amd.errors import CM1Refusal
from gm1_amd.fields import CaseConfig
from gm1_amd.large_step import host_initial_state
from helpers import to_dev, to_host
from oracle import adjoint_sound as oad
from oracle import stage as ost
from test_binary64 import oracle_binary64, promote

PERIODIC = (1, 1, 1, 1)


def c5_like(ni=14, nj=12, nk=10, seed=11):
    """C5's geometry and state family at a small size: dx 500 m, dz 250 m, dt 3 s, 6 substeps."""
    return CaseConfig(f"c5-{ni}x{nj}x{nk}", ni, nj, nk, 500.0, 500.0, 250.0, 3.0, 6, PERIODIC, False, "linear15",
                      ((2.0, ni * 250.0, nj * 250.0, 1400.0),), bubble_radius=(ni * 150.0, 1400.0), noise=0.1,
                      seed=seed)


def stage_case(nrk, seed=0):
    """A binary64 state after the stage terms, with nonzero winds, pressure and forcing everywhere."""
    o64 = oracle_binary64()
    cfg = c5_like()
    host = host_initial_state(cfg)
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    a = promote({k: v.copy() for k, v in host["arrays"].items()})
    m = promote(host["metrics"])
    base = promote(host["base"])
    rng = np.random.default_rng(100 + nrk + seed)
    for name, scale in (("u3d", 2.0), ("v3d", 2.0), ("w3d", 0.5), ("ua", 2.0), ("va", 2.0), ("wa", 0.5),
                        ("uten", 1e-3), ("vten", 1e-3), ("wten", 1e-3), ("ppten", 1e-6), ("ppx", 1e-4),
                        ("ppi", 1e-4), ("pp3d", 1e-4)):
        a[name] = a[name] + scale * rng.standard_normal(a[name].shape)
    o64.stage.stage_terms(a, base, a["th3d"], None, ni=ni, nj=nj, nk=nk, sides=PERIODIC)
    prof = o64.sound.terrain_coordinate_profiles(nk, cfg.dz)
    return o64, cfg, a, m, prof


def run_stage(o64, cfg, a, m, prof, nrk):
    b = {k: v.copy() for k, v in a.items()}
    o64.sound.cpu_psolver3_stage_inplace(b, m, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, dt=cfg.dt, nsound=cfg.nsound,
                                         nrk=nrk, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, lateral_sides=PERIODIC,
                                         prof=prof, zt=np.float64(cfg.nk * cfg.dz), radiation=False)
    return b


def inputs_of(nrk):
    if nrk == 1:
        return ("u3d", "v3d", "w3d", "pp3d", "ppx", "wa", "uten", "vten", "wten", "ppten")
    return ("ua", "va", "wa", "ppi", "ppx", "uten", "vten", "wten", "ppten")


@pytest.mark.parametrize("nrk", [1, 2, 3])
def test_stage_adjoint_dot_product_binary64(nrk):
    o64, cfg, a, m, prof = stage_case(nrk)
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    rng = np.random.default_rng(7 + nrk)
    y0 = run_stage(o64, cfg, a, m, prof, nrk)
    d = {n: rng.standard_normal(a[n].shape) * np.abs(a[n]).max() * 1e-3 for n in inputs_of(nrk)}
    a1 = {k: v.copy() for k, v in a.items()}
    for n, dv in d.items():
        a1[n] = a1[n] + dv
    y1 = run_stage(o64, cfg, a1, m, prof, nrk)
    ybar = {n: rng.standard_normal(y0[n].shape) for n in o64.adjoint_sound.STATE}
    lhs = sum(float(np.sum(ybar[n] * (y1[n] - y0[n]))) for n in ybar)
    coef = {n: a[n] for n in ("thv", "ppterm", "piadv")}
    xbar = o64.adjoint_sound.stage_adjoint(ybar, coef, m, ni=ni, nj=nj, nk=nk, dt=cfg.dt, nsound=cfg.nsound,
                                           nrk=nrk, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, prof=prof)
    rhs = sum(float(np.sum(xbar[n] * d[n])) for n in d)
    scale = sum(float(np.sum(np.abs(ybar[n] * (y1[n] - y0[n])))) for n in ybar)
    print(f"nrk={nrk} lhs={lhs:.12e} rhs={rhs:.12e} scale={scale:.3e}")
    assert abs(lhs - rhs) <= 1e-9 * scale, (lhs, rhs, scale)


def test_adjoint_refuses_open_sides():
    with pytest.raises(CM1Refusal):
        launch_cuda_psolver3_stage_adjoint({}, {}, {}, ni=8, nj=8, nk=8, nrk=1, dt=3.0, nsound=6, dx=500.0,
                                           dy=500.0, dz=250.0, lateral_sides=(2, 2, 2, 2))


# ------------------------------------------------------------------ GPU

def stage_case32(nrk, seed=0):
    """The binary32 counterpart of ``stage_case``."""
    cfg = c5_like()
    host = host_initial_state(cfg)
    a = {k: v.copy() for k, v in host["arrays"].items()}
    rng = np.random.default_rng(100 + nrk + seed)
    for name, scale in (("u3d", 2.0), ("v3d", 2.0), ("w3d", 0.5), ("ua", 2.0), ("va", 2.0), ("wa", 0.5),
                        ("uten", 1e-3), ("vten", 1e-3), ("wten", 1e-3), ("ppten", 1e-6), ("ppx", 1e-4),
                        ("ppi", 1e-4), ("pp3d", 1e-4)):
        a[name] = (a[name] + scale * rng.standard_normal(a[name].shape)).astype(np.float32)
    ost.stage_terms(a, host["base"], a["th3d"], None, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, sides=PERIODIC)
    return cfg, a, host["metrics"]


def kw_of(cfg, nrk):
    return {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "nrk": nrk, "dt": cfg.dt, "nsound": cfg.nsound, "dx": cfg.dx,
            "dy": cfg.dy, "dz": cfg.dz}


def forward_names():
    from gm1_amd.psolver3 import BOUNDARY_NAMES, PRESSURE_NAMES, U_NAMES, V_NAMES, W_NAMES

    return (*PRESSURE_NAMES, *U_NAMES, *V_NAMES, *W_NAMES, *BOUNDARY_NAMES)


def gpu_stage(device, cfg, a, m, nrk):
    from gm1_amd.psolver3 import launch_cuda_psolver3_terrain_open_stage

    da = to_dev({n: a[n] for n in forward_names()}, device)
    launch_cuda_psolver3_terrain_open_stage(da, to_dev(m, device), lateral_sides=PERIODIC, **kw_of(cfg, nrk))
    return to_host({n: da[n] for n in STATE})


def output_adjoints(a, seed):
    rng = np.random.default_rng(seed)
    return {n: rng.standard_normal(a[n].shape).astype(np.float32) for n in STATE}


@pytest.mark.gpu
@pytest.mark.parametrize("nrk", [1, 2, 3])
def test_gpu_stage_adjoint_matches_the_binary64_adjoint(device, nrk):
    cfg, a, m = stage_case32(nrk)
    ybar = output_adjoints(a, 40 + nrk)
    coef = {n: a[n] for n in ("thv", "ppterm", "piadv")}
    ref = oad.stage_adjoint(ybar, coef, m, ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, dt=cfg.dt, nsound=cfg.nsound, nrk=nrk,
                            dx=cfg.dx, dy=cfg.dy, dz=cfg.dz)
    got, ready = launch_cuda_psolver3_stage_adjoint(to_dev(ybar, device), to_dev(coef, device), to_dev(m, device),
                                                    **kw_of(cfg, nrk))
    ready.synchronize()
    got = to_host(got)
    assert set(got) == set(ref) - ({"ua", "va", "ppi"} if nrk == 1 else set())
    errs = {}
    for n, g in got.items():
        scale = float(np.abs(ref[n]).max())
        assert scale > 0.0, n
        errs[n] = float(np.abs(g.astype(np.float64) - ref[n]).max()) / scale
    print(f"nrk={nrk}", {n: f"{e:.1e}" for n, e in errs.items()})
    for n, e in errs.items():
        assert e <= 1e-5, (n, e)


@pytest.mark.gpu
@pytest.mark.parametrize("nrk", [1, 2, 3])
def test_gpu_stage_adjoint_dot_product_against_the_gpu_stage(device, nrk):
    cfg, a, m = stage_case32(nrk)
    rng = np.random.default_rng(60 + nrk)
    d = {n: (rng.standard_normal(a[n].shape) * np.abs(a[n]).max() * 0.5).astype(np.float32) for n in inputs_of(nrk)}
    a1 = dict(a)
    for n, dv in d.items():
        a1[n] = a[n] + dv
    y0 = gpu_stage(device, cfg, a, m, nrk)
    y1 = gpu_stage(device, cfg, a1, m, nrk)
    ybar = output_adjoints(a, 80 + nrk)
    dy = {n: y1[n].astype(np.float64) - y0[n] for n in STATE}
    lhs = sum(float(np.sum(ybar[n] * dy[n])) for n in STATE)
    scale = sum(float(np.sum(np.abs(ybar[n] * dy[n]))) for n in STATE)
    coef = {n: a[n] for n in ("thv", "ppterm", "piadv")}
    xbar, ready = launch_cuda_psolver3_stage_adjoint(to_dev(ybar, device), to_dev(coef, device), to_dev(m, device),
                                                     **kw_of(cfg, nrk))
    ready.synchronize()
    xbar = to_host(xbar)
    rhs = sum(float(np.sum(xbar[n].astype(np.float64) * d[n])) for n in d)
    print(f"nrk={nrk} lhs={lhs:.9e} rhs={rhs:.9e} rel={abs(lhs - rhs) / scale:.1e}")
    assert abs(lhs - rhs) <= 1e-5 * scale, (lhs, rhs, scale)


@pytest.mark.gpu
@pytest.mark.parametrize("nrk", [1, 3])
def test_gpu_autograd_stage_matches_the_direct_adjoint(device, nrk):
    import torch

    cfg, a, m = stage_case32(nrk)
    dm = to_dev(m, device)
    fixed = to_dev({n: a[n] for n in forward_names() if n not in INPUTS}, device)
    stage = {"arrays": fixed, "metrics": dm, **kw_of(cfg, nrk)}
    xs = [to_dev(a[n], device).requires_grad_(True) for n in INPUTS]
    ys = PeriodicAcousticStage.apply(stage, *xs)
    # the forward is the device stage
    want = gpu_stage(device, cfg, a, m, nrk)
    for n, y in zip(STATE, ys):
        assert np.array_equal(y.detach().cpu().numpy(), want[n]), n
    ybar = to_dev(output_adjoints(a, 90 + nrk), device)
    loss = sum((ybar[n] * y).sum() for n, y in zip(STATE, ys))
    grads = torch.autograd.grad(loss, xs)
    direct, ready = launch_cuda_psolver3_stage_adjoint(ybar, {n: fixed[n] for n in ("thv", "ppterm", "piadv")}, dm,
                                                       **kw_of(cfg, nrk))
    ready.synchronize()
    for n, g in zip(INPUTS, grads):
        if n in direct:
            assert torch.equal(g, direct[n]), n
        else:
            assert not bool(torch.any(g != 0)), n
    # the inputs were not written
    for n, x in zip(INPUTS, xs):
        assert np.array_equal(x.detach().cpu().numpy(), a[n]), n
