"""Momentum advection5: oracle pins and GPU parity.

Reference behaviour: ``cuda_momentum_advection5.py`` (CPU :753-1011) and its
tests ``tests/test_cuda_momentum_advection5.py`` (constant flow gives zero
increments, normal open faces keep their seed).
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, NG, assert_bits_equal, golden, seeded_metrics, shapes, to_dev, to_host
from oracle import momentum5 as om


def _golden_case(g, c):
    ni, nj, nk = (int(v) for v in g[f"c{c}_dims"])
    sides = tuple(int(v) for v in g[f"c{c}_sides"])
    a = {k[len(f"c{c}_in_"):]: g[k].copy() for k in g.files if k.startswith(f"c{c}_in_")}
    m = {k[len(f"c{c}_metric_"):]: g[k] for k in g.files if k.startswith(f"c{c}_metric_")}
    return ni, nj, nk, sides, a, m


@pytest.mark.parametrize("case", [0, 1, 2])
def test_oracle_matches_reference_golden_vectors(case):
    g = golden("momentum5")
    ni, nj, nk, sides, a, m = _golden_case(g, case)
    divx = np.zeros((nk, nj, ni), F)
    om.cpu_terrain_open_momentum_advection5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, nrk=2, dx=250.0, dy=200.0,
                                                    dz=150.0, lateral_sides=sides)
    for k in ("rru", "rrv", "rrw", "uten", "vten", "wten"):
        assert_bits_equal(a[k], g[f"c{case}_out_{k}"], f"{k} sides={sides}")
    assert_bits_equal(divx, g[f"c{case}_out_divx"], "divx")


def _case(rng, ni, nj, nk, constant=False):
    sh = shapes(ni, nj, nk)
    m = seeded_metrics(rng, ni, nj, nk, terrain=not constant)
    if constant:
        for name in ("rho0_3d", "rr0", "rf0", "rrf0", "c1", "c2", "gz", "gzu", "gzv", "rgzu", "rgzv"):
            m[name][:] = F(1.0) if name not in ("c1", "c2") else F(0.5)
        a = {"u3d": np.full(sh["u"], F(4.0)), "v3d": np.full(sh["v"], F(-3.0)), "w3d": np.zeros(sh["w"], F)}
    else:
        a = {"u3d": (3 * rng.standard_normal(sh["u"])).astype(F), "v3d": (3 * rng.standard_normal(sh["v"])).astype(F),
             "w3d": rng.standard_normal(sh["w"]).astype(F)}
    a.update({
        "uten": np.zeros(sh["u"], F) if constant else (0.01 * rng.standard_normal(sh["u"])).astype(F),
        "vten": np.zeros(sh["v"], F) if constant else (0.01 * rng.standard_normal(sh["v"])).astype(F),
        "wten": np.zeros(sh["w"], F) if constant else (0.01 * rng.standard_normal(sh["w"])).astype(F),
        "rru": np.zeros(sh["u"], F), "rrv": np.zeros(sh["v"], F), "rrw": np.zeros(sh["w"], F),
        "uh": np.ones(ni + 2 * NG, F), "vh": np.ones(nj + 2 * NG, F),
    })
    return a, m


@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_constant_flat_flow_has_zero_divergence_and_increment(sides):
    ni, nj, nk = 12, 10, 9
    a, m = _case(np.random.default_rng(1), ni, nj, nk, constant=True)
    divx = np.ones((nk, nj, ni), F)
    om.cpu_terrain_open_momentum_advection5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, nrk=1, dx=100.0, dy=100.0,
                                                    dz=100.0, lateral_sides=sides)
    assert not np.count_nonzero(divx)
    for k in ("uten", "vten", "wten", "rrw"):
        assert not np.count_nonzero(a[k]), k


def test_open_normal_faces_keep_their_seed():
    ni, nj, nk = 11, 9, 8
    a, m = _case(np.random.default_rng(2), ni, nj, nk)
    useed, vseed = a["uten"].copy(), a["vten"].copy()
    divx = np.zeros((nk, nj, ni), F)
    om.cpu_terrain_open_momentum_advection5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, nrk=1, dx=100.0, dy=100.0,
                                                    dz=100.0)
    assert np.array_equal(a["uten"][:, :, NG], useed[:, :, NG])
    assert np.array_equal(a["uten"][:, :, NG + ni], useed[:, :, NG + ni])
    assert np.array_equal(a["vten"][:, NG, :], vseed[:, NG, :])
    assert np.array_equal(a["vten"][:, NG + nj, :], vseed[:, NG + nj, :])
    assert np.count_nonzero(a["uten"] != useed)


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1), (1, 1, 2, 3), (4, 3, 1, 1)])
@pytest.mark.parametrize("dims", [(10, 9, 8), (67, 41, 17)])
def test_gpu_momentum_advection_bit_exact(device, sides, dims):
    from gm1_amd.momentum_advection5 import launch_cuda_terrain_open_momentum_advection5

    ni, nj, nk = dims
    a, m = _case(np.random.default_rng(ni * 31 + nk), ni, nj, nk)
    da, dm = to_dev(a, device), to_dev(m, device)
    divx = np.zeros((nk, nj, ni), F)
    om.cpu_terrain_open_momentum_advection5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, nrk=1, dx=250.0, dy=200.0,
                                                    dz=150.0, lateral_sides=sides)
    launch = launch_cuda_terrain_open_momentum_advection5(da, dm, ni=ni, nj=nj, nk=nk, nrk=1, dx=250.0, dy=200.0,
                                                          dz=150.0, lateral_sides=sides)
    for k in ("rru", "rrv", "rrw", "uten", "vten", "wten"):
        assert_bits_equal(to_host(da[k]), a[k], f"{k} sides={sides}")
    assert_bits_equal(to_host(launch.divx), divx, "divx")


@pytest.mark.gpu
def test_gpu_hook_split_matches_fused(device):
    from gm1_amd.momentum_advection5 import launch_cuda_terrain_open_momentum_advection5

    ni, nj, nk = 33, 21, 12
    a, m = _case(np.random.default_rng(9), ni, nj, nk)
    d1, d2, dm = to_dev(a, device), to_dev(a, device), to_dev(m, device)
    calls = []

    def hook(arrays, **kw):
        calls.append(kw["nrk"])
        return {"performed": True}

    l1 = launch_cuda_terrain_open_momentum_advection5(d1, dm, ni=ni, nj=nj, nk=nk, nrk=2, dx=100.0, dy=100.0, dz=100.0)
    l2 = launch_cuda_terrain_open_momentum_advection5(d2, dm, ni=ni, nj=nj, nk=nk, nrk=2, dx=100.0, dy=100.0, dz=100.0,
                                                      post_divergence_hook=hook)
    assert calls == [2]
    for k in ("uten", "vten", "wten"):
        assert_bits_equal(to_host(d2[k]), to_host(d1[k]), k)
    assert_bits_equal(to_host(l2.divx), to_host(l1.divx), "divx")


@pytest.mark.gpu
def test_gpu_golden_vectors_through_c_abi(device):
    from gm1_amd.momentum_advection5 import launch_cuda_terrain_open_momentum_advection5

    g = golden("momentum5")
    for case in range(3):
        ni, nj, nk, sides, a, m = _golden_case(g, case)
        full = seeded_metrics(np.random.default_rng(0), ni, nj, nk)
        full.update(m)
        da = to_dev(a, device)
        launch = launch_cuda_terrain_open_momentum_advection5(da, to_dev(full, device), ni=ni, nj=nj, nk=nk, nrk=2,
                                                              dx=250.0, dy=200.0, dz=150.0, lateral_sides=sides)
        for k in ("rru", "rrv", "rrw", "uten", "vten", "wten"):
            assert_bits_equal(to_host(da[k]), g[f"c{case}_out_{k}"], f"{k} case={case}")
        assert_bits_equal(to_host(launch.divx), g[f"c{case}_out_divx"], "divx")
