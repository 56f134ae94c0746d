"""Scalar advection5 and terrain mass flux: oracle pins and GPU parity.

Reference behaviour: ``cuda_advection5.py`` (CPU transcription :270-579,
kernels :603-670) and its tests ``tests/test_cuda_advection5.py``.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import (
    F, NG, assert_bits_equal, golden, seeded_metrics, shapes, to_dev, to_host,
)
from oracle import advection5 as oa
from oracle.common import flx5, terrain_static_shapes, weno5


# ---------------------------------------------------------------- CPU: oracle

def _golden_case(g, c):
    ni, nj, nk = (int(v) for v in g[f"c{c}_dims"])
    a = {k: g[f"c{c}_in_{k}"].copy() for k in ("u3d", "v3d", "w3d", "rru", "rrv", "rrw")}
    m = {k[len(f"c{c}_metric_"):]: g[k] for k in g.files if k.startswith(f"c{c}_metric_")}
    return ni, nj, nk, a, m


@pytest.mark.parametrize("case", [0, 1])
def test_oracle_matches_reference_golden_vectors(case):
    g = golden("advection5")
    ni, nj, nk, a, m = _golden_case(g, case)
    divx = np.zeros((nk, nj, ni), F)
    oa.cpu_terrain_mass_flux5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, dx=250.0, dy=200.0, dz=150.0)
    for k in ("rru", "rrv", "rrw"):
        assert_bits_equal(a[k], g[f"c{case}_out_{k}"], k)
    assert_bits_equal(divx, g[f"c{case}_out_divx"], "divx")
    for nrk in (1, 3):
        fl = {"x": np.zeros_like(a["rru"]), "y": np.zeros_like(a["rrv"]), "z": np.zeros_like(a["rrw"])}
        t = np.zeros((nk, nj, ni), F)
        oa.cpu_scalar_advection5_inplace(
            g[f"c{case}_scalar"], a, m, divx, t, fl, ni=ni, nj=nj, nk=nk, nrk=nrk,
            dx=250.0, dy=200.0, dz=150.0, weps=1.0e-17)
        assert_bits_equal(t, g[f"c{case}_nrk{nrk}_tend"], f"tendency nrk={nrk}")
        for d in "xyz":
            assert_bits_equal(fl[d], g[f"c{case}_nrk{nrk}_flux_{d}"], f"flux {d} nrk={nrk}")


def _reference_fixture():
    """The 8x8x8 fixture of tests/test_cuda_advection5.py:36-75 (rebuilt as data)."""
    ni = nj = nk = 8
    sh = shapes(ni, nj, nk)
    arrays = {n: np.zeros(sh[s], F) for n, s in (("u3d", "u"), ("v3d", "v"), ("w3d", "w"), ("rru", "u"), ("rrv", "v"), ("rrw", "w"))}
    for name in ("u3d", "v3d", "w3d"):
        z, y, x = np.indices(arrays[name].shape)
        arrays[name][:] = (F(2.0) + F(0.01) * x.astype(F) - F(0.02) * y.astype(F) + F(0.03) * z.astype(F))
    metrics = {n: np.ones(s, F) for n, s in terrain_static_shapes(ni, nj, nk).items()}
    for k in range(nk):
        metrics["rho0_3d"][k] = F(1.1 - 0.01 * k)
        metrics["rr0"][k] = F(1.0) / metrics["rho0_3d"][k]
    for k in range(nk + 1):
        metrics["rf0"][k] = F(1.05 - 0.009 * k)
        metrics["c1"][k] = F(0.5)
        metrics["c2"][k] = F(0.5)
    y, x = np.indices((nj, ni))
    metrics["dzdx"][:] = F(0.001) * (x - 3).astype(F)
    metrics["dzdy"][:] = F(-0.001) * (y - 3).astype(F)
    metrics["gz"][:] = F(1.01)
    scalar = np.zeros(sh["s"], F)
    z, y, x = np.indices(scalar.shape)
    scalar[:] = (np.where(x < NG + 4, F(0.1), F(1.0)) + F(0.001) * z.astype(F) + F(0.002) * y.astype(F))
    return arrays, metrics, scalar


def _tend(scalar, arrays, metrics, divx, nrk, wflag=True):
    t = np.zeros((8, 8, 8), F)
    fl = {"x": np.zeros_like(arrays["rru"]), "y": np.zeros_like(arrays["rrv"]), "z": np.zeros_like(arrays["rrw"])}
    oa.cpu_scalar_advection5_inplace(scalar, arrays, metrics, divx, t, fl, ni=8, nj=8, nk=8, nrk=nrk,
                                     dx=100.0, dy=100.0, dz=200.0, weps=1.0e-17, wflag=wflag)
    return t


def test_reference_known_answer_weno_changes_exactly_32_words():
    """tests/test_cuda_advection5.py:170-198 known answers."""
    arrays, metrics, scalar = _reference_fixture()
    divx = np.zeros((8, 8, 8), F)
    oa.cpu_terrain_mass_flux5_inplace(arrays, metrics, divx, ni=8, nj=8, nk=8, dx=100.0, dy=100.0, dz=200.0)
    s1 = _tend(scalar, arrays, metrics, divx, 1)
    s2 = _tend(scalar, arrays, metrics, divx, 2)
    s3 = _tend(scalar, arrays, metrics, divx, 3)
    off = _tend(scalar, arrays, metrics, divx, 3, wflag=False)
    assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32))
    assert np.array_equal(s1.view(np.uint32), off.view(np.uint32))
    assert np.count_nonzero(s1)
    assert np.count_nonzero(s3.view(np.uint32) != s2.view(np.uint32)) == 32
    mask = np.ones((8, 8), bool)
    mask[3:5, 3:5] = False
    assert not np.count_nonzero(s1[:, mask])


def test_weno_selector():
    assert oa.scalar_advection_uses_weno(1) is False
    assert oa.scalar_advection_uses_weno(2) is False
    assert oa.scalar_advection_uses_weno(3) is True
    assert oa.scalar_advection_uses_weno(3, wflag=False) is False


def test_flx5_is_exact_for_quartic_cell_averages():
    """Closed form: the 5-point upwind-biased face value is exact for degree <= 4.

    The weights are binary32-rounded, so exactness holds to ~1e-7 relative;
    a wrong weight or a transposed stencil gives an O(1) error.
    """
    rng = np.random.default_rng(3)
    for _ in range(20):
        c = rng.standard_normal(5)
        prim = lambda x: sum(c[p] * x ** (p + 1) / (p + 1) for p in range(5))  # noqa: E731
        avg = [prim(m + 1) - prim(m) for m in range(-3, 2)]  # cells -3..1, face at x=0
        face = sum(c[p] * 0.0 ** p for p in range(5))
        got = flx5(*(np.float64(v) for v in avg))
        assert abs(float(got) - face) < 1e-6 * (1 + sum(abs(c)) * 10)


def test_weno_tends_to_flx5_on_smooth_data():
    x = np.linspace(0.0, 0.05, 5).astype(F)
    s = (F(1.0) + F(0.3) * x + F(0.1) * x * x).astype(F)
    a = flx5(*s)
    b = weno5(*s, 1.0e-17)
    assert abs(float(a) - float(b)) < 1e-5


def test_constant_scalar_in_nondivergent_flow_has_zero_tendency():
    ni, nj, nk = 16, 12, 10
    rng = np.random.default_rng(0)
    sh = shapes(ni, nj, nk)
    m = seeded_metrics(rng, ni, nj, nk, terrain=False)
    arrays = {"rru": np.full(sh["u"], F(2.0)), "rrv": np.full(sh["v"], F(-1.5)), "rrw": np.zeros(sh["w"], F)}
    divx = np.zeros((nk, nj, ni), F)
    scalar = np.full(sh["s"], F(300.0))
    fl = {"x": np.zeros(sh["u"], F), "y": np.zeros(sh["v"], F), "z": np.zeros(sh["w"], F)}
    for nrk in (1, 3):
        t = np.ones((nk, nj, ni), F)
        oa.cpu_scalar_advection5_inplace(scalar, arrays, m, divx, t, fl, ni=ni, nj=nj, nk=nk, nrk=nrk,
                                         dx=100.0, dy=100.0, dz=100.0, weps=1e-17, lateral_sides=(1, 1, 1, 1))
        assert np.all(t == 0.0)


# ---------------------------------------------------------------- GPU parity

def _random_case(rng, ni, nj, nk, terrain=True):
    sh = shapes(ni, nj, nk)
    m = seeded_metrics(rng, ni, nj, nk, terrain=terrain)
    a = {
        "u3d": (3.0 * rng.standard_normal(sh["u"])).astype(F),
        "v3d": (3.0 * rng.standard_normal(sh["v"])).astype(F),
        "w3d": (1.0 * rng.standard_normal(sh["w"])).astype(F),
        "rru": np.zeros(sh["u"], F), "rrv": np.zeros(sh["v"], F), "rrw": np.zeros(sh["w"], F),
    }
    scalar = (300.0 + 2.0 * rng.standard_normal(sh["s"])).astype(F)
    scalar[:, :, : NG + ni // 3] += F(4.0)
    return a, m, scalar


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(12, 10, 9), (70, 45, 20), (130, 9, 8)])
def test_gpu_terrain_mass_flux_bit_exact(device, dims):
    from gm1_amd.advection5 import launch_cuda_terrain_mass_flux5

    ni, nj, nk = dims
    rng = np.random.default_rng(sum(dims))
    a, m, _ = _random_case(rng, ni, nj, nk)
    divx = np.zeros((nk, nj, ni), F)
    da, dm, dd = to_dev(a, device), to_dev(m, device), to_dev(divx, device)
    oa.cpu_terrain_mass_flux5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, dx=250.0, dy=200.0, dz=150.0)
    launch_cuda_terrain_mass_flux5(da, dm, dd, ni=ni, nj=nj, nk=nk, dx=250.0, dy=200.0, dz=150.0)
    for k in ("rru", "rrv", "rrw"):
        assert_bits_equal(to_host(da[k]), a[k], k)
    assert_bits_equal(to_host(dd), divx, "divx")


@pytest.mark.gpu
@pytest.mark.parametrize("nrk", [1, 3])
@pytest.mark.parametrize("sides", [None, (2, 2, 2, 2), (1, 1, 1, 1), (2, 1, 3, 2)])
@pytest.mark.parametrize("dims", [(12, 10, 9), (70, 45, 20)])
def test_gpu_scalar_advection_bit_exact(device, nrk, sides, dims):
    import torch

    from gm1_amd.advection5 import launch_cuda_scalar_advection5

    ni, nj, nk = dims
    rng = np.random.default_rng(7 * ni + nj)
    a, m, scalar = _random_case(rng, ni, nj, nk)
    divx = np.zeros((nk, nj, ni), F)
    oa.cpu_terrain_mass_flux5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, dx=250.0, dy=200.0, dz=150.0)
    # fill the mass-flux edges the interior routine leaves, so every face is exercised
    for k in ("rru", "rrv", "rrw"):
        edge = a[k] == 0
        a[k][edge] = (rng.standard_normal(int(edge.sum())) * 2.0).astype(F)
    divx = (rng.standard_normal((nk, nj, ni)) * 1e-3).astype(F)
    sh = shapes(ni, nj, nk)
    want_t = np.zeros((nk, nj, ni), F)
    fl = {"x": np.zeros(sh["u"], F), "y": np.zeros(sh["v"], F), "z": np.zeros(sh["w"], F)}
    oa.cpu_scalar_advection5_inplace(scalar, a, m, divx, want_t, fl, ni=ni, nj=nj, nk=nk, nrk=nrk,
                                     dx=250.0, dy=200.0, dz=150.0, weps=1.0e-17, lateral_sides=sides)
    dev_t = torch.zeros((nk, nj, ni), dtype=torch.float32, device=device)
    dfl = to_dev({"x": np.zeros(sh["u"], F), "y": np.zeros(sh["v"], F), "z": np.zeros(sh["w"], F)}, device)
    launch_cuda_scalar_advection5(
        to_dev(scalar, device), to_dev(a, device), to_dev(m, device), to_dev(divx, device), dev_t, dfl,
        ni=ni, nj=nj, nk=nk, nrk=nrk, dx=250.0, dy=200.0, dz=150.0, weps=1.0e-17, lateral_sides=sides)
    assert_bits_equal(to_host(dev_t), want_t, "tendency")
    for d in "xyz":
        assert_bits_equal(to_host(dfl[d]), fl[d], f"flux {d}")
    # the flux-free launch gives the same tendency
    dev_t2 = torch.zeros_like(dev_t)
    launch_cuda_scalar_advection5(
        to_dev(scalar, device), to_dev(a, device), to_dev(m, device), to_dev(divx, device), dev_t2, None,
        ni=ni, nj=nj, nk=nk, nrk=nrk, dx=250.0, dy=200.0, dz=150.0, weps=1.0e-17, lateral_sides=sides)
    assert torch.equal(dev_t2.view(torch.int32), dev_t.view(torch.int32))


@pytest.mark.gpu
def test_gpu_golden_vectors_through_c_abi(device):
    from gm1_amd.advection5 import launch_cuda_scalar_advection5, launch_cuda_terrain_mass_flux5
    import torch

    g = golden("advection5")
    for case in (0, 1):
        ni, nj, nk, a, m = _golden_case(g, case)
        da, dm = to_dev(a, device), to_dev(m, device)
        dd = torch.zeros((nk, nj, ni), dtype=torch.float32, device=device)
        full_m = {k: to_dev(v, device) for k, v in m.items()}
        launch_cuda_terrain_mass_flux5(da, full_m, dd, ni=ni, nj=nj, nk=nk, dx=250.0, dy=200.0, dz=150.0)
        for k in ("rru", "rrv", "rrw"):
            assert_bits_equal(to_host(da[k]), g[f"c{case}_out_{k}"], k)
        for nrk in (1, 3):
            t = torch.zeros((nk, nj, ni), dtype=torch.float32, device=device)
            launch_cuda_scalar_advection5(to_dev(g[f"c{case}_scalar"], device), da, dm, dd, t, None,
                                          ni=ni, nj=nj, nk=nk, nrk=nrk, dx=250.0, dy=200.0, dz=150.0, weps=1.0e-17)
            assert_bits_equal(to_host(t), g[f"c{case}_nrk{nrk}_tend"], f"tend nrk={nrk}")
