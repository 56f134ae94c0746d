"""The flat, uniform, doubly periodic acoustic stage (cuda_sound3_periodic.py).

CPU: the oracle reproduces the reference transcription's golden vectors word
for word (three fixtures: the reference test's smooth 5x4x7 case, and two
seeded cases with non-unit mh/mf and, in one, declared alph/kdiv), and the
reference test's known answers hold (cadence and fill counts, Thomas
residual <= 2e-7 and pivots > 0.9, duplicate periodic faces equal, level
divergence sums < 2e-8, a horizontally uniform state has no divergence or
pressure advection).  GPU: the C-ABI stage reproduces every compared output
bit for bit, on the fixtures and at the 64x64x40 dry-case size.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import sound3_periodic as osp

GOLDEN = Path(__file__).resolve().parent / "golden" / "sound3_periodic.npz"
F = np.float32


def golden_case(tag: str):
    g = np.load(GOLDEN)
    ni, nj, nk = (int(x) for x in g[f"{tag}_dims"])
    p = g[f"{tag}_params"]
    params = dict(dt=float(p[0]), nsound=int(p[1]), rdx=float(p[2]), rdy=float(p[3]), rdz=float(p[4]),
                  alph=float(p[5]), kdiv=float(p[6]))
    inputs = {n: g[f"{tag}_in_{n}"].copy() for n in osp.ALL_ARRAY_NAMES}
    outputs = {n: g[f"{tag}_out_{n}"] for n in osp.COMPARE_OUTPUT_NAMES}
    return (ni, nj, nk), params, inputs, outputs, g[f"{tag}_cadence"]


def seeded_case(ni: int, nj: int, nk: int, seed: int, *, uniform: bool = False) -> dict[str, np.ndarray]:
    """A periodic state of our own construction with every halo filled."""
    rng = np.random.default_rng(seed)
    a = {n: np.zeros(s, F) for n, s in osp.array_shapes(ni, nj, nk).items()}
    a["dtu0"], a["dtv0"] = osp.default_periodic_face_weights(ni, nj)
    for n in ("uf", "vf", "uh", "vh"):
        a[n].fill(F(1.0))
    z = (np.arange(nk + 2) / (nk + 1))[:, None, None]
    h = np.zeros((1, nj, ni)) if uniform else rng.uniform(-1.0, 1.0, (1, nj, ni))
    inner = (slice(None), slice(1, nj + 1), slice(1, ni + 1))
    for name, value in (("mh", 1.0 + 0.2 * np.sin(3 * z)), ("mf", 1.0 + 0.15 * np.cos(2 * z)),
                        ("rr0", 1.0 + 0.01 * z), ("rf0", 0.99 + 0.012 * z), ("c1", 0.45 + 0.05 * z),
                        ("c2", 0.55 - 0.05 * z), ("piadv", 1.0 + 3e-4 * z + 1e-5 * h),
                        ("ppterm", 0.014 + 2e-4 * z), ("thv", 299.0 + 1.8 * z + 0.03 * h),
                        ("ppten", 2e-6 * h + 1e-6 * z), ("ppi", 2e-4 * h + 4e-5 * z),
                        ("ppx", 4e-5 * h + 1e-6 * z)):
        a[name][inner] = np.broadcast_to(value, (nk + 2, nj, ni)).astype(F)
        osp.fill_periodic_scalar(a[name], ni=ni, nj=nj, nk=nk)
    zk = (np.arange(1, nk + 1) / nk)[:, None, None]
    a["wa"][1 : nk + 1] = (3e-3 * h * np.sin(np.pi * zk)).astype(F)
    a["wten"][1 : nk + 1] = np.broadcast_to(2e-5 * h, (nk, nj, ni)).astype(F)
    hu = 0.0 if uniform else 1.0
    a["ua"][1 : nk + 1, 1 : nj + 1, :ni] = (4.0 + hu * 0.15 * rng.uniform(-1, 1, (nk, nj, ni))).astype(F)
    a["ua"][1 : nk + 1, 1 : nj + 1, ni] = a["ua"][1 : nk + 1, 1 : nj + 1, 0]
    a["va"][1 : nk + 1, :nj, 1 : ni + 1] = (-1.5 + hu * 0.12 * rng.uniform(-1, 1, (nk, nj, ni))).astype(F)
    a["va"][1 : nk + 1, nj, 1 : ni + 1] = a["va"][1 : nk + 1, 0, 1 : ni + 1]
    a["uten"][1 : nk + 1, 1 : nj + 1, :ni] = (hu * 1e-4 * rng.uniform(-1, 1, (nk, nj, ni))).astype(F)
    a["uten"][1 : nk + 1, 1 : nj + 1, ni] = a["uten"][1 : nk + 1, 1 : nj + 1, 0]
    a["vten"][1 : nk + 1, :nj, 1 : ni + 1] = (hu * 8e-5 * rng.uniform(-1, 1, (nk, nj, ni))).astype(F)
    a["vten"][1 : nk + 1, nj, 1 : ni + 1] = a["vten"][1 : nk + 1, 0, 1 : ni + 1]
    osp.fill_periodic_uv(a["ua"], a["va"], ni=ni, nj=nj, nk=nk)
    for dst, src in (("u3d", "ua"), ("v3d", "va"), ("w3d", "wa"), ("pp3d", "ppi"), ("ppd", "ppx")):
        np.copyto(a[dst], a[src])
    return a


def differing(a: dict, b: dict, names=osp.COMPARE_OUTPUT_NAMES) -> dict[str, int]:
    out = {}
    for n in names:
        d = int((np.asarray(a[n]).view(np.uint32) != np.asarray(b[n]).view(np.uint32)).sum())
        if d:
            out[n] = d
    return out


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_oracle_matches_reference_golden_vectors(tag):
    dims, params, a, expected, cadence = golden_case(tag)
    ni, nj, nk = dims
    stages = osp.three_stage(a, ni=ni, nj=nj, nk=nk, **params)
    assert differing(a, expected) == {}
    got = [[s["nrk"], s["nloop"], s["ppd_boundary_fills"], s["stage_pressure_boundary_fills"]] for s in stages]
    assert got == cadence.tolist()


def test_reference_known_answers_on_the_smooth_fixture():
    (ni, nj, nk), params, a, _, _ = golden_case("a")
    u0, v0 = a["u3d"].copy(), a["v3d"].copy()
    stages = osp.three_stage(a, ni=ni, nj=nj, nk=nk, **params)
    assert [(s["nrk"], s["nloop"]) for s in stages] == [(1, 2), (2, 3), (3, 6)]
    assert [s["ppd_boundary_fills"] for s in stages] == [1, 2, 5]
    assert [s["stage_pressure_boundary_fills"] for s in stages] == [1, 1, 0]
    assert not np.array_equal(a["u3d"].view(np.uint32), u0.view(np.uint32))
    assert not np.array_equal(a["v3d"].view(np.uint32), v0.view(np.uint32))
    I, J, K = slice(1, ni + 1), slice(1, nj + 1), slice(1, nk + 1)
    assert np.any(a["divh"][K, J, I] != 0) and np.any(a["hadvp"][K, J, I] != 0)
    assert float(np.max(np.abs(a["tri_residual"][2 : nk + 1, J, I]))) <= 2.0e-7
    assert float(np.min(np.abs(a["tri_pivot"][2 : nk + 1, J, I]))) > 0.9
    assert np.array_equal(a["u3d"][K, J, 0].view(np.uint32), a["u3d"][K, J, ni].view(np.uint32))
    assert np.array_equal(a["v3d"][K, 0, I].view(np.uint32), a["v3d"][K, nj, I].view(np.uint32))
    sums = np.sum(a["divh"][K, J, I], axis=(1, 2), dtype=np.float64)
    assert float(np.max(np.abs(sums))) < 2.0e-8


def test_thomas_solution_matches_a_dense_binary64_solve():
    """The stored tridiagonal rows, solved densely in binary64, give the returned w to binary32 accuracy."""
    ni, nj, nk = 6, 5, 12
    a = seeded_case(ni, nj, nk, 11)
    osp.stage(a, ni=ni, nj=nj, nk=nk, dt=3.0, nsound=6, nrk=1, rdx=1 / 500, rdy=1 / 500, rdz=1 / 250)
    n = nk - 1
    for j in range(1, nj + 1):
        for i in range(1, ni + 1):
            m = np.zeros((n, n))
            rhs = np.zeros(n)
            for r, k in enumerate(range(2, nk + 1)):
                m[r, r] = a["tri_bb"][k, j, i]
                if r > 0:
                    m[r, r - 1] = a["tri_aa"][k, j, i]
                if r < n - 1:
                    m[r, r + 1] = a["tri_cc"][k, j, i]
                rhs[r] = a["tri_dd"][k, j, i]
            exact = np.linalg.solve(m, rhs)
            got = a["w3d"][2 : nk + 1, j - 1, i - 1].astype(np.float64)
            assert np.allclose(got, exact, rtol=1e-5, atol=1e-9 * max(1.0, np.abs(exact).max()))


def test_uniform_state_has_no_horizontal_terms_and_identical_columns():
    ni, nj, nk = 5, 4, 9
    a = seeded_case(ni, nj, nk, 3, uniform=True)
    osp.three_stage(a, ni=ni, nj=nj, nk=nk, dt=2.0, nsound=6, rdx=1e-3, rdy=1e-3, rdz=2e-3)
    I, J, K = slice(1, ni + 1), slice(1, nj + 1), slice(1, nk + 1)
    assert not np.any(a["divh"][K, J, I]) and not np.any(a["hadvp"][K, J, I])
    for name in ("pp3d", "ppd", "fpk", "tri_pivot"):
        block = a[name][K, J, I]
        assert np.array_equal(block.view(np.uint32), np.broadcast_to(block[:, :1, :1], block.shape).view(np.uint32))
    assert np.array_equal(a["w3d"].view(np.uint32), np.broadcast_to(a["w3d"][:, :1, :1], a["w3d"].shape).view(np.uint32))


def test_a_perturbed_coefficient_reaches_the_outputs():
    """Changing one interior mh word by 1 % changes the compared outputs (the fixture is sensitive)."""
    (ni, nj, nk), params, a, expected, _ = golden_case("b")
    a["mh"][4, 3, 5] = F(a["mh"][4, 3, 5] * F(1.01))
    osp.fill_periodic_scalar(a["mh"], ni=ni, nj=nj, nk=nk)
    osp.three_stage(a, ni=ni, nj=nj, nk=nk, **params)
    assert differing(a, expected)


def test_host_shapes_and_face_weights_match_the_oracle():
    from gm1_amd import sound3_periodic as sp

    assert sp.ALL_ARRAY_NAMES == osp.ALL_ARRAY_NAMES
    assert sp.array_shapes(9, 7, 10) == osp.array_shapes(9, 7, 10)
    for got, want in zip(sp.default_periodic_face_weights(9, 7), osp.default_periodic_face_weights(9, 7)):
        assert np.array_equal(got, want)
    with pytest.raises(ValueError, match="ni must be"):
        sp.array_shapes(2, 7, 10)


# ----------------------------------------------------------------------------- GPU


def _to_device(a: dict, device) -> dict:
    import torch

    return {n: torch.from_numpy(np.ascontiguousarray(v)).to(device) for n, v in a.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_device_stage_reproduces_reference_golden_vectors(tag):
    import torch

    from gm1_amd.sound3_periodic import launch_cuda_sound3_periodic_three_stage, validate_periodic_contract

    (ni, nj, nk), params, a, expected, _ = golden_case(tag)
    d = _to_device(a, "cuda")
    validate_periodic_contract(d, ni=ni, nj=nj, nk=nk)
    print(f"periodic stage golden {tag} {ni}x{nj}x{nk}", flush=True)
    launch = launch_cuda_sound3_periodic_three_stage(d, ni=ni, nj=nj, nk=nk, **params)
    launch.ready_event.synchronize()
    got = {n: d[n].cpu().numpy() for n in osp.COMPARE_OUTPUT_NAMES}
    assert differing(got, expected) == {}
    assert [c["nloop"] for c in launch.receipt["cadence"]] == ([3, 4, 8] if tag == "c" else [2, 3, 6])
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_device_stage_matches_oracle_at_the_dry_case_size():
    from gm1_amd.sound3_periodic import launch_cuda_sound3_periodic_stage

    ni, nj, nk = 64, 64, 40
    a = seeded_case(ni, nj, nk, 64)
    d = _to_device(a, "cuda")
    params = dict(dt=6.0, nsound=6, rdx=1e-3, rdy=1e-3, rdz=1 / 500)
    print(f"periodic stage {ni}x{nj}x{nk} vs oracle", flush=True)
    for nrk in (1, 2, 3):
        launch_cuda_sound3_periodic_stage(d, ni=ni, nj=nj, nk=nk, nrk=nrk, **params)
        osp.stage(a, ni=ni, nj=nj, nk=nk, nrk=nrk, **params)
        got = {n: d[n].cpu().numpy() for n in osp.COMPARE_OUTPUT_NAMES}
        assert differing(got, a) == {}, f"stage {nrk}"


@pytest.mark.gpu
def test_device_contract_refuses_a_broken_halo():
    from gm1_amd.sound3_periodic import validate_periodic_contract

    (ni, nj, nk), _, a, _, _ = golden_case("a")
    a["ppx"][2, 0, 2] = np.nextafter(a["ppx"][2, 0, 2], F(np.inf), dtype=F)
    with pytest.raises(ValueError, match="south periodic halo"):
        validate_periodic_contract(_to_device(a, "cuda"), ni=ni, nj=nj, nk=nk)
