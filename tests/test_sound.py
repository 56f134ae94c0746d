"""psolver=3 acoustic stage: oracle pins, closed-form checks and GPU parity.

Reference behaviour: ``cuda_psolver3_terrain_open.py`` (CPU :300-893) and the
cadence of ``cuda_sound3_column.py:391-424`` pinned by
``tests/test_cuda_sound3_periodic.py:230-236``.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, NG, assert_bits_equal, golden, seeded_metrics, to_dev, to_host
from oracle import boundary as obc
from oracle import sound as osd

OUT = ("u3d", "v3d", "w3d", "pp3d", "ppd", "ppx", "rru", "rrv", "rrw", "radbcw", "radbce", "radbcs", "radbcn",
       "scalar", "qk", "pk1", "pk2", "mm")


def _golden_case(g, c):
    ni, nj, nk, ns, dt, dx, dy, dz = g[f"c{c}_params"]
    a = {k[len(f"c{c}_in_"):]: g[k].copy() for k in g.files if k.startswith(f"c{c}_in_")}
    m = {k[len(f"c{c}_metric_"):]: g[k] for k in g.files if k.startswith(f"c{c}_metric_")}
    return int(ni), int(nj), int(nk), int(ns), float(dt), float(dx), float(dy), float(dz), a, m


@pytest.mark.parametrize("case", [0, 1])
def test_oracle_matches_reference_three_stages(case):
    g = golden("psolver3")
    ni, nj, nk, ns, dt, dx, dy, dz, a, m = _golden_case(g, case)
    for nrk in (1, 2, 3):
        osd.cpu_psolver3_stage_inplace(a, m, ni=ni, nj=nj, nk=nk, dt=dt, nsound=ns, nrk=nrk, dx=dx, dy=dy, dz=dz)
        for k in OUT:
            assert_bits_equal(a[k], g[f"c{case}_nrk{nrk}_{k}"], f"{k} nrk={nrk}")


@pytest.mark.parametrize("nsound,loops", [(6, (2, 3, 6)), (8, (3, 4, 8)), (10, (4, 5, 10))])
def test_cadence_known_answers(nsound, loops):
    got = tuple(osd.sound3_stage_schedule_f32(6.0, nsound, nrk)[0] for nrk in (1, 2, 3))
    assert got == loops
    from gm1_amd.grid import sound3_stage_schedule_f32, small_steps_per_large_step

    assert tuple(sound3_stage_schedule_f32(6.0, nsound, nrk)[0] for nrk in (1, 2, 3)) == loops
    assert small_steps_per_large_step(6.0, nsound) == sum(loops)


def test_profiles_and_fornberg_weights():
    prof = osd.terrain_coordinate_profiles(10, 250.0)
    # quadratic extrapolation to the ground from three equally spaced nodes
    assert np.allclose(prof["cgs"], [1.875, -1.25, 0.375])
    assert np.allclose(prof["cgt"], [1.875, -1.25, 0.375])
    from gm1_amd.grid import terrain_coordinate_profiles

    p2 = terrain_coordinate_profiles(10, 250.0)
    for name in ("sigma", "sigmaf", "rds", "rdsf", "cgs", "cgt"):
        assert_bits_equal(getattr(p2, name), prof[name], name)
    # stretched column: weights still extrapolate a quadratic exactly
    sf = np.concatenate([[0.0], np.cumsum(np.linspace(20.0, 280.0, 12))]).astype(np.float64)
    sigmaf = np.zeros(14, np.float32)
    sigmaf[1:14] = sf
    sigma = np.zeros(14, np.float32)
    sigma[1:13] = 0.5 * (sf[:-1] + sf[1:])
    pr = osd.terrain_coordinate_profiles(12, 20.0, sigma=sigma, sigmaf=sigmaf)
    quad = lambda z: 1.0 + 0.003 * z - 1e-6 * z * z  # noqa: E731
    est = sum(float(pr["cgs"][n]) * quad(float(sigma[1 + n])) for n in range(3))
    assert abs(est - quad(0.0)) < 1e-4


def _acoustic_case(rng, ni, nj, nk, terrain=True):
    from oracle.sound import terrain_coordinate_profiles  # noqa: F401

    shp = {
        "p": (nk + 2, nj + 2, ni + 2), "u": (nk + 2, nj + 2 * NG, ni + 1 + 2 * NG),
        "v": (nk + 2, nj + 1 + 2 * NG, ni + 2 * NG), "w": (nk + 3, nj + 2 * NG, ni + 2 * NG),
    }
    a = {}
    for n in ("ppi", "pp3d", "ppx", "ppd", "piadv", "ppterm", "ppten", "thv", "qk", "pk1", "pk2", "mm", "fpk",
              "elim_c", "elim_d"):
        a[n] = np.zeros(shp["p"], F)
    a["thv"][:] = (300.0 + rng.standard_normal(shp["p"])).astype(F)
    a["piadv"][:] = (1.0 - 0.01 * np.arange(nk + 2)[:, None, None] + 1e-4 * rng.standard_normal(shp["p"])).astype(F)
    a["ppterm"][:] = (0.4 + 0.01 * rng.standard_normal(shp["p"])).astype(F)
    a["ppten"][:] = (1e-5 * rng.standard_normal(shp["p"])).astype(F)
    for n in ("ppi", "pp3d", "ppx"):
        a[n][:] = (1e-4 * rng.standard_normal(shp["p"])).astype(F)
    for n, key, scale in (("ua", "u", 5.0), ("u3d", "u", 5.0), ("uten", "u", 1e-3), ("rru", "u", 0.0),
                          ("va", "v", 5.0), ("v3d", "v", 5.0), ("vten", "v", 1e-3), ("rrv", "v", 0.0),
                          ("wa", "w", 1.0), ("w3d", "w", 1.0), ("wten", "w", 1e-3), ("rrw", "w", 0.0)):
        a[n] = (scale * rng.standard_normal(shp[key])).astype(F)
    for n, shape in (("radbcw", (nk + 2, nj + 2 * NG)), ("radbce", (nk + 2, nj + 2 * NG)),
                     ("radbcs", (nk + 2, ni + 2 * NG)), ("radbcn", (nk + 2, ni + 2 * NG))):
        a[n] = (-5.0 * np.abs(rng.standard_normal(shape))).astype(F)
    a["uh"] = np.ones(ni + 2 * NG, F)
    a["vh"] = np.ones(nj + 2 * NG, F)
    a["ruf"] = np.ones(ni + 1 + 2 * NG, F)
    a["rvf"] = np.ones(nj + 1 + 2 * NG, F)
    a["scalar"] = (1.0 + rng.standard_normal((nk + 2, nj + 2 * NG, ni + 2 * NG))).astype(F)
    m = seeded_metrics(rng, ni, nj, nk, terrain=terrain)
    return a, m


def test_tridiagonal_rows_solve_like_a_dense_binary64_solve():
    """The Thomas sweep's w equals a dense LAPACK solve of the same rows."""
    ni, nj, nk = 9, 8, 10
    a, m = _acoustic_case(np.random.default_rng(4), ni, nj, nk)
    prof = osd.terrain_coordinate_profiles(nk, 200.0)
    nloop, dts = osd.sound3_stage_schedule_f32(3.0, 6, 3)
    dts = F(dts)
    osd.prepare_coefficients(a, m, prof, ni=ni, nj=nj, nk=nk, dts=dts)
    a["ppd"][1:nk + 1] = a["ppx"][1:nk + 1]
    rec = {}
    osd.acoustic_small_step(a, m, prof, (2, 2, 2, 2), ni=ni, nj=nj, nk=nk, dts=dts, rdx=F(1 / 250.0),
                            rdy=F(1 / 250.0), dx=250.0, dy=250.0, zt=F(nk * 200.0), nrk=3, n=1, nloop=nloop,
                            alph=F(0.6), beta=F(0.4), kdiv=F(0.1), tridiagonal=rec)
    rows = {k: (aa, bb, cc, dd) for k, aa, bb, cc, dd in rec["rows"]}
    n = nk - 1
    worst = 0.0
    for j in range(nj):
        for i in range(ni):
            A = np.zeros((n, n))
            rhs = np.zeros(n)
            for r, k in enumerate(range(2, nk + 1)):
                aa, bb, cc, dd = (np.asarray(x)[j, i] if np.ndim(x) else x for x in rows[k])
                A[r, r] = bb
                if r > 0:
                    A[r, r - 1] = aa
                if r < n - 1:
                    A[r, r + 1] = cc
                rhs[r] = dd
            exact = np.linalg.solve(A, rhs)
            got = a["w3d"][2 : nk + 1, j + NG, i + NG].astype(np.float64)
            worst = max(worst, float(np.max(np.abs(got - exact)) / (np.max(np.abs(exact)) + 1e-30)))
            assert np.all(np.abs(np.diag(A)) > 0.9)
    assert worst < 1e-5


def _fill_periodic(a, ni, nj, nk):
    sides = (1, 1, 1, 1)
    obc.fill(a["u3d"], obc.KIND_U, sides, ni=ni, nj=nj, nk=nk)
    obc.fill(a["ua"], obc.KIND_U, sides, ni=ni, nj=nj, nk=nk)
    obc.fill(a["uten"], obc.KIND_U, sides, ni=ni, nj=nj, nk=nk)
    for n in ("v3d", "va", "vten"):
        obc.fill(a[n], obc.KIND_V, sides, ni=ni, nj=nj, nk=nk)
    for n in ("w3d", "wa", "wten"):
        obc.fill(a[n], obc.KIND_W, sides, ni=ni, nj=nj, nk=nk)
    for n in ("ppi", "pp3d", "ppx", "piadv", "ppterm", "ppten", "thv"):
        obc.bcp(a[n], sides, ni=ni, nj=nj, nk=nk)
    # periodic duplicate faces: face ni equals face 0
    a["u3d"][:, :, NG + ni] = a["u3d"][:, :, NG]
    a["ua"][:, :, NG + ni] = a["ua"][:, :, NG]
    a["uten"][:, :, NG + ni] = a["uten"][:, :, NG]
    a["v3d"][:, NG + nj, :] = a["v3d"][:, NG, :]
    a["va"][:, NG + nj, :] = a["va"][:, NG, :]
    a["vten"][:, NG + nj, :] = a["vten"][:, NG, :]


def test_periodic_stage_is_translation_equivariant_and_keeps_duplicate_faces():
    """Shifting a periodic state by whole cells shifts the stage result, bit for bit."""
    ni, nj, nk = 10, 9, 8
    rng = np.random.default_rng(6)
    a, m = _acoustic_case(rng, ni, nj, nk, terrain=False)
    for n in ("rho0_3d", "rr0", "rf0", "rrf0", "c1", "c2"):
        m[n][:] = m[n][:, :1, :1]  # horizontally uniform base
    _fill_periodic(a, ni, nj, nk)
    b = {k: v.copy() for k, v in a.items()}
    s = 3
    for n, key in (("u3d", "u"), ("ua", "u"), ("uten", "u"), ("v3d", "v"), ("va", "v"), ("vten", "v"),
                   ("w3d", "w"), ("wa", "w"), ("wten", "w")):
        core = b[n][:, :, NG:NG + ni].copy()
        b[n][:, :, NG:NG + ni] = np.roll(core, s, axis=2)
    for n in ("ppi", "pp3d", "ppx", "piadv", "ppterm", "ppten", "thv"):
        b[n][:, :, 1:ni + 1] = np.roll(a[n][:, :, 1:ni + 1], s, axis=2)
    _fill_periodic(b, ni, nj, nk)
    for st in (a, b):
        for nrk in (1, 2, 3):
            osd.cpu_psolver3_stage_inplace(st, m, ni=ni, nj=nj, nk=nk, dt=4.0, nsound=6, nrk=nrk, dx=500.0, dy=500.0,
                                           dz=250.0, lateral_sides=(1, 1, 1, 1), radiation=False)
    assert_bits_equal(np.roll(a["pp3d"][1:nk + 1, 1:nj + 1, 1:ni + 1], s, axis=2), b["pp3d"][1:nk + 1, 1:nj + 1, 1:ni + 1], "pp3d")
    assert_bits_equal(np.roll(a["w3d"][:, NG:NG + nj, NG:NG + ni], s, axis=2), b["w3d"][:, NG:NG + nj, NG:NG + ni], "w3d")
    assert_bits_equal(np.roll(a["u3d"][:, NG:NG + nj, NG:NG + ni], s, axis=2), b["u3d"][:, NG:NG + nj, NG:NG + ni], "u3d")
    # the two copies of the periodic boundary face stay equal
    assert_bits_equal(a["u3d"][1:nk + 1, NG:NG + nj, NG + ni], a["u3d"][1:nk + 1, NG:NG + nj, NG], "u face ni vs 0")
    assert_bits_equal(a["v3d"][1:nk + 1, NG + nj, NG:NG + ni], a["v3d"][1:nk + 1, NG, NG:NG + ni], "v face nj vs 0")


# ------------------------------------------------------------------ GPU parity

def _gpu_stage(device, a, m, *, ni, nj, nk, dt, ns, dx, dy, dz, sides, nrks=(1, 2, 3), flags=None,
               column_pass="staged", slopes_zero=False):
    from gm1_amd.psolver3 import TerrainOpenSoundAblations, launch_cuda_psolver3_terrain_open_stage

    da, dm = to_dev(a, device), to_dev(m, device)
    abl = TerrainOpenSoundAblations() if flags is None else flags
    for nrk in nrks:
        launch_cuda_psolver3_terrain_open_stage(da, dm, ni=ni, nj=nj, nk=nk, nrk=nrk, dt=dt, nsound=ns, dx=dx, dy=dy,
                                                dz=dz, lateral_sides=sides, ablations=abl, column_pass=column_pass,
                                                terrain_slopes_zero=slopes_zero)
    return to_host(da)


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
@pytest.mark.parametrize("rest", [False, True])
def test_gpu_flat_slope_pressure_gradient_bit_exact(device, sides, rest):
    """terrain_slopes_zero on flat metrics gives the general text's words, zeros and their signs included
    (a state at rest has h = 0 on every face, where the terrain term is still formed)."""
    from gm1_amd.large_step import slopes_are_zero

    ni, nj, nk = 40, 33, 12
    a, m = _acoustic_case(np.random.default_rng(5), ni, nj, nk, terrain=False)
    assert slopes_are_zero(m)
    if rest:
        for n in ("u3d", "v3d", "w3d", "ua", "va", "wa", "uten", "vten", "wten", "ppten", "ppi", "pp3d", "ppx",
                  "ppd"):
            a[n][...] = F(0.0)
    want = {k: v.copy() for k, v in a.items()}
    got = _gpu_stage(device, a, m, ni=ni, nj=nj, nk=nk, dt=2.0, ns=8, dx=250.0, dy=200.0, dz=150.0, sides=sides,
                     slopes_zero=True)
    for nrk in (1, 2, 3):
        osd.cpu_psolver3_stage_inplace(want, m, ni=ni, nj=nj, nk=nk, dt=2.0, nsound=8, nrk=nrk, dx=250.0, dy=200.0,
                                       dz=150.0, lateral_sides=sides)
    for k in OUT:
        assert_bits_equal(got[k], want[k], f"{k} sides={sides} rest={rest}")


@pytest.mark.gpu
def test_gpu_golden_three_stages_through_c_abi(device):
    g = golden("psolver3")
    for case in (0, 1):
        ni, nj, nk, ns, dt, dx, dy, dz, a, m = _golden_case(g, case)
        for k in ("fpk", "elim_c", "elim_d"):
            a.setdefault(k, np.zeros_like(a["qk"]))
        got = _gpu_stage(device, a, m, ni=ni, nj=nj, nk=nk, dt=dt, ns=ns, dx=dx, dy=dy, dz=dz, sides=(2, 2, 2, 2))
        for k in OUT:
            assert_bits_equal(got[k], g[f"c{case}_nrk3_{k}"], f"{k} case={case}")


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1), (1, 1, 2, 2), (3, 4, 1, 1)])
@pytest.mark.parametrize("dims", [(9, 8, 8), (70, 45, 20)])
def test_gpu_acoustic_stage_bit_exact(device, sides, dims):
    ni, nj, nk = dims
    a, m = _acoustic_case(np.random.default_rng(ni + 7 * nk), ni, nj, nk, terrain=True)
    want = {k: v.copy() for k, v in a.items()}
    got = _gpu_stage(device, a, m, ni=ni, nj=nj, nk=nk, dt=2.0, ns=8, dx=250.0, dy=200.0, dz=150.0, sides=sides)
    for nrk in (1, 2, 3):
        osd.cpu_psolver3_stage_inplace(want, m, ni=ni, nj=nj, nk=nk, dt=2.0, nsound=8, nrk=nrk, dx=250.0, dy=200.0,
                                       dz=150.0, lateral_sides=sides)
    for k in OUT:
        assert_bits_equal(got[k], want[k], f"{k} sides={sides}")


@pytest.mark.gpu
@pytest.mark.parametrize("sides", [(2, 2, 2, 2), (1, 1, 1, 1)])
def test_gpu_staged_and_register_column_passes_agree(device, sides):
    """The LDS-staged column passes (every workgroup shape) and the
    one-thread-per-column pass give the same words."""
    ni, nj, nk = 67, 33, 40
    a, m = _acoustic_case(np.random.default_rng(17), ni, nj, nk, terrain=True)
    got = {}
    for mode in ("staged", "pipelined", "staged16", "pipelined16", "register"):
        got[mode] = _gpu_stage(device, {k: v.copy() for k, v in a.items()}, m, ni=ni, nj=nj, nk=nk, dt=2.0, ns=8,
                               dx=250.0, dy=200.0, dz=150.0, sides=sides, column_pass=mode)
    for k in OUT:
        assert_bits_equal(got["staged"][k], got["register"][k], f"{k} sides={sides}")
        assert_bits_equal(got["staged16"][k], got["register"][k], f"{k} sides={sides} (16 columns)")
        for mode in ("pipelined", "pipelined16"):
            assert_bits_equal(got[mode][k], got["register"][k], f"{k} sides={sides} ({mode})")
