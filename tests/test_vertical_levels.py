"""CM1's vertical coordinate (stretch_z 0 and 1) as the stretched family consumes it.

Pinned by the reference's published per-array SHA-256 digests of its 15 km
stretch_z=1 child grid (stretch_z1_child_15km.py:22-48: nk 75, dz 100,
ztop 15 km, 20 m below 200 m, 320 m above 7 km), by param.F's uniform column
written straight out (zf(k) = (k-1)*dz, mh = mf = 1) for stretch_z=0, and by
the refusals CM1 stops on.
"""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

from gm1_amd.grid import VerticalGridError, build_vertical_levels

F = np.float32

CHILD_15KM = dict(stretch_z=1, nk=75, dz=100.0, ztop=15000.0, str_bot=200.0, str_top=7000.0, dz_bot=20.0,
                  dz_top=320.0)
CHILD_15KM_SHA256 = {
    "zf": "ceb1624e4dc10eb31562c7e8003f939d6d157e1ecd73d1dee479de5459f2a90e",
    "zh": "e0c800cd95ac68f526efef46de91db5463784e91860f52ea90ad76c9e18037c3",
    "sigmaf": "ceb1624e4dc10eb31562c7e8003f939d6d157e1ecd73d1dee479de5459f2a90e",
    "sigma": "e0c800cd95ac68f526efef46de91db5463784e91860f52ea90ad76c9e18037c3",
    "mh": "bc520e91ecab72cd55662dbb301c29f78cc04864ec371b3abba5dca8d8c6cbcc",
    "mf": "e686a95701fb4694da25878934e7da4713b512831a06adf549ef61e1704e245f",
    "rmh": "54833ee529867d7c005060a52c3d9e883bfe34731fd5fb74cdba15afbd039b35",
    "rmf": "ab5f395ab5c7726528eb7f7f4dd1946911d03537eb94ea9b38e1ed1dc8551515",
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).astype("<f4").tobytes()).hexdigest()


def test_stretch_z1_reproduces_the_reference_child_grid_digests():
    levels = build_vertical_levels(**CHILD_15KM)
    for name, digest in CHILD_15KM_SHA256.items():
        assert sha(getattr(levels, name)) == digest, name
    assert levels.zh[1] == F(10.0) and levels.zf[76] == F(15000.0)
    assert np.all(levels.spacings[:11] == 20.0)


@pytest.mark.parametrize("nk,dz", [(4, 1.0), (16, 100.0), (40, 500.0), (80, 250.0), (48, 333.0)])
def test_stretch_z0_is_the_uniform_column_written_out(nk, dz):
    levels = build_vertical_levels(stretch_z=0, nk=nk, dz=dz)
    d = F(dz)
    zf = np.array([F(F(k - 1) * d) for k in range(nk + 3)], np.float32)
    zf[0] = F(-zf[2])
    zf[nk + 2] = F(zf[nk + 1] + F(zf[nk + 1] - zf[nk]))
    assert np.array_equal(levels.zf.view(np.uint32), zf.view(np.uint32))
    mh = np.array([F(d / F(zf[k + 1] - zf[k])) for k in range(nk + 2)], np.float32)
    assert np.array_equal(levels.mh.view(np.uint32), mh.view(np.uint32))
    assert np.all(levels.mh[1 : nk + 1] == F(1.0)) and np.all(levels.mf[2 : nk + 1] == F(1.0))
    # the terrain-coordinate profiles of the uniform column are the uniform ones
    from gm1_amd.grid import terrain_coordinate_profiles

    a, b = levels.profiles(), terrain_coordinate_profiles(nk, dz)
    for name in ("sigma", "sigmaf", "rds", "rdsf", "cgs", "cgt"):
        assert np.array_equal(getattr(a, name)[1 : nk + 1], getattr(b, name)[1 : nk + 1]), name


def test_large_domain_grid_is_feasible_and_graded():
    from gm1_amd.fields import large_stretched_c3

    cfg = large_stretched_c3()
    levels = cfg.levels()
    assert levels.nk == 80 and levels.zf[81] == F(17850.0)
    s = levels.spacings
    assert np.all(s[:10] == 20.0) and np.all(s[-55:] == 280.0)
    assert np.all(np.diff(s) >= -1e-3)  # monotone grading
    assert levels.zh[1] == F(10.0)


def test_infeasible_columns_are_refused():
    # 20 m layers up to 500 m and 280 m layers to 18 km leave no room for the stretch in 80 levels
    with pytest.raises(VerticalGridError, match="nk2>=2"):
        build_vertical_levels(stretch_z=1, nk=80, dz=100.0, ztop=18000.0, str_bot=500.0, str_top=2450.0, dz_bot=20.0,
                              dz_top=280.0)
    with pytest.raises(VerticalGridError, match="feasibility"):
        build_vertical_levels(stretch_z=1, nk=80, dz=100.0, ztop=17850.0, str_bot=200.0, str_top=2500.0, dz_bot=20.0,
                              dz_top=280.0)
    with pytest.raises(VerticalGridError, match="stretch_z=2"):
        build_vertical_levels(stretch_z=2, nk=80, dz=100.0, ztop=18000.0)
    with pytest.raises(VerticalGridError, match="dz\\*nk"):
        build_vertical_levels(stretch_z=0, nk=40, dz=500.0, ztop=19000.0)
