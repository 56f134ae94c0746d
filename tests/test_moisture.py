"""Moisture transport: the final-stage positive-definite chain and the moist RK stage.

CPU tests pin ``oracle/pdef.py`` to the reference's ``cpu_pdef_direction_inplace``
(cuda_scalar_moist_rk.py:352-520) on its exact interior and to closed forms;
GPU tests compare the gfx950 kernels with the oracle bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import F, NG, assert_bits_equal, golden, seeded_metrics, shapes, to_dev, to_host
from oracle import pdef as op


def _metrics_for_repair(g, ni, nj, nk):
    return {"rho0_3d": g["rho0"], "gz": g["gz"], "rgz": g["rgz"]}


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_repair_matches_reference_interior(axis):
    g = golden("pdef")
    ni, nj, nk = (int(v) for v in g["dims"])
    adv, cand = op.repair_direction(
        g["qa"], g["s"], g["rho0"], g["gz"], g["rgz"], g[f"{axis}_flux"], g[f"{axis}_in_directional"].copy(),
        axis=axis, ni=ni, nj=nj, nk=nk, dt=1.5, spacing=float(g[f"{axis}_spacing"]), sides=(1, 1, 1, 1))
    assert_bits_equal(cand, g[f"{axis}_out_candidate"], f"{axis} candidate")
    want = g[f"{axis}_out_directional"]
    if axis == "x":
        sl = (slice(None), slice(None), slice(2, ni - 2))
    elif axis == "y":
        sl = (slice(None), slice(2, nj - 2), slice(None))
    else:
        sl = (slice(None), slice(None), slice(None))
    assert_bits_equal(adv[sl], want[sl], f"{axis} corrected tendency (exact interior)")
    # the repair was exercised
    assert (adv[sl] != g[f"{axis}_in_directional"][sl]).mean() > 0.2


def test_repair_conserves_a_periodic_row_and_leaves_positive_rows_alone():
    rng = np.random.default_rng(3)
    n = 16
    vals = rng.normal(1e-3, 1e-3, (4, n)).astype(F)
    vals[0] = np.abs(vals[0])  # a row with no negatives is untouched
    mass, flag = op._row_mass(vals, op._family_index(n, 1, 1))
    assert not flag[0].any()
    repaired = np.where(flag, vals + mass, vals).astype(np.float64)
    np.testing.assert_allclose(repaired.sum(axis=1), vals.astype(np.float64).sum(axis=1), rtol=0, atol=1e-9)
    # every negative cell that had a positive donor is raised
    has_donor = (np.maximum(np.roll(vals, 1, 1), 0) + np.maximum(np.roll(vals, -1, 1), 0)) > 1e-30
    assert (repaired[(vals < 0) & has_donor] > vals[(vals < 0) & has_donor]).all()


def test_family_index_maps_halos_like_the_fills():
    n = 6
    np.testing.assert_array_equal(op._family_index(n, 1, 1), [4, 5, 0, 1, 2, 3, 4, 5, 0, 1])
    np.testing.assert_array_equal(op._family_index(n, 2, 2), [0, 0, 0, 1, 2, 3, 4, 5, 5, 5])
    np.testing.assert_array_equal(op._family_index(n, 3, 4), [1, 0, 0, 1, 2, 3, 4, 5, 5, 4])
