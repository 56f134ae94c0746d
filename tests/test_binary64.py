"""Rounding error of the binary32 large step against the same statements in binary64.

The GPU large step equals the oracle word for word (the GPU parity tests), so
the oracle's binary32 error is the GPU's.  Here the oracle runs a second time
with its arithmetic type switched to binary64 (GM1_ORACLE_BINARY64=1 at
import, a separate copy of the package) from the same binary32 initial words,
and the binary32 result's relative max-norm error max|x32 - x64| / max|x64|
is measured after one large step:

* a dry C1-like periodic box (uniform dz) and the same box with
  open-radiative sides;
* a dry flat stretched (stretch_z=1) column, with the Thomas elimination run
  once more in binary64 inside the binary32 step (the restatement's
  TRIDIAGONAL_BINARY64 knob) to measure how much of the error the elimination
  contributes.

The bounds asserted are regression guards around the measured values.
"""

from __future__ import annotations

import importlib
import importlib.util
import os
import sys
import types
from pathlib import Path

import numpy as np
import pytest

from gm1_amd.fields import CaseConfig
from gm1_amd.large_step import host_initial_state
from oracle import large_step as ol32
from oracle import sound as sound32

ORACLE = Path(__file__).resolve().parents[1] / "oracle"
FIELDS = ("u3d", "w3d", "th3d", "pp3d")


def oracle_binary64():
    """A second copy of the oracle package whose arithmetic type is binary64."""
    name = "oracle_binary64"
    if name in sys.modules:
        return sys.modules[name]
    saved = os.environ.get("GM1_ORACLE_BINARY64")
    os.environ["GM1_ORACLE_BINARY64"] = "1"
    try:
        spec = importlib.util.spec_from_file_location(name, ORACLE / "__init__.py",
                                                      submodule_search_locations=[str(ORACLE)])
        pkg = importlib.util.module_from_spec(spec)
        sys.modules[name] = pkg
        spec.loader.exec_module(pkg)
        for sub in ("common", "boundary", "stage", "sound", "advection5", "momentum5", "pdef", "kessler",
                    "large_step", "adjoint_sound", "adjoint_large_step"):
            importlib.import_module(f"{name}.{sub}")
    finally:
        if saved is None:
            del os.environ["GM1_ORACLE_BINARY64"]
        else:
            os.environ["GM1_ORACLE_BINARY64"] = saved
    assert pkg.common.F is np.float64
    return pkg


def promote(x):
    if isinstance(x, dict):
        return {k: promote(v) for k, v in x.items()}
    if isinstance(x, np.ndarray) and x.dtype == np.float32:
        return x.astype(np.float64)
    return x


def rel_err(x32, x64):
    return float(np.abs(x32.astype(np.float64) - x64).max() / max(np.abs(x64).max(), 1e-30))


def cfg_dict(cfg):
    return {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
            "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}


def dry_box(sides):
    ni, nj, nk = 32, 28, 20
    return CaseConfig(f"dry-{ni}x{nj}x{nk}", ni, nj, nk, 1000.0, 1000.0, 500.0, 6.0, 6, sides, False, "linear15",
                      ((2.0, ni * 450.0, nj * 520.0, 1400.0),), bubble_radius=(ni * 300.0, 1400.0), noise=0.05,
                      seed=3)


def stretched_dry(sides):
    ni, nj, nk = 32, 28, 16
    return CaseConfig(f"stretched-dry-{ni}x{nj}x{nk}", ni, nj, nk, 100.0, 100.0, 100.0, 0.6, 10, sides, False,
                      "hodograph", ((3.0, ni * 50.0, nj * 50.0, 600.0),), bubble_radius=(ni * 30.0, 500.0),
                      noise=0.1, seed=7, stretched=(1600.0, 100.0, 700.0, 20.0, 180.0))


@pytest.mark.parametrize("sides,bounds", [
    ((1, 1, 1, 1), {"u3d": 1e-6, "w3d": 5e-5, "th3d": 5e-6, "pp3d": 5e-5}),
    ((2, 2, 2, 2), {"u3d": 1e-6, "w3d": 5e-4, "th3d": 5e-6, "pp3d": 5e-4}),
])
def test_uniform_dry_step_error_against_binary64(sides, bounds):
    o64 = oracle_binary64()
    cfg = dry_box(sides)
    host = host_initial_state(cfg)
    c = cfg_dict(cfg)
    a32 = {k: v.copy() for k, v in host["arrays"].items()}
    a64 = promote(host["arrays"])
    ol32.large_step(a32, host["metrics"], host["base"], sound32.terrain_coordinate_profiles(cfg.nk, cfg.dz), c,
                    zt=np.float32(cfg.nk * cfg.dz))
    o64.large_step.large_step(a64, promote(host["metrics"]), promote(host["base"]),
                              o64.sound.terrain_coordinate_profiles(cfg.nk, cfg.dz), c, zt=np.float64(cfg.nk * cfg.dz))
    errs = {n: rel_err(a32[n], a64[n]) for n in FIELDS}
    print(sides, {n: f"{e:.2e}" for n, e in errs.items()})
    assert all(a64[n].dtype == np.float64 for n in FIELDS)
    for n in FIELDS:
        assert 0.0 < errs[n] <= bounds[n], (n, errs[n])


def test_stretched_dry_step_error_and_the_elimination_share():
    o64 = oracle_binary64()
    cfg = stretched_dry((2, 2, 2, 2))
    host = host_initial_state(cfg)
    lv = host["levels"]
    c = cfg_dict(cfg)
    lv64 = types.SimpleNamespace(mh=lv.mh.astype(np.float64), mf=lv.mf.astype(np.float64), rdz=float(lv.rdz))
    a32 = {k: v.copy() for k, v in host["arrays"].items()}
    a32t = {k: v.copy() for k, v in host["arrays"].items()}
    a64 = promote(host["arrays"])
    ol32.large_step(a32, host["metrics"], host["base"], host["prof"], c, zt=np.float32(lv.ztop), flat=lv)
    sound32.TRIDIAGONAL_BINARY64 = True
    try:
        ol32.large_step(a32t, host["metrics"], host["base"], host["prof"], c, zt=np.float32(lv.ztop), flat=lv)
    finally:
        sound32.TRIDIAGONAL_BINARY64 = False
    o64.large_step.large_step(a64, promote(host["metrics"]), promote(host["base"]), promote(host["prof"]), c,
                              zt=np.float64(lv.ztop), flat=lv64)
    errs = {n: rel_err(a32[n], a64[n]) for n in FIELDS}
    errs_t = {n: rel_err(a32t[n], a64[n]) for n in FIELDS}
    print("binary32", {n: f"{e:.2e}" for n, e in errs.items()})
    print("binary32, elimination in binary64", {n: f"{e:.2e}" for n, e in errs_t.items()})
    # the knob does change the words ...
    assert not np.array_equal(a32["w3d"], a32t["w3d"])
    # ... but not the error: the elimination is not where the binary32 error comes from
    for n in FIELDS:
        assert abs(errs_t[n] - errs[n]) <= 0.05 * errs[n], (n, errs[n], errs_t[n])
    assert errs["u3d"] <= 1e-5 and errs["th3d"] <= 1e-6
    assert errs["w3d"] <= 2e-4 and errs["pp3d"] <= 1e-3
