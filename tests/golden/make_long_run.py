"""Write the long-run records of the moist warm-bubble case from the oracle.

Run in the build container (about twelve minutes per mode on one core):

    python tests/golden/make_long_run.py replay        # -> moist_long_run.npz
    python tests/golden/make_long_run.py independent   # -> moist_long_run_independent.npz

The case is the 320x320x80 moist configuration's physics (dx = dy = dz =
250 m, open-radiative laterals, dt 1.5 s, 8 substeps, Kessler, the
Weisman-Klemp sounding and hodograph, +3 K bubble with seeded noise) on a
48x48x40 grid with the bubble radius scaled to the domain (2 km), stepped
600 times.  Every 50 steps each record holds the domain diagnostics (max w,
min pi', total condensate mass) and a SHA-256 of the carriers' words.

* ``replay``: the dry-mass correction of each step is the device's binary32
  word (``moist_long_run_device.npz``, written on a HIP device by
  ``record_long_run_device.py``), applied after the oracle checks that its
  own ledger agrees with the device's (dry mass and Exner sum within 1e-13,
  its own binary32 correction within what that ledger difference moves it:
  the correction cancels, (mass1/mass2)^r - 1, so a last-bit difference of
  the dry mass moves it by several binary32 ulps).  The
  device's run must then equal this record word for word.
* ``independent``: the oracle's own correction every step (its ledger sums
  in its own order).  The two runs differ only where a ledger sum rounds to
  a neighbouring binary32 correction; the difference of their diagnostics
  is the run's sensitivity to one-ulp perturbations.

``tests/test_configs.py`` defines the case, the diagnostics and the digest.
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
for p in (ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

STEPS = 600
EVERY = 50


def main(mode: str) -> None:
    from gm1_amd.large_step import host_initial_state
    from oracle import kessler as ok
    from oracle import large_step as ol
    from oracle import sound as osound
    from test_configs import LONG_RUN_CARRIERS, carrier_digest, long_run_case, long_run_diagnostics

    if mode not in ("replay", "independent"):
        raise SystemExit("mode must be 'replay' or 'independent'")
    cfg = long_run_case()
    host = host_initial_state(cfg)
    a = {k: v.copy() for k, v in host["arrays"].items()}
    initial = carrier_digest(a)
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    c = {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}
    mass1 = ok.dry_mass(a["rho"], a["ruh"], a["rvh"], a["rmh"])
    dev = None
    if mode == "replay":
        dev = np.load(HERE / "moist_long_run_device.npz")
        assert str(dev["initial_digest"]) == initial, "the device started from other initial words"
        assert abs(float(dev["mass1"]) - mass1) <= 1e-13 * mass1
        mass1 = float(dev["mass1"])
    corrections = np.zeros(STEPS, np.float32)
    own = np.zeros(STEPS, np.float32)
    ledger_rel = np.zeros((STEPS, 2))
    steps, diag, digests = [], [], []
    t0 = time.time()
    for step in range(1, STEPS + 1):
        rec: dict = {}

        def hook(arr, _rec=rec, _step=step):
            led = ok.apmasscon_correction(arr, mass1)
            _rec.update(led)
            corr = led["correction"]
            if dev is not None:
                d2, dp = dev["ledger"][_step - 1]
                _rec["rel"] = (abs(led["mass2"] - d2) / abs(d2), abs(led["pressure_sum"] - dp) / abs(dp))
                corr = float(dev["corrections"][_step - 1])
            _rec["applied"] = corr
            ok.apply_mass_correction(arr, np.float32(corr))

        ol.large_step(a, host["metrics"], host["base"], prof, c, zt=np.float32(cfg.nk * cfg.dz), mass_hook=hook)
        corrections[step - 1] = np.float32(rec["applied"])
        own[step - 1] = np.float32(rec["correction"])
        if dev is not None:
            ledger_rel[step - 1] = rec["rel"]
            assert max(rec["rel"]) <= 1e-13, (step, rec["rel"])
            # corr = ((mass1/mass2)^r - 1) * psum/n cancels: a mass2 that differs by dm moves it by about
            # r*(psum/n)*dm/mass2, far more than one binary32 ulp of corr when mass2 ~ mass1
            n = cfg.ni * cfg.nj * cfg.nk
            d2 = float(dev["ledger"][step - 1][0])
            slope = ok.RDDCV * abs(rec["pressure_sum"]) / n / d2
            bound = slope * (abs(rec["mass2"] - d2) + 4.0 * np.finfo(np.float64).eps * d2)
            bound += 2.0 * float(np.spacing(abs(corrections[step - 1])))
            assert abs(float(own[step - 1]) - float(corrections[step - 1])) <= bound, (step, own[step - 1],
                                                                                   corrections[step - 1], bound)
        if step % EVERY == 0:
            steps.append(step)
            diag.append(long_run_diagnostics(a["w3d"], a["pp3d"], a["q3d"], a["rho"], cfg.ni, cfg.nj, cfg.nk,
                                             cfg.dx, cfg.dy, cfg.dz))
            digests.append(carrier_digest(a))
            print(f"step {step}: max w {diag[-1][0]:.4f} min pp {diag[-1][1]:.6e} condensate {diag[-1][2]:.6e} "
                  f"({time.time() - t0:.0f} s)", flush=True)
    name = "moist_long_run.npz" if mode == "replay" else "moist_long_run_independent.npz"
    extra = {}
    if dev is not None:
        extra = {"own_corrections": own, "ledger_rel": ledger_rel,
                 "own_differs": np.int64(np.count_nonzero(own.view(np.uint32) != corrections.view(np.uint32)))}
    np.savez_compressed(HERE / name, steps=np.array(steps, np.int64), diagnostics=np.array(diag),
                        corrections=corrections, mass1=np.float64(mass1), initial_digest=np.array(initial),
                        digests=np.array(digests, dtype="U64"), carriers=np.array(LONG_RUN_CARRIERS, dtype="U8"),
                        **extra)
    print("wrote", name, extra.get("own_differs", ""))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "replay")
