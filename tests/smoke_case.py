"""One small invocation of the hot path on the GPU, checked against the oracle.

A 24x20x16 moist warm-bubble state (saturated bubble, Kessler warm rain,
dry-mass correction) on open-radiative laterals takes one full CM1 large
step (three RK stages, 15 acoustic small steps, pdef moisture, the Kessler
transition) through the C ABI; every prognostic carrier must equal the
oracle's bit for bit.  The oracle takes the device's dry-mass correction (the
one reduction summed in another order) after checking its ledger.
"""

from __future__ import annotations

import numpy as np

FIELDS = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ppd", "q3d", "rho", "prs", "rain", "qpten", "qtten",
          "qvten", "qcten")


def run_smoke(device) -> dict:
    import torch

    from gm1_amd.fields import CaseConfig
    from gm1_amd.large_step import create_gm1_large_step_state, host_initial_state, launch_cuda_gm1_large_step
    from oracle import large_step as ol
    from oracle import sound as osound

    ni, nj, nk = 24, 20, 16
    cfg = CaseConfig("smoke-moist-open-24x20x16", ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 8, (2, 2, 2, 2), True,
                     "hodograph", ((3.0, ni * 120.0, nj * 130.0, 1000.0),), bubble_radius=(ni * 80.0, 1000.0),
                     noise=0.1, seed=7, extra={"bubble_rh": 1.03})
    host = host_initial_state(cfg)
    st = create_gm1_large_step_state(host, device)
    launch_cuda_gm1_large_step(st)
    torch.cuda.synchronize()
    corr = float(st.dry_mass.correction.cpu().numpy()[0])
    a = {k: v.copy() for k, v in host["arrays"].items()}
    prof = osound.terrain_coordinate_profiles(nk, cfg.dz)
    c = {"ni": ni, "nj": nj, "nk": nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}
    ol.large_step(a, host["metrics"], host["base"], prof, c, zt=np.float32(nk * cfg.dz), mass_correction=corr)
    report = {"ok": True, "case": cfg.name}
    for name in FIELDS:
        got = st.a[name].detach().cpu().numpy()
        differ = int((got.view(np.uint32) != a[name].view(np.uint32)).sum())
        if differ:
            report["ok"] = False
            report.setdefault("mismatch", {})[name] = differ
    report["max_w"] = float(np.abs(a["w3d"]).max())
    report["max_qc"] = float(a["q3d"][1].max())
    return report
