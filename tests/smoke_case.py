"""One small invocation of the hot path on the GPU, checked against the oracle.

A 24x20x12 dry warm-bubble state on open-radiative laterals takes one full
CM1 large step (three RK stages, 11 acoustic small steps) through the C ABI;
every prognostic carrier must equal the oracle's bit for bit.
"""

from __future__ import annotations

import numpy as np

PROGNOSTIC = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ppd")


def run_smoke(device) -> dict:
    import torch

    from gm1_amd.fields import CaseConfig
    from gm1_amd.large_step import create_gm1_large_step_state, host_initial_state, launch_cuda_gm1_large_step
    from oracle import large_step as ol
    from oracle import sound as osound

    ni, nj, nk = 24, 20, 12
    cfg = CaseConfig("smoke-dry-open-24x20x12", ni, nj, nk, 1000.0, 1000.0, 500.0, 6.0, 6, (2, 2, 2, 2), False,
                     "hodograph", ((2.0, ni * 450.0, nj * 520.0, 1400.0),), bubble_radius=(ni * 300.0, 1400.0),
                     noise=0.05, seed=7)
    host = host_initial_state(cfg)
    st = create_gm1_large_step_state(host, device)
    launch_cuda_gm1_large_step(st)
    torch.cuda.synchronize()
    a = {k: v.copy() for k, v in host["arrays"].items()}
    prof = osound.terrain_coordinate_profiles(nk, cfg.dz)
    c = {"ni": ni, "nj": nj, "nk": nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}
    ol.large_step(a, host["metrics"], host["base"], prof, c, zt=np.float32(nk * cfg.dz))
    report = {"ok": True, "case": cfg.name}
    for name in PROGNOSTIC:
        got = st.a[name].detach().cpu().numpy()
        differ = int((got.view(np.uint32) != a[name].view(np.uint32)).sum())
        if differ:
            report["ok"] = False
            report.setdefault("mismatch", {})[name] = differ
    report["max_w"] = float(np.abs(a["w3d"]).max())
    return report
