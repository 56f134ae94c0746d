"""One small invocation of the hot path on the GPU, checked against the oracle."""

from __future__ import annotations

import numpy as np


def run_smoke(device) -> dict:
    import torch

    from helpers import F, assert_bits_equal, seeded_metrics, shapes, to_dev, to_host
    from oracle import advection5 as oa
    from gm1_amd.advection5 import launch_cuda_scalar_advection5, launch_cuda_terrain_mass_flux5

    ni, nj, nk = 24, 20, 12
    rng = np.random.default_rng(5)
    sh = shapes(ni, nj, nk)
    m = seeded_metrics(rng, ni, nj, nk)
    a = {"u3d": (3 * rng.standard_normal(sh["u"])).astype(F), "v3d": (3 * rng.standard_normal(sh["v"])).astype(F),
         "w3d": rng.standard_normal(sh["w"]).astype(F), "rru": np.zeros(sh["u"], F), "rrv": np.zeros(sh["v"], F),
         "rrw": np.zeros(sh["w"], F)}
    scalar = (300 + rng.standard_normal(sh["s"])).astype(F)
    divx = np.zeros((nk, nj, ni), F)
    da, dm, dd = to_dev(a, device), to_dev(m, device), to_dev(divx, device)
    oa.cpu_terrain_mass_flux5_inplace(a, m, divx, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0)
    launch_cuda_terrain_mass_flux5(da, dm, dd, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=250.0)
    t = np.zeros((nk, nj, ni), F)
    fl = {"x": np.zeros(sh["u"], F), "y": np.zeros(sh["v"], F), "z": np.zeros(sh["w"], F)}
    oa.cpu_scalar_advection5_inplace(scalar, a, m, divx, t, fl, ni=ni, nj=nj, nk=nk, nrk=3, dx=250.0, dy=250.0,
                                     dz=250.0, weps=1e-17, lateral_sides=(1, 1, 1, 1))
    dt = torch.zeros((nk, nj, ni), dtype=torch.float32, device=device)
    launch_cuda_scalar_advection5(to_dev(scalar, device), da, dm, dd, dt, None, ni=ni, nj=nj, nk=nk, nrk=3,
                                  dx=250.0, dy=250.0, dz=250.0, weps=1e-17, lateral_sides=(1, 1, 1, 1))
    torch.cuda.synchronize()
    ok = True
    try:
        assert_bits_equal(to_host(dt), t, "tendency")
        assert_bits_equal(to_host(dd), divx, "divx")
    except AssertionError as exc:
        ok = False
        return {"ok": ok, "error": str(exc)}
    return {"ok": ok, "case": "scalar advection5 WENO stage 24x20x12 periodic", "max_abs": float(np.abs(t).max())}
