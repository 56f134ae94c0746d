"""The no-allocation Kessler driver against the workspace driver, through the C ABI.

``gm1hip_kessler_driver_noalloc`` runs the same kernels with the library's
file-scope device array as scratch.  Two calls in a row, the second on a
smaller grid, must give the workspace driver's words (fields, binary64
budgets, status); an input too large for the array is refused before any
launch.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_bits_equal, golden
from oracle import kessler as ok

FIELDS = ("rho", "rr", "pp3d", "th3d", "prs", "qv", "qc", "qr", "fall_velocity", "rain", "prate")
POINTERS = ("ruh", "rvh", "rmh", "mh", "pi0", "th0", "rho", "rr", "pp3d", "th3d", "prs", "qv", "qc", "qr",
            "fall_velocity", "rain", "prate", "asq", "qbudget")


def _cases():
    import test_kessler as tk

    cfg, big = tk.c2_like_kessler_fields(48, 40, 80, seed=11)
    g = golden("kessler")
    small = {k[len("d0_in_"):]: g[k].copy() for k in g.files if k.startswith("d0_in_")}
    dt, dz, _ = (float(v) for v in g["d0_params"])
    ni, nj, nk = (int(v) for v in g["dims"])
    return [((48, 40, 80), big, cfg.dz, cfg.dt), ((ni, nj, nk), small, dz, dt)]


def _device(a, device):
    import torch

    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in a.items() if isinstance(v, np.ndarray)}


@pytest.mark.gpu
def test_noalloc_driver_matches_the_workspace_driver_on_two_calls(device):
    import torch

    from gm1_amd import _native
    from gm1_amd._device import ptr, stream_handle
    from gm1_amd.kessler import Gm1KesslerDriverSelection, launch_cuda_gm1_kessler_driver

    for (ni, nj, nk), host, dz, dt in _cases():
        print(f"no-alloc Kessler driver {ni}x{nj}x{nk}", flush=True)
        ref = _device(host, device)
        got = _device(host, device)
        ws = launch_cuda_gm1_kessler_driver(ref, ni=ni, nj=nj, nk=nk, dx=250.0, dy=250.0, dz=dz, dt=dt,
                                            selection=Gm1KesslerDriverSelection(equation_set=2))
        status = torch.zeros(2, dtype=torch.int32, device=device)
        _native.call("gm1hip_kessler_driver_noalloc", *(ptr(got[n]) for n in POINTERS), ptr(status),
                     None, None, None, None, ni, nj, nk, 250.0, 250.0, float(dz), float(dt), 2,
                     stream_handle(None))
        torch.cuda.synchronize()
        for name in FIELDS:
            assert_bits_equal(got[name].cpu().numpy(), ref[name].cpu().numpy(), f"{ni}x{nj}x{nk} {name}")
        np.testing.assert_array_equal(got["asq"].cpu().numpy(), ref["asq"].cpu().numpy())
        np.testing.assert_array_equal(got["qbudget"].cpu().numpy(), ref["qbudget"].cpu().numpy())
        np.testing.assert_array_equal(status.cpu().numpy(), ws.status.cpu().numpy())
        assert int(status[0]) >= 1
        # and both are the oracle's (binary64-rounded transcendentals)
        b = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in host.items()}
        with ok.transcendentals("rounded"):
            ok.kessler_driver(b, dx=250.0, dy=250.0, dz=dz, dt=dt, equation_set=2, budgets=False)
        for name in FIELDS:
            assert_bits_equal(got[name].cpu().numpy(), b[name], f"{ni}x{nj}x{nk} {name} vs oracle")


@pytest.mark.gpu
def test_noalloc_driver_refuses_a_grid_larger_than_its_array(device):
    """4096 x 4096 x 80 needs more than the 64 MiB array: refused before any launch, inputs untouched."""
    import torch

    from gm1_amd import _native
    from gm1_amd._device import ptr, stream_handle
    from gm1_amd.kessler import KESSLER_NOALLOC_MAX_BYTES

    ni, nj, nk = 4096, 4096, 80
    assert int(_native.query("gm1hip_kessler_scratch_bytes", ni, nj, nk)) > KESSLER_NOALLOC_MAX_BYTES
    dummy = torch.zeros(16, dtype=torch.float32, device=device)
    status = torch.zeros(2, dtype=torch.int32, device=device)
    code = getattr(_native.lib(), "gm1hip_kessler_driver_noalloc")(
        *([ptr(dummy)] * 19), ptr(status), None, None, None, None, ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 2,
        stream_handle(None))
    torch.cuda.synchronize()
    assert code != 0
    assert not torch.any(dummy) and not torch.any(status)
