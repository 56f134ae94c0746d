"""Seeded fixtures and numpy<->torch plumbing shared by the tests."""

from __future__ import annotations

from pathlib import Path

import numpy as np

F = np.float32
NG = 3
GOLDEN = Path(__file__).resolve().parent / "golden"


def golden(name: str):
    return np.load(GOLDEN / f"{name}.npz")


def shapes(ni, nj, nk):
    return {
        "u": (nk + 2, nj + 2 * NG, ni + 1 + 2 * NG),
        "v": (nk + 2, nj + 1 + 2 * NG, ni + 2 * NG),
        "w": (nk + 3, nj + 2 * NG, ni + 2 * NG),
        "s": (nk + 2, nj + 2 * NG, ni + 2 * NG),
        "p": (nk + 2, nj + 2, ni + 2),
    }


def metric_shapes(ni, nj, nk):
    from oracle.common import terrain_static_shapes

    return terrain_static_shapes(ni, nj, nk)


def seeded_metrics(rng, ni, nj, nk, *, terrain=True):
    """Positive, mildly varying metrics (same construction as the golden script)."""
    m = {name: np.ones(shape, F) for name, shape in metric_shapes(ni, nj, nk).items()}
    kk = np.arange(nk, dtype=np.float64)
    kf = np.arange(nk + 1, dtype=np.float64)
    m["rho0_3d"][:] = (1.15 * np.exp(-kk / 30.0)[:, None, None] * (1.0 + 0.01 * rng.standard_normal((nk, nj, ni)))).astype(F)
    m["rr0"][:] = (F(1.0) / m["rho0_3d"]).astype(F)
    m["rf0"][:] = (1.18 * np.exp(-kf / 30.0)[:, None, None] * (1.0 + 0.01 * rng.standard_normal((nk + 1, nj, ni)))).astype(F)
    m["rrf0"][:] = (F(1.0) / m["rf0"]).astype(F)
    c2 = (0.5 + 0.05 * rng.standard_normal((nk + 1, nj, ni))).astype(F)
    m["c2"][:] = c2
    m["c1"][:] = (F(1.0) - c2).astype(F)
    if terrain:
        m["dzdx"][:] = (0.01 * rng.standard_normal((nj, ni))).astype(F)
        m["dzdy"][:] = (0.01 * rng.standard_normal((nj, ni))).astype(F)
        m["gz"][:] = (1.0 + 0.02 * rng.standard_normal((nj, ni))).astype(F)
        m["rgz"][:] = (F(1.0) / m["gz"]).astype(F)
        m["gzu"][:] = (1.0 + 0.02 * rng.standard_normal((nj, ni + 1))).astype(F)
        m["rgzu"][:] = (F(1.0) / m["gzu"]).astype(F)
        m["gzv"][:] = (1.0 + 0.02 * rng.standard_normal((nj + 1, ni))).astype(F)
        m["rgzv"][:] = (F(1.0) / m["gzv"]).astype(F)
        m["gxu"][:] = (0.01 * rng.standard_normal((nk + 1, nj, ni + 1))).astype(F)
        m["gyv"][:] = (0.01 * rng.standard_normal((nk + 1, nj + 1, ni))).astype(F)
        m["gx"][:] = (0.01 * rng.standard_normal((nk + 1, nj, ni))).astype(F)
        m["gy"][:] = (0.01 * rng.standard_normal((nk + 1, nj, ni))).astype(F)
    else:
        for name in ("dzdx", "dzdy", "gxu", "gyv", "gx", "gy"):
            m[name][:] = F(0.0)
    return m


def to_dev(d, device):
    import torch

    if isinstance(d, dict):
        return {k: to_dev(v, device) for k, v in d.items()}
    return torch.from_numpy(np.ascontiguousarray(d)).to(device)


def to_host(d):
    if isinstance(d, dict):
        return {k: to_host(v) for k, v in d.items()}
    return d.detach().cpu().numpy()


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def assert_bits_equal(got, want, name=""):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    diff = bits(got) != bits(want)
    if diff.any():
        idx = np.argwhere(diff)
        first = tuple(idx[0])
        raise AssertionError(
            f"{name}: {int(diff.sum())} of {diff.size} words differ; first at {first}: "
            f"got {got[first]!r} want {want[first]!r}"
        )
