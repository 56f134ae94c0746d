"""CM1FCP1 restarts of the large step: the envelope, and resuming word for word.

The envelope follows the reference's forecast checkpoint (rust/cm1-state/src/
forecast.rs): magic, u64 header length, canonical-JSON header with per-carrier
SHA-256, payload of binary32 carriers in contract order.  The carried set is
CM1's restart inventory for this path (cartesian_restart.py:1-60,
cuda_gm1_kessler_large_step.py:86-95).  The resume test is the one
README.restart states: N steps equal k steps, a restart, a restore into a
freshly initialised state, and N - k more steps, bit for bit.
"""

from __future__ import annotations

import json
import struct

import numpy as np
import pytest

from helpers import F, assert_bits_equal, to_host
from gm1_amd import restart as rs
from gm1_amd.fields import CaseConfig
from gm1_amd.large_step import host_initial_state
from oracle import large_step as ol
from oracle import sound as osound
from oracle import stage as ost

FIELDS = ("u3d", "v3d", "w3d", "th3d", "pp3d", "ppx", "ppi", "ua", "va", "wa", "tha", "q3d", "qa", "rho", "prs",
          "rr", "rf", "rain", "prate", "qpten", "qtten", "qvten", "qcten", "radbcw", "radbce", "radbcs", "radbcn")


def moist_case(ni, nj, nk, sides=(2, 2, 2, 2), seed=5):
    return CaseConfig(f"moist-{ni}x{nj}x{nk}", ni, nj, nk, 250.0, 250.0, 250.0, 1.5, 8, tuple(sides), True,
                      "hodograph", ((3.0, ni * 120.0, nj * 130.0, 1000.0),), bubble_radius=(ni * 80.0, 1000.0),
                      noise=0.1, seed=seed, extra={"bubble_rh": 1.03})


# ----------------------------------------------------------------------------- the envelope


def _tiny_contract():
    specs = (rs.ArraySpec("prognostic.a", "a field", "1", "mass", (2, 3)),
             rs.ArraySpec("prognostic.b", "b field", "K", "mass", (4,)))
    return rs.Contract("test/v1", "11" * 32, "22" * 32, specs, (("nstep", "u64"), ("mass1", "f64-bits")))


def _tiny_blob():
    arrays = {"prognostic.a": np.arange(6, dtype=np.float32).reshape(2, 3),
              "prognostic.b": np.array([1.5, -0.0, np.inf, 3e-42], np.float32)}
    return rs.encode_forecast_checkpoint(_tiny_contract(), {"step": 7, "model_time": 10.5, "dt": 1.5}, arrays,
                                         {"nstep": 7, "mass1": 1.25e13}), arrays


def test_envelope_layout_is_the_forecast_checkpoint():
    blob, arrays = _tiny_blob()
    assert blob[:8] == b"CM1FCP1\0"
    hlen = struct.unpack("<Q", blob[8:16])[0]
    header = json.loads(blob[16 : 16 + hlen])
    assert sorted(header) == ["arrays", "authority_sha256", "capability", "claim_limit", "clock", "component_blobs",
                              "contract", "contract_sha256", "live_state_wiring", "payload_nbytes", "payload_order",
                              "payload_sha256", "representation", "scheduler_scalars", "schema"]
    assert header["schema"] == "gpuwm-cm1.forecast-checkpoint/generic-v1"
    assert header["payload_order"] == ["array:prognostic.a", "array:prognostic.b"]
    assert header["clock"] == {"dt_f64_bits": "3ff8000000000000", "model_time_f64_bits": "4025000000000000",
                               "step": 7}
    assert header["contract"]["array_specs"][0]["bounds"] == [1, 2, 1, 3]
    # the payload is the carriers' little-endian words back to back
    payload = blob[16 + hlen :]
    assert payload == arrays["prognostic.a"].astype("<f4").tobytes() + arrays["prognostic.b"].astype("<f4").tobytes()
    assert header["arrays"][1] == {"name": "prognostic.b", "nbytes": 16, "offset": 24,
                                   "sha256": __import__("hashlib").sha256(payload[24:]).hexdigest()}
    # canonical: sorted keys, no whitespace between tokens, ASCII
    assert blob[16 : 16 + hlen] == json.dumps(header, sort_keys=True, separators=(",", ":")).encode("ascii")
    d = rs.decode_forecast_checkpoint(blob)
    for k, v in arrays.items():
        assert_bits_equal(d.arrays[k], v, k)
    assert d.scalars == {"nstep": 7, "mass1": 1.25e13} and d.clock == {"step": 7, "model_time": 10.5, "dt": 1.5}


@pytest.mark.parametrize("where", ["payload", "header", "magic", "truncated"])
def test_a_damaged_envelope_is_refused(where):
    blob, _ = _tiny_blob()
    b = bytearray(blob)
    hlen = struct.unpack("<Q", blob[8:16])[0]
    if where == "payload":
        b[16 + hlen + 5] ^= 1
    elif where == "header":
        i = bytes(b).index(b"prognostic.b", 16)
        b[i] = ord("P")
    elif where == "magic":
        b[0] = ord("X")
    else:
        b = b[:-4]
    with pytest.raises(rs.RestartError):
        rs.decode_forecast_checkpoint(bytes(b))


# ----------------------------------------------------------------------------- resuming on the oracle


def _oracle_steps(a, host, n, mass1):
    cfg = host["cfg"]
    prof = osound.terrain_coordinate_profiles(cfg.nk, cfg.dz)
    c = {"ni": cfg.ni, "nj": cfg.nj, "nk": cfg.nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": cfg.lateral_sides}
    for _ in range(n):
        ol.large_step(a, host["metrics"], host["base"], prof, c, zt=F(cfg.nk * cfg.dz), mass1=mass1, budgets=True)


def test_oracle_resumes_word_for_word(tmp_path):
    """3 steps = 1 step + restart + restore into a fresh state + 2 steps, every carried and rebuilt array."""
    from oracle import kessler as ok

    cfg = moist_case(16, 14, 12)
    host = host_initial_state(cfg)
    a0 = host["arrays"]
    mass1 = ok.dry_mass(a0["rho"], a0["ruh"], a0["rvh"], a0["rmh"])
    full = {k: v.copy() for k, v in a0.items()}
    _oracle_steps(full, host, 3, mass1)
    part = {k: v.copy() for k, v in a0.items()}
    _oracle_steps(part, host, 1, mass1)
    blob = rs.snapshot_restart_words(part, cfg, step=1, mass1=mass1)
    receipt = rs.publish_restart(blob, tmp_path, restart_number=1)
    restored = rs.read_cartesian_restart(receipt["directory"])
    assert restored.step == 1
    fresh = {k: v.copy() for k, v in host_initial_state(cfg)["arrays"].items()}
    s = rs.restore_restart_words(fresh, cfg, restored.decoded)
    fresh["rr"][...], fresh["rf"][...] = ost.solve3_density(fresh["rho"], host["metrics"]["c1"], host["metrics"]["c2"])
    _oracle_steps(fresh, host, 2, s["mass1"])
    for name in FIELDS:
        assert_bits_equal(fresh[name], full[name], name)
    np.testing.assert_array_equal(fresh["asq"], full["asq"])
    np.testing.assert_array_equal(fresh["qbudget"], full["qbudget"])


def test_restore_refuses_another_configuration(tmp_path):
    cfg = moist_case(16, 14, 12)
    host = host_initial_state(cfg)
    blob = rs.snapshot_restart_words(host["arrays"], cfg, step=0, mass1=1.0)
    other = moist_case(16, 14, 12, sides=(1, 1, 1, 1))
    with pytest.raises(rs.RestartError, match="another configuration"):
        rs.restore_restart_words({k: v.copy() for k, v in host["arrays"].items()}, other,
                                 rs.decode_forecast_checkpoint(blob))
    rs.publish_restart(blob, tmp_path, restart_number=2)
    with pytest.raises(rs.RestartError, match="never overwritten"):
        rs.publish_restart(blob, tmp_path, restart_number=2)


# ----------------------------------------------------------------------------- resuming on the device


@pytest.mark.gpu
@pytest.mark.parametrize("stretched", [False, True])
def test_gpu_resume_is_word_for_word(device, tmp_path, stretched):
    """4 device steps = 2 steps + write_cartesian_restart + restore into a fresh state + 2 steps."""
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, launch_cuda_gm1_large_step

    if stretched:
        cfg = CaseConfig("stretched-24x20x16", 24, 20, 16, 100.0, 100.0, 100.0, 0.6, 10, (2, 2, 2, 2), True,
                         "hodograph", ((3.0, 1200.0, 1000.0, 600.0),), bubble_radius=(720.0, 500.0), noise=0.1,
                         seed=7, extra={"bubble_rh": 1.03}, stretched=(1600.0, 100.0, 700.0, 20.0, 180.0))
    else:
        cfg = moist_case(24, 20, 16)
    print(f"restart resume {cfg.name}", flush=True)
    host = host_initial_state(cfg)
    full = create_gm1_large_step_state(host, device)
    for _ in range(4):
        launch_cuda_gm1_large_step(full)
    part = create_gm1_large_step_state(host, device)
    for _ in range(2):
        launch_cuda_gm1_large_step(part)
    receipt = rs.write_cartesian_restart(part, tmp_path)
    restored = rs.read_cartesian_restart(receipt["directory"])
    fresh = create_gm1_large_step_state(host_initial_state(cfg), device)
    rs.restore_cartesian_restart(fresh, restored)
    assert fresh.step_index == 2 and fresh.mass1 == part.mass1
    for _ in range(2):
        launch_cuda_gm1_large_step(fresh)
    torch.cuda.synchronize()
    for name in FIELDS:
        assert_bits_equal(to_host(fresh.a[name]), to_host(full.a[name]), name)
    np.testing.assert_array_equal(to_host(fresh.a["asq"]), to_host(full.a["asq"]))
    np.testing.assert_array_equal(to_host(fresh.a["qbudget"]), to_host(full.a["qbudget"]))
    assert fresh.step_index == full.step_index == 4
