"""CM1FCP1 restart checkpoints of the large step's carried state (host side).

Host mirror of the reference's restart pair for this path
(``cartesian_restart.py``: ``write_cartesian_restart`` :1341,
``read_cartesian_restart`` :1516, ``restore_cartesian_restart`` :1866) and of
the envelope its Rust publisher writes (``rust/cm1-state/src/forecast.rs``):

* the artifact is one ``CM1FCP1`` forecast checkpoint: eight magic bytes
  ``CM1FCP1\\0``, the header length as a little-endian u64, the header as
  canonical JSON (keys sorted, no whitespace, non-ASCII escaped as
  ``\\uXXXX``), then the payload: every carrier's raw little-endian binary32
  words back to back in contract order (forecast.rs:1603-1690);
* the header records each carrier's offset, byte count and SHA-256, the
  payload's SHA-256, the contract (array specs with CM1 bounds, scheduler
  scalar specs) and its SHA-256, and the clock (step, model time and dt as
  binary64 bit patterns);
* what is carried is CM1's restart inventory for this path
  (restart_write.F:692-840 as cartesian_restart.py:1-60 lists it):
  rho, prs, ua, va, wa, ppi, tha, ppx, the moisture species qa, the eqtset=2
  held rates qpten/qtten/qvten/qcten, the Kessler rain and prate
  (cuda_gm1_kessler_large_step.py:86-95), the radiative lateral speeds
  radbcw/e/s/n; and as scheduler scalars the step, the model time, the
  dry-mass reference mass1 and the binary64 budget accumulators asq and
  qbudget.  Everything else is rebuilt at restore exactly as solve3 leaves
  it (u3d = ua, v3d = va, w3d = wa, th3d = tha, q3d = qa, pp3d = ppi,
  rr = 1/rho, rf by c1/c2), or is rewritten by the step before it is read.

A restored state continues the integration word for word.
"""

from __future__ import annotations

import hashlib
import json
import os
import struct
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Mapping, Sequence

import numpy as np

MAGIC = b"CM1FCP1\0"
PREFIX_BYTES = 16
MAX_HEADER_BYTES = 1024 * 1024
SCHEMA = "gpuwm-cm1.forecast-checkpoint/generic-v1"
CAPABILITY = "caller-contracted-complete-step-forecast-state"
REPRESENTATION = ("arrays are raw IEEE-754 binary32 little-endian in caller-contracted dimension order; "
                  "component blobs are opaque bytes")
PROFILE_ID = "gm1_amd.large-step-restart/v1"
RESTART_DIRECTORY_FORMAT = "restart-{number:06d}"
ENVELOPE_FILENAME = "restart.cm1fcp"
RECEIPT_FILENAME = "restart-publication.json"

# carrier name -> (state key, meaning, units); CM1's restart_write.F order
_CARRIERS = (
    ("prognostic.rho", "rho", "dry-air density", "kg m-3"),
    ("prognostic.prs", "prs", "pressure", "Pa"),
    ("prognostic.ua", "ua", "x wind at the large step", "m s-1"),
    ("prognostic.va", "va", "y wind at the large step", "m s-1"),
    ("prognostic.wa", "wa", "z wind at the large step", "m s-1"),
    ("prognostic.ppi", "ppi", "Exner perturbation at the large step", "1"),
    ("prognostic.tha", "tha", "potential temperature perturbation at the large step", "K"),
    ("prognostic.ppx", "ppx", "Exner tendency of the divergence filter", "1"),
    ("prognostic.qa", "qa", "moisture mixing ratios qv, qc, qr", "kg kg-1"),
    ("held.qpten", "qpten", "held Exner rate of the Kessler transition", "s-1"),
    ("held.qtten", "qtten", "held theta rate of the Kessler transition", "K s-1"),
    ("held.qvten", "qvten", "held qv rate of the Kessler transition", "kg kg-1 s-1"),
    ("held.qcten", "qcten", "held qc rate of the Kessler transition", "kg kg-1 s-1"),
    ("persistent.rain", "rain", "accumulated surface rain", "cm"),
    ("persistent.prate", "prate", "surface precipitation rate", "kg m-2 s-1"),
    ("lateral.radbcw", "radbcw", "west radiative phase speed", "m s-1"),
    ("lateral.radbce", "radbce", "east radiative phase speed", "m s-1"),
    ("lateral.radbcs", "radbcs", "south radiative phase speed", "m s-1"),
    ("lateral.radbcn", "radbcn", "north radiative phase speed", "m s-1"),
)
_MOIST_ONLY = {"prognostic.qa", "held.qpten", "held.qtten", "held.qvten", "held.qcten", "persistent.rain",
               "persistent.prate"}
# what solve3 leaves equal to a carried array (cuda_gm1_forecast.py:1414-1457)
_REBUILT_COPIES = (("u3d", "ua"), ("v3d", "va"), ("w3d", "wa"), ("th3d", "tha"), ("q3d", "qa"), ("pp3d", "ppi"))


class RestartError(ValueError):
    """A restart artifact that does not prove itself complete, or one this state cannot take."""


# ----------------------------------------------------------------------------- canonical JSON


def _json_string(value: str) -> str:
    out = ['"']
    for ch in value:
        c = ord(ch)
        if ch == '"':
            out.append('\\"')
        elif ch == "\\":
            out.append("\\\\")
        elif ch == "\b":
            out.append("\\b")
        elif ch == "\f":
            out.append("\\f")
        elif ch == "\n":
            out.append("\\n")
        elif ch == "\r":
            out.append("\\r")
        elif ch == "\t":
            out.append("\\t")
        elif c < 0x20 or 0x7F <= c <= 0xFFFF:
            out.append(f"\\u{c:04x}")
        elif c > 0xFFFF:
            s = c - 0x10000
            out.append(f"\\u{0xD800 + (s >> 10):04x}\\u{0xDC00 + (s & 0x3FF):04x}")
        else:
            out.append(ch)
    out.append('"')
    return "".join(out)


class Number(str):
    """A JSON number carried as its exact decimal text (forecast.rs numbers are decimal strings)."""


def canonical_json(value: Any) -> bytes:
    """forecast.rs / json.rs canonical form: sorted keys, no whitespace, ASCII with \\u escapes."""

    def enc(v: Any) -> str:
        if v is None:
            return "null"
        if v is True:
            return "true"
        if v is False:
            return "false"
        if isinstance(v, Number):
            return str(v)
        if isinstance(v, bool):
            raise TypeError("unreachable")
        if isinstance(v, int):
            return str(v)
        if isinstance(v, str):
            return _json_string(v)
        if isinstance(v, (list, tuple)):
            return "[" + ",".join(enc(x) for x in v) + "]"
        if isinstance(v, Mapping):
            return "{" + ",".join(f"{_json_string(k)}:{enc(v[k])}" for k in sorted(v)) + "}"
        raise TypeError(f"no canonical JSON for {type(v).__name__}")

    return enc(value).encode("ascii")


def f64_bits(value: float) -> str:
    return f"{struct.unpack('<Q', struct.pack('<d', float(value)))[0]:016x}"


def from_f64_bits(text: str) -> float:
    if len(text) != 16 or any(ch not in "0123456789abcdef" for ch in text):
        raise RestartError(f"f64 bits must be 16 lowercase hexadecimal characters, got {text!r}")
    return struct.unpack("<d", struct.pack("<Q", int(text, 16)))[0]


def _sha256(data: bytes | memoryview) -> str:
    return hashlib.sha256(data).hexdigest()


# ----------------------------------------------------------------------------- the generic envelope


@dataclass(frozen=True)
class ArraySpec:
    name: str
    meaning: str
    units: str
    stagger: str
    shape: tuple[int, ...]

    def record(self) -> dict[str, Any]:
        return {"bounds": [v for extent in self.shape for v in (1, int(extent))], "meaning": self.meaning,
                "name": self.name, "shape": [int(x) for x in self.shape], "stagger": self.stagger,
                "units": self.units}


@dataclass(frozen=True)
class Contract:
    profile_id: str
    config_sha256: str
    execution_identity_sha256: str
    arrays: tuple[ArraySpec, ...]
    scalars: tuple[tuple[str, str], ...]  # (name, type: "i64" | "u64" | "f64-bits" | "bool")

    def record(self) -> dict[str, Any]:
        return {"array_specs": [a.record() for a in self.arrays], "component_blob_specs": [],
                "config_sha256": self.config_sha256, "execution_identity_sha256": self.execution_identity_sha256,
                "profile_id": self.profile_id,
                "scheduler_scalar_specs": [{"name": n, "type": t} for n, t in self.scalars]}

    @property
    def sha256(self) -> str:
        return _sha256(canonical_json(self.record()))


def _scalar_value(kind: str, value: Any) -> Any:
    if kind == "f64-bits":
        return f64_bits(value)
    if kind in ("i64", "u64"):
        v = int(value)
        if kind == "u64" and not 0 <= v < 1 << 64:
            raise RestartError(f"u64 scalar out of range: {v}")
        return Number(str(v))
    if kind == "bool":
        return bool(value)
    raise RestartError(f"unknown scheduler scalar type {kind!r}")


def encode_forecast_checkpoint(contract: Contract, clock: Mapping[str, Any], arrays: Mapping[str, np.ndarray],
                               scalars: Mapping[str, Any]) -> bytes:
    """prefix + header + payload of one CM1FCP1 checkpoint (forecast.rs:1603-1690)."""
    if set(arrays) != {a.name for a in contract.arrays}:
        raise RestartError(f"carriers {sorted(arrays)} are not the contract's {[a.name for a in contract.arrays]}")
    if set(scalars) != {n for n, _ in contract.scalars}:
        raise RestartError(f"scalars {sorted(scalars)} are not the contract's {[n for n, _ in contract.scalars]}")
    chunks, headers, order = [], [], []
    offset = 0
    for spec in contract.arrays:
        a = np.asarray(arrays[spec.name])
        if a.dtype != np.float32 or tuple(a.shape) != tuple(spec.shape):
            raise RestartError(f"carrier {spec.name} is {a.dtype}{tuple(a.shape)}, the contract says "
                               f"float32{tuple(spec.shape)}")
        raw = np.ascontiguousarray(a).astype("<f4", copy=False).tobytes()
        headers.append({"name": spec.name, "nbytes": Number(str(len(raw))), "offset": Number(str(offset)),
                        "sha256": _sha256(raw)})
        order.append(f"array:{spec.name}")
        chunks.append(raw)
        offset += len(raw)
    payload = b"".join(chunks)
    header = {
        "arrays": headers,
        "authority_sha256": _sha256(b"gm1_amd CM1FCP1 restart of the CM1 large step"),
        "capability": CAPABILITY,
        "claim_limit": "the large step's carried state as restart_write.F names it for this path",
        "clock": {"dt_f64_bits": f64_bits(clock["dt"]), "model_time_f64_bits": f64_bits(clock["model_time"]),
                  "step": Number(str(int(clock["step"])))},
        "component_blobs": [],
        "contract": contract.record(),
        "contract_sha256": contract.sha256,
        "live_state_wiring": "gm1_amd device state restored through restore_cartesian_restart",
        "payload_nbytes": Number(str(len(payload))),
        "payload_order": order,
        "payload_sha256": _sha256(payload),
        "representation": REPRESENTATION,
        "scheduler_scalars": [{"name": n, "type": t, "value": _scalar_value(t, scalars[n])}
                              for n, t in contract.scalars],
        "schema": SCHEMA,
    }
    hbytes = canonical_json(header)
    if len(hbytes) > MAX_HEADER_BYTES:
        raise RestartError(f"header of {len(hbytes)} bytes exceeds {MAX_HEADER_BYTES}")
    return MAGIC + struct.pack("<Q", len(hbytes)) + hbytes + payload


@dataclass
class DecodedCheckpoint:
    header: dict[str, Any]
    arrays: dict[str, np.ndarray]
    scalars: dict[str, Any]
    clock: dict[str, Any]
    header_sha256: str


def decode_forecast_checkpoint(blob: bytes | memoryview) -> DecodedCheckpoint:
    """Validate and split one CM1FCP1 checkpoint: magic, header size, payload size, every digest."""
    blob = memoryview(blob)
    if len(blob) < PREFIX_BYTES or bytes(blob[:8]) != MAGIC:
        raise RestartError("the artifact does not begin with the CM1FCP1 magic")
    hlen = struct.unpack("<Q", bytes(blob[8:16]))[0]
    if not 0 < hlen <= MAX_HEADER_BYTES or PREFIX_BYTES + hlen > len(blob):
        raise RestartError(f"the artifact declares a {hlen}-byte header it does not hold")
    hbytes = bytes(blob[PREFIX_BYTES : PREFIX_BYTES + hlen])
    header = json.loads(hbytes)
    if canonical_json(_renumber(header)) != hbytes:
        raise RestartError("the header is not in canonical form")
    if header.get("schema") != SCHEMA:
        raise RestartError(f"schema {header.get('schema')!r} is not {SCHEMA!r}")
    payload = blob[PREFIX_BYTES + hlen :]
    if len(payload) != int(header["payload_nbytes"]):
        raise RestartError(f"payload of {len(payload)} bytes, the header declares {header['payload_nbytes']}")
    if _sha256(payload) != header["payload_sha256"]:
        raise RestartError("the payload's SHA-256 is not the header's")
    if _sha256(canonical_json(_renumber(header["contract"]))) != header["contract_sha256"]:
        raise RestartError("the contract's SHA-256 is not the header's")
    names = [s["name"] for s in header["contract"]["array_specs"]]
    if [row["name"] for row in header["arrays"]] != names:
        raise RestartError("the header's carriers are not the contract's, in the contract's order")
    if header["payload_order"] != [f"array:{n}" for n in names]:
        raise RestartError("the payload order is not the contract's carrier order")
    specs = {s["name"]: s for s in header["contract"]["array_specs"]}
    arrays = {}
    expect = 0
    for row in header["arrays"]:
        spec = specs[row["name"]]
        off, n = int(row["offset"]), int(row["nbytes"])
        if off != expect:
            raise RestartError(f"carrier {row['name']} starts at {off}, not at {expect}")
        expect = off + n
        raw = payload[off : off + n]
        if len(raw) != n or _sha256(raw) != row["sha256"]:
            raise RestartError(f"carrier {row['name']} does not match its SHA-256")
        shape = tuple(int(x) for x in spec["shape"])
        if n != 4 * int(np.prod(shape)):
            raise RestartError(f"carrier {row['name']} holds {n} bytes for shape {shape}")
        arrays[row["name"]] = np.frombuffer(raw, dtype="<f4").astype(np.float32).reshape(shape)
    scalars = {}
    for row in header["scheduler_scalars"]:
        kind, v = row["type"], row["value"]
        scalars[row["name"]] = from_f64_bits(v) if kind == "f64-bits" else (bool(v) if kind == "bool" else int(v))
    c = header["clock"]
    clock = {"step": int(c["step"]), "model_time": from_f64_bits(c["model_time_f64_bits"]),
             "dt": from_f64_bits(c["dt_f64_bits"])}
    return DecodedCheckpoint(header, arrays, scalars, clock, _sha256(hbytes))


def _renumber(value: Any) -> Any:
    """json.loads' ints back to exact decimal Numbers, so the canonical bytes can be re-derived."""
    if isinstance(value, bool) or value is None:
        return value
    if isinstance(value, int):
        return Number(str(value))
    if isinstance(value, list):
        return [_renumber(v) for v in value]
    if isinstance(value, dict):
        return {k: _renumber(v) for k, v in value.items()}
    return value


# ----------------------------------------------------------------------------- the large step's restart


def _to_host(x: Any) -> np.ndarray:
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _stagger(key: str) -> str:
    return {"ua": "padded-u", "va": "padded-v", "wa": "padded-w", "ppi": "padded-pressure",
            "ppx": "padded-pressure", "qa": "mass-species", "rain": "surface", "prate": "surface"}.get(
        key, "lateral" if key.startswith("radbc") else "mass")


def _config_record(cfg: Any) -> dict[str, Any]:
    rec = {}
    for name in ("name", "ni", "nj", "nk", "dx", "dy", "dz", "dt", "nsound", "lateral_sides", "moist", "wind",
                 "stretched"):
        v = getattr(cfg, name, None)
        if isinstance(v, float):
            v = f64_bits(v)
        elif isinstance(v, tuple):
            v = [f64_bits(x) if isinstance(x, float) else x for x in v]
        rec[name] = v
    return rec


def _execution_identity() -> str:
    receipt = Path(__file__).resolve().parent / "libgm1hip.receipt.json"
    return _sha256(receipt.read_bytes()) if receipt.exists() else _sha256(b"gm1_amd")


def restart_contract(arrays: Mapping[str, Any], cfg: Any, *, moist: bool, nasq: int = 3, nqbudget: int = 10) -> Contract:
    specs = []
    for name, key, meaning, units in _CARRIERS:
        if name in _MOIST_ONLY and not moist:
            continue
        shape = tuple(int(x) for x in arrays[key].shape)
        specs.append(ArraySpec(name, meaning, units, _stagger(key), shape))
    scalars = [("nstep", "u64"), ("mtime", "f64-bits"), ("nrst", "u64")]
    if moist:
        scalars.append(("mass1", "f64-bits"))
        scalars += [(f"asq{i}", "f64-bits") for i in range(nasq)]
        scalars += [(f"qbudget{i}", "f64-bits") for i in range(nqbudget)]
    return Contract(PROFILE_ID, _sha256(canonical_json(_config_record(cfg))), _execution_identity(), tuple(specs),
                    tuple(scalars))


def snapshot_restart_words(arrays: Mapping[str, Any], cfg: Any, *, step: int, mass1: float | None,
                           restart_number: int = 1, model_time: float | None = None) -> bytes:
    """The CM1FCP1 bytes of one state (host NumPy arrays or device tensors) after ``step`` large steps."""
    moist = "q3d" in arrays
    contract = restart_contract(arrays, cfg, moist=moist)
    key = {name: k for name, k, _, _ in _CARRIERS}
    host = {spec.name: _to_host(arrays[key[spec.name]]) for spec in contract.arrays}
    scalars: dict[str, Any] = {"nstep": step, "nrst": restart_number,
                               "mtime": float(step) * float(cfg.dt) if model_time is None else model_time}
    if moist:
        if mass1 is None:
            raise RestartError("a moist restart carries the dry-mass reference mass1; none was given")
        scalars["mass1"] = float(mass1)
        asq, qb = _to_host(arrays["asq"]).astype(np.float64), _to_host(arrays["qbudget"]).astype(np.float64)
        scalars.update({f"asq{i}": float(v) for i, v in enumerate(asq)})
        scalars.update({f"qbudget{i}": float(v) for i, v in enumerate(qb)})
    clock = {"step": step, "model_time": scalars["mtime"], "dt": float(cfg.dt)}
    return encode_forecast_checkpoint(contract, clock, host, scalars)


def write_cartesian_restart(st: Any, directory: Path | str, *, restart_number: int = 1,
                            model_time: float | None = None) -> dict[str, Any]:
    """Write one restart of a device large-step state into ``directory/restart-NNNNNN`` (atomically).

    The device is synchronised once to read the state; the envelope is written
    into a staging directory and renamed into place only once it is on disk.
    """
    import torch

    if st.stream is not None:
        st.stream.synchronize()
    torch.cuda.synchronize()
    blob = snapshot_restart_words(st.a, st.cfg, step=st.step_index, mass1=st.mass1, restart_number=restart_number,
                                  model_time=model_time)
    return publish_restart(blob, directory, restart_number=restart_number)


def publish_restart(blob: bytes, directory: Path | str, *, restart_number: int = 1) -> dict[str, Any]:
    """Write the envelope and its receipt into a new ``restart-NNNNNN`` directory, renamed into place."""
    base = Path(directory)
    base.mkdir(parents=True, exist_ok=True)
    final = base / RESTART_DIRECTORY_FORMAT.format(number=restart_number)
    if final.exists():
        raise RestartError(f"{final} exists; a restart is never overwritten")
    staging = base / (final.name + ".staging")
    staging.mkdir()
    decoded = decode_forecast_checkpoint(blob)
    receipt = {"profile_id": PROFILE_ID, "envelope": ENVELOPE_FILENAME, "artifact_sha256": _sha256(blob),
               "header_sha256": decoded.header_sha256, "payload_sha256": decoded.header["payload_sha256"],
               "contract_sha256": decoded.header["contract_sha256"], "step": decoded.clock["step"],
               "artifact_nbytes": len(blob)}
    for name, data in ((ENVELOPE_FILENAME, blob), (RECEIPT_FILENAME, json.dumps(receipt, sort_keys=True).encode())):
        with open(staging / name, "xb") as fh:
            fh.write(data)
            fh.flush()
            os.fsync(fh.fileno())
    os.rename(staging, final)
    return {**receipt, "directory": str(final)}


@dataclass
class RestoredCartesianRestart:
    directory: Path
    decoded: DecodedCheckpoint
    receipt: dict[str, Any] = field(default_factory=dict)

    @property
    def step(self) -> int:
        return self.decoded.clock["step"]


def read_cartesian_restart(directory: Path | str) -> RestoredCartesianRestart:
    """Read and validate one restart directory (envelope digests against the receipt and the header)."""
    d = Path(directory)
    blob = (d / ENVELOPE_FILENAME).read_bytes()
    receipt = json.loads((d / RECEIPT_FILENAME).read_text())
    if _sha256(blob) != receipt.get("artifact_sha256"):
        raise RestartError(f"{d / ENVELOPE_FILENAME} is not the artifact its receipt names")
    decoded = decode_forecast_checkpoint(blob)
    if decoded.header["contract"]["profile_id"] != PROFILE_ID:
        raise RestartError(f"profile {decoded.header['contract']['profile_id']!r} is not {PROFILE_ID!r}")
    return RestoredCartesianRestart(d, decoded, receipt)


def restore_restart_words(arrays: dict[str, Any], cfg: Any, decoded: DecodedCheckpoint) -> dict[str, Any]:
    """Load a decoded checkpoint into a state's arrays (NumPy or device tensors); returns its scalars.

    Refuses a checkpoint written for another configuration or another carrier
    set before touching any array.
    """
    moist = "q3d" in arrays
    contract = restart_contract(arrays, cfg, moist=moist)
    if decoded.header["contract"]["config_sha256"] != contract.config_sha256:
        raise RestartError("the restart was written for another configuration")
    want = {s.name: s.shape for s in contract.arrays}
    got = {k: tuple(v.shape) for k, v in decoded.arrays.items()}
    if want != got:
        raise RestartError(f"the restart carries {got}, this state needs {want}")
    key = {name: k for name, k, _, _ in _CARRIERS}

    def put(dst_key: str, src: np.ndarray) -> None:
        dst = arrays[dst_key]
        if hasattr(dst, "copy_"):
            import torch

            dst.copy_(torch.from_numpy(np.ascontiguousarray(src)).to(dst.device))
        else:
            dst[...] = src

    for name, values in decoded.arrays.items():
        put(key[name], values)
    for dst, src in _REBUILT_COPIES:
        if dst in arrays and key.get(f"prognostic.{src}") is not None and f"prognostic.{src}" in decoded.arrays:
            put(dst, decoded.arrays[f"prognostic.{src}"])
    s = decoded.scalars
    if moist:
        put("asq", np.array([s[f"asq{i}"] for i in range(len(_to_host(arrays["asq"])))], np.float64))
        put("qbudget", np.array([s[f"qbudget{i}"] for i in range(len(_to_host(arrays["qbudget"])))], np.float64))
    return s


def restore_cartesian_restart(st: Any, restored: RestoredCartesianRestart) -> None:
    """Upload a restart into a device large-step state; the next step continues the saved run word for word.

    rr and rf are rebuilt from rho by the solve3 density pass on the device.
    """
    import torch

    from . import _native
    from ._device import ptr, stream_handle

    torch.cuda.synchronize()
    s = restore_restart_words(st.a, st.cfg, restored.decoded)
    st.step_index = int(s["nstep"])
    if "q3d" in st.a:
        st.mass1 = float(s["mass1"])
        cfg, a, m = st.cfg, st.a, st.metrics
        _native.call("gm1hip_solve3_density", ptr(a["rho"]), ptr(m["c1"]), ptr(m["c2"]), ptr(a["rr"]), ptr(a["rf"]),
                     cfg.ni, cfg.nj, cfg.nk, stream_handle(st.stream))
    torch.cuda.synchronize()
