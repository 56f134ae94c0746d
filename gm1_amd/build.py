"""Ahead-of-time build of the gfx950 HIP library ``libgm1hip.so``.

Every ``csrc/*.hip`` translation unit is compiled by ``hipcc`` into an object
and the objects are linked into one shared library next to this file.  The
flags fix the arithmetic contract the kernels rely on: no floating-point
contraction, no fast math, IEEE denormals, correctly rounded binary32 divide
and sqrt, and no packed-FP32 instructions.
"""

from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
BUILD_DIR = PACKAGE_DIR / "_build"
LIBRARY = PACKAGE_DIR / "libgm1hip.so"
RECEIPT = PACKAGE_DIR / "libgm1hip.receipt.json"
ARCH = "gfx950"

HIPCC_FLAGS = (
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
)


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD_DIR / (src.stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{proc.stderr[-4000:]}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source and link ``libgm1hip.so`` (always rebuilds)."""

    BUILD_DIR.mkdir(exist_ok=True)
    srcs = sources()
    if not srcs:
        raise RuntimeError("no HIP sources under gm1_amd/csrc")
    jobs = jobs or min(8, len(srcs))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr[-4000:]}")
    os.replace(tmp, LIBRARY)
    digest = hashlib.sha256()
    for src in srcs + sorted(CSRC.glob("*.h")):
        digest.update(src.name.encode())
        digest.update(src.read_bytes())
    RECEIPT.write_text(json.dumps({
        "arch": ARCH,
        "flags": list(HIPCC_FLAGS),
        "sources_sha256": digest.hexdigest(),
        "sources": [s.name for s in srcs],
    }, indent=1))
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
