"""Record the device's dry-mass ledger over the 600-step moist run (run on a HIP device).

    python tests/golden/record_long_run_device.py OUT.npz

Steps the 48x48x40 moist case (``tests/test_configs.py``'s
``long_run_case``) 600 large steps on ``cuda:0`` and writes, per step, the
binary32 Exner correction the device applied and its ledger (the dry mass
before the correction and the Exner sum, binary64), the device's ledger
``mass1``, and the digest of the initial carriers.  ``make_long_run.py``
replays these corrections in the oracle after checking that its own ledger
agrees, so that the oracle's run and the device's differ in nothing but
the order of the ledger's sums.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
for p in (ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

STEPS = 600


def main(out: str) -> None:
    import torch

    from gm1_amd.large_step import create_gm1_large_step_state, host_initial_state, launch_cuda_gm1_large_step
    from test_configs import carrier_digest, long_run_case

    cfg = long_run_case()
    host = host_initial_state(cfg)
    initial = carrier_digest(host["arrays"])
    st = create_gm1_large_step_state(host, "cuda:0")
    del host
    corrections = np.zeros(STEPS, np.float32)
    ledger = np.zeros((STEPS, 2), np.float64)
    for step in range(STEPS):
        launch_cuda_gm1_large_step(st)
        corrections[step] = st.dry_mass.correction[0].item()
        ledger[step] = st.dry_mass.ledger.cpu().numpy()
        if (step + 1) % 100 == 0:
            print(f"step {step + 1}", flush=True)
    torch.cuda.synchronize()
    np.savez_compressed(out, corrections=corrections, ledger=ledger, mass1=np.float64(st.mass1),
                        initial_digest=np.array(initial))
    print("wrote", out, "initial digest", initial)


if __name__ == "__main__":
    main(sys.argv[1])
