"""Live per-kernel GPU time (HIP event pairs recorded around each launch).

``arm(kernel, capacity)`` makes the library bracket every launch of one
kernel with events on the launch stream; ``collect()`` waits for them and
returns ``(total_ms, launches)``.  Used for throughput and roofline reports.
"""

from __future__ import annotations

import ctypes

from . import _native
from ._native import I, signature

signature("gm1hip_kernel_timer_arm", I, I)
signature("gm1hip_kernel_timer_collect", ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int))

KERNELS = {
    "k_snd_column": 1,
    "k_snd_pgrad": 2,
    "k_scalar_adv5": 3,
    "k_mom_advu": 4,
    "k_mom_advv": 5,
    "k_mom_advw": 6,
    "k_snd_small_step": 7,
}


def arm(kernel: str, capacity: int) -> None:
    if kernel not in KERNELS:
        raise ValueError(f"no timer hook for kernel {kernel!r}; choose from {sorted(KERNELS)}")
    _native.call("gm1hip_kernel_timer_arm", KERNELS[kernel], int(capacity))


def collect() -> tuple[float, int]:
    total = ctypes.c_double(0.0)
    count = ctypes.c_int(0)
    _native.call("gm1hip_kernel_timer_collect", ctypes.byref(total), ctypes.byref(count))
    return float(total.value), int(count.value)


def disarm() -> None:
    _native.call("gm1hip_kernel_timer_arm", 0, 0)
