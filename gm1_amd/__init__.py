"""MI355X-native CM1 dynamical-core large step (gfx950 HIP kernels, C ABI).

Operator modules mirror the reference's GM1 entry points; every one runs in
``libgm1hip.so`` and raises when that library is missing.
"""

from . import adjoint_sound, adjoint_step, advection5, autotune, boundary, ensemble, kessler, large_step, momentum_advection5, psolver3, scalar_moist_rk, sound3_periodic, stretched, timing  # noqa: F401,E501  (register C-ABI signatures)

__all__ = ["adjoint_sound", "adjoint_step", "advection5", "boundary", "kessler", "large_step", "momentum_advection5", "psolver3", "scalar_moist_rk", "timing"]
