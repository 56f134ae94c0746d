"""CPU oracle for the CM1 large-step path -- TEST INFRASTRUCTURE ONLY.

This package restates, in vectorised NumPy binary32 arithmetic, the
reference's own float32 CPU transcriptions of the operators on the CM1
dynamical-core large step (GM1's ``cpu_*_inplace`` functions).  Each module
cites the reference file and lines it follows.  Every binary32 ``+ - * /`` is
a single rounded NumPy float32 array operation in the association order the
reference spells, WENO weights are formed in binary64, and no operation is
contracted.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker or as
the timed CPU baseline.  The product path (``gm1_amd``) never imports it and
fails loudly when its HIP library is missing.

Pinning (what checks this restatement, besides the GPU comparison):
  * golden vectors written by ``tests/golden/make_golden.py`` from the
    reference's own ``cpu_*`` functions on small seeded fixtures;
  * the reference tests' known answers (e.g. "WENO changes exactly 32 words"
    on the 8x8x8 advection fixture, the acoustic cadence 2/3/6);
  * closed-form invariants (constant fields, non-divergent flow, polynomial
    exactness, tridiagonal residuals against a dense binary64 solve).
Functions whose pinning is partial say so in their docstring
("parity unpinned" where nothing but the GPU comparison checks them).
"""
