"""FP64 CPU restatement of HydroEXA's per-level advance (test infrastructure).

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this package, and only as the checker: the product path in
``hydroexa_amd`` never imports it.

Parity pinning status per function (the checks live in tests/test_oracle_*.py):

* ``roe_face``           pinned: consistency, upwind (supersonic) limits,
                          mirror/rotation symmetry, lake-at-rest balance,
                          friction closed form.
* ``hroe_face``          pinned: consistency, upwind limits, symmetry.
* ``vanleer_slope``      pinned: van Leer's harmonic-mean form, brute force.
* ``hydro_face``         pinned: flat-bed reduction to ``hroe_face``,
                          lake-at-rest balance, NODATA wall.
* ``roe_step``           pinned: mass conservation, y-invariance, lake at rest,
                          SWASHES dam-break closed form, transpose symmetry.
* ``hydro_step``         pinned: lake at rest with NODATA and dry islands,
                          mass conservation, SWASHES closed form, implicit
                          Manning equation.
* ``level_dt``           pinned: closed form per cell, brute force.
* ``tag``                pinned: linear fields (exact central differences),
                          brute-force numpy restatement, NODATA skip.
* ``masked_average_down`` pinned: closed forms (all valid, 50 % rule, NaN).
* ``prune_boxes``        pinned: brute-force numpy restatement.
* AMReX ghost-fill internals (``FilccCell``): restated, parity unpinned (no
  reference test or vector covers them).
"""
from .oracle import *  # noqa: F401,F403
