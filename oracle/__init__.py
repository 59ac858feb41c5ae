"""FP64 CPU restatement of HydroEXA's per-level advance (test infrastructure).

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this package, and only as the checker: the product path in
``hydroexa_amd`` never imports it.

Parity pinning status per function (the checks live in tests/test_oracle_*.py):

* ``roe_face``           pinned: consistency, upwind (supersonic) limits,
                          mirror/rotation symmetry, lake-at-rest balance,
                          friction closed form; branch by branch: the
                          entropy fix on stationary expansion and
                          admissible shocks, the positivity guard's star
                          depths read back from the outputs (both branches
                          and the late-wave case), the emerged-step thrust
                          on either side.
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
* ``step_padded``        pinned: equals ``step`` when the ghosts hold the
                          physical halo.
* ``amr`` (two-level coupling, one fine box at ratio 2):
  - ``fill_fine``        pinned: children average to the parent, linear data
                          reproduced, flat at an extremum, time blends.
  - ``face_flux``        pinned: the side-selected fluxes reproduce the Roe
                          update; hydrostatic mass flux drives the depth update.
  - ``crse_init`` / ``fine_add`` / ``reflux`` / ``average_down``: pinned by
                          composite mass conservation with reflux (and its
                          failure without), uniform flow, lake at rest, child
                          means.  The exact time-blend weight and summation
                          order of AMReX's FillPatchTwoLevels / FluxRegister
                          are restated from its published algorithm: parity
                          unpinned at the last bit (AMReX is absent).
* ``run32`` (the restatement compiled in single precision, hx_oracle_f32.c):
                          pinned: within float rounding of ``run`` on smooth
                          flows, lake at rest kept to float round-off.
* ``amr.AmrOracle`` regrid replay: pinned by the hierarchy tests it checks
  (recorded events must fall where the regrid cadence puts them).
* AMReX ghost-fill internals (``FilccCell``): restated, parity unpinned (no
  reference test or vector covers them).
"""
from .oracle import *  # noqa: F401,F403
