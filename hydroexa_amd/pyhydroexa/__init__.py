"""PyHydroEXA-side differentiable interface: rollouts whose backward is the
library's hand-written adjoint, and roughness / bathymetry calibration.

The reference's PyHydroEXA package (PyHydroEXA/src/pyhydroexa/) is its
pre- and post-processing layer; this submodule adds what it does not have:
``Rollout`` / ``ShallowWaterRollout`` (a ``torch.autograd.Function``),
``RolloutBatch`` (many scenarios differentiated together) and the
calibration problem built on them.
"""
from .differentiable import Rollout, RolloutBatch, ShallowWaterRollout  # noqa: F401
from .calibration import CalibrationProblem, calibrate, synthetic_problem  # noqa: F401
