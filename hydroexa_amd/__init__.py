"""hydroexa_amd: MI355X-native per-level shallow-water advance (HydroEXA's hot path).

The arithmetic lives in hand-written HIP kernels for gfx950
(``csrc/``, built into ``lib/libhydroexa.so``) behind the C-ABI of
``include/hydroexa.h``; this package is the host-side mirror of the
reference's solver interface.
"""
from .solver import (  # noqa: F401
    Level,
    SolverVariant,
    context,
    masked_average_down,
    post_process_base_grids,
    set_arith,
    set_tuning,
)
from .ensemble import Ensemble  # noqa: F401,E402
from .amr import TwoLevel  # noqa: F401,E402
from . import _lib  # noqa: F401
from . import workloads  # noqa: F401,E402

__all__ = ["Ensemble", "Level", "TwoLevel", "SolverVariant", "masked_average_down", "post_process_base_grids",
           "set_tuning", "set_arith", "workloads"]
