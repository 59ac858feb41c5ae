"""Conservation diagnostics over the levels of a run.

Mirrors AmrMeshState::StoreInitialMassMomentum and ComputeMassDiagnostics
(src/state/AmrMeshState_Utils.cpp:213-460): per level one fused device
reduction (``hx_conservation``) over the valid cells whose roughness is not
NODATA and whose depth is positive, then totals over levels in level order
(sums) and the minimum depth, and the reference's log lines:

    [Initial] mass=...  px=...  py=...  KE=...
    [Conservation] t=...  iter=...  mass=...  px=...  py=...  KE=...  min_h=...
    [Conservation error] dMass=...%  dPx=...%  dPy=...%  dKE=...%

Numbers are printed as a default C++ ostream prints a double (6 significant
digits, %g style), which is what the reference's acceptance scripts grep
(tests/LakeAtRestHydrostaticRoe/run_t.sh:21-22, 54).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .solver import Level, context

DBL_MAX = 1.7976931348623157e308


def level_conservation(level: Level, U: Optional[torch.Tensor] = None) -> np.ndarray:
    """(mass, px, py, KE, min_h) of one level (its U_new unless U is given)."""
    out = np.zeros(5)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(L.load().hx_conservation(context(), C.byref(level.patch(U)),
                                     out.ctypes.data_as(C.POINTER(C.c_double)), st),
            "hx_conservation")
    return out


def _g(x: float) -> str:
    """A double as a default std::ostream prints it (precision 6, %g)."""
    if x == DBL_MAX:
        return "1.79769e+308"
    return format(x, ".6g")


class ConservationMonitor:
    """Baseline and per-step conservation report over a list of levels."""

    def __init__(self):
        self.initial: Optional[np.ndarray] = None

    @staticmethod
    def totals(levels: Sequence[Level]) -> np.ndarray:
        tot = np.zeros(4)
        mn = DBL_MAX
        for lv in levels:              # level order, as the IO rank sums
            r = level_conservation(lv)
            tot += r[:4]
            mn = min(mn, r[4])
        return np.concatenate([tot, [mn]])

    def store_initial(self, levels: Sequence[Level]) -> str:
        t = self.totals(levels)
        self.initial = t[:4].copy()
        return (f"[Initial] mass={_g(t[0])}  px={_g(t[1])}  py={_g(t[2])}  KE={_g(t[3])}")

    def report(self, levels: Sequence[Level], time: float, iteration: int) -> Dict:
        t = self.totals(levels)
        lines = [f"[Conservation] t={_g(time)}  iter={iteration}  mass={_g(t[0])}  px={_g(t[1])}"
                 f"  py={_g(t[2])}  KE={_g(t[3])}  min_h={_g(t[4])}"]
        out = {"mass": t[0], "px": t[1], "py": t[2], "KE": t[3], "min_h": t[4]}
        if self.initial is not None and t[0] > 0.0:
            with np.errstate(divide="ignore", invalid="ignore"):
                err = (t[:4] - self.initial) / self.initial * 100.0
            out["error_pct"] = err
            lines.append(f"[Conservation error] dMass={_g(err[0])}%  dPx={_g(err[1])}%"
                         f"  dPy={_g(err[2])}%  dKE={_g(err[3])}%")
        out["log"] = "\n".join(lines)
        return out
