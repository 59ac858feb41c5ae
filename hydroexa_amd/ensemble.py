"""Batched ensemble of independent levels (flood scenarios) over the C-ABI.

One launch per step advances every scenario, each with its own CFL-limited
dt; a 6-step cycle can be captured in a HIP graph and replayed.  Each
scenario's result equals ``SolverVariant.run`` on that scenario alone.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np
import torch

from . import _lib as L
from .solver import Level, MODELS, _DT, context


class Ensemble:
    """Scenarios share a solver and a dtype; sizes may differ."""

    def __init__(self, levels: Sequence[Level], model: str):
        if model not in MODELS:
            raise ValueError(f"unknown model {model!r}")
        if not levels:
            raise ValueError("an ensemble needs at least one scenario")
        dtypes = {lv.dtype for lv in levels}
        if len(dtypes) != 1:
            raise ValueError("all scenarios must share one dtype")
        self.levels: List[Level] = list(levels)
        self.model = model
        self.dtype = dtypes.pop()
        n = len(self.levels)
        patches = (L.HxPatch * n)()
        alt = (C.c_void_p * (3 * n))()
        for s, lv in enumerate(self.levels):
            patches[s] = lv.patch(lv.U_new)
            plane = lv.nx * lv.ny * lv.U_old.element_size()
            b = lv.U_old.data_ptr()
            alt[3 * s], alt[3 * s + 1], alt[3 * s + 2] = b, b + plane, b + 2 * plane
        self._h = C.c_void_p()
        L.check(L.load().hx_ensemble_create(context(), MODELS[model], _DT[self.dtype], n, patches,
                                            alt, C.byref(self._h)), "hx_ensemble_create")

    @property
    def cells(self) -> int:
        return int(sum(lv.nx * lv.ny for lv in self.levels))

    def set_tuning(self, rows: int = 0, occupancy: int = 0, persistent: bool = False) -> None:
        L.check(L.load().hx_ensemble_set_tuning(self._h, int(rows), int(occupancy), int(persistent)),
                "hx_ensemble_set_tuning")

    def run(self, cfl: float, nsteps: int, graph: bool = True, stream: torch.cuda.Stream = None):
        """nsteps steps of every scenario; returns (dt history [nscen, nsteps],
        steps done per scenario).  A scenario whose dt collapses stops and the
        others go on; DtCollapse is raised after the run, with .done and
        .dt_hist (a stopped scenario's state is not defined).  Graph capture needs a non-default stream:
        one is made when none is given.  The final state is left in each
        level's U_new."""
        n = len(self.levels)
        dth = np.zeros((n, max(nsteps, 1)))
        done = np.zeros(n, dtype=np.int32)
        st = stream
        if graph and st is None:
            st = torch.cuda.Stream()
        cur = torch.cuda.current_stream()
        if st is not None:
            st.wait_stream(cur)
            handle = C.c_void_p(st.cuda_stream)
        else:
            handle = C.c_void_p(cur.cuda_stream)
        rc = L.load().hx_ensemble_run(self._h, float(cfl), int(nsteps), int(bool(graph)),
                                      dth.ctypes.data_as(C.POINTER(C.c_double)),
                                      done.ctypes.data_as(C.POINTER(C.c_int32)), handle)
        if st is not None:
            cur.wait_stream(st)
        if rc in (L.HX_OK, L.HX_E_DT_COLLAPSE) and nsteps % 2 == 1:
            for lv in self.levels:
                lv.swap()
        if rc == L.HX_E_DT_COLLAPSE:
            # the other scenarios ran every step: the exception carries the
            # steps done per scenario and the dt history
            e = L.DtCollapse(rc, "hx_ensemble_run (scenarios "
                             f"{[int(s) for s in np.nonzero(done < nsteps)[0]]} stopped)")
            e.done, e.dt_hist = done, dth[:, :nsteps]
            raise e
        L.check(rc, "hx_ensemble_run")
        return dth[:, :nsteps], done

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.load().hx_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
