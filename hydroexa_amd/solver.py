"""Host-side mirror of HydroEXA's per-level solver interface over the C-ABI.

``SolverVariant`` keeps the reference's entry points and their meaning
(include/solvers/SolverVariant.H:45-191, include/solvers/BaseSolver.H:41-193):
``compute_fluxes`` advances one level by dt, swapping U_new/U_old first;
``compute_dt`` runs the multi-level dt protocol; ``tag_cells`` marks cells for
refinement.  ``masked_average_down`` and ``post_process_base_grids`` mirror
include/utils/MaskedAverageDown.H:124-127 and
src/state/AmrMeshState_AMR.cpp:318-405.  All arithmetic runs in the HIP
library; torch only owns device memory and the stream.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L

MODELS = {"Roe": L.HX_SOLVER_ROE, "HydrostaticRoe": L.HX_SOLVER_HYDROSTATIC_ROE}
_DT = {torch.float64: L.HX_F64, torch.float32: L.HX_F32}

_ctx = {}


def context(device: int = None) -> C.c_void_p:
    """Per-device library context (created on first use, kept for the process)."""
    if device is None:
        device = torch.cuda.current_device()
    if device not in _ctx:
        h = C.c_void_p()
        L.check(L.load().hx_ctx_create(C.byref(h), device), "hx_ctx_create")
        _ctx[device] = h
    return _ctx[device]


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_cuda(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor (the solver has no CPU path)")


class Level:
    """One box covering a level: U_new/U_old as (3, ny, nx) planes h, hu, hv
    and terrain as (2, ny, nx) planes z, n (the reference's U and
    DynamicTerrain MultiFabs for a single FAB, component-major, i fastest).

    bc_lo/bc_hi hold the input-file codes per direction (0 reflective,
    1 transmissive; src/state/AmrMeshState_Init.cpp:42-70)."""

    def __init__(self, h, hu, hv, z, n, dx: float, dy: float,
                 bc_lo: Sequence[int] = (0, 0), bc_hi: Sequence[int] = (0, 0),
                 dtype: torch.dtype = torch.float64, device: str = "cuda"):
        def dev(a):
            return torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a,
                                   dtype=dtype).to(device)
        h, hu, hv, z, n = (dev(a) for a in (h, hu, hv, z, n))
        if h.dim() != 2 or any(a.shape != h.shape for a in (hu, hv, z, n)):
            raise ValueError("h, hu, hv, z, n must share one (ny, nx) shape")
        self.ny, self.nx = h.shape
        self.dtype = dtype
        self.U_new = torch.stack([h, hu, hv]).contiguous()
        self.U_old = torch.empty_like(self.U_new)
        self.terrain = torch.stack([z, n]).contiguous()
        self.dx, self.dy = float(dx), float(dy)
        self.bc_lo, self.bc_hi = tuple(bc_lo), tuple(bc_hi)
        _require_cuda(self.U_new, "state")

    def patch(self, U: torch.Tensor = None) -> L.HxPatch:
        U = self.U_new if U is None else U
        es = U.element_size()
        plane = self.nx * self.ny * es
        p = L.HxPatch()
        p.nx, p.ny, p.ng, p.dtype = self.nx, self.ny, 0, _DT[self.dtype]
        p.pitch = self.nx
        base = U.data_ptr()
        p.h, p.hu, p.hv = base, base + plane, base + 2 * plane
        p.eta = None
        tb = self.terrain.data_ptr()
        p.z, p.n = tb, tb + plane
        p.dx, p.dy = self.dx, self.dy
        p.bc_lo[0], p.bc_lo[1] = self.bc_lo
        p.bc_hi[0], p.bc_hi[1] = self.bc_hi
        return p

    def swap(self) -> None:
        self.U_new, self.U_old = self.U_old, self.U_new

    @property
    def h(self):
        return self.U_new[0]

    @property
    def hu(self):
        return self.U_new[1]

    @property
    def hv(self):
        return self.U_new[2]


class SolverVariant:
    """std::variant<monostate, Roe, HydrostaticRoe> dispatch, by model name."""

    def __init__(self, model: str = None):
        if model is not None and model not in MODELS:
            raise ValueError(f"unknown model {model!r}; expected one of {sorted(MODELS)}")
        self.model = model

    def _kind(self) -> int:
        if self.model is None:
            raise RuntimeError("[SolverVariant Error] Solver not initialized.")
        return MODELS[self.model]

    def has_solver(self) -> bool:
        return self.model is not None

    # -- SolverVariant::compute_fluxes ------------------------------------
    def compute_fluxes(self, level: Level, dt: float, time: float = 0.0,
                       edge_fluxes: torch.Tensor = None) -> None:
        """Advance ``level`` by dt.  Swaps U_new/U_old, then writes U_new from
        U_old (src/solvers/Roe.cpp:255, HydrostaticRoe.cpp:405).

        edge_fluxes: optional (2 ny + 2 nx, 3) device tensor of the level's
        dtype that receives, from the same kernel, the effective fluxes of the
        box's boundary faces as its own cells took them -- what the reference
        hands to the fine side of a flux register (Roe.cpp:544-566,
        HydrostaticRoe.cpp:612-627): x-low faces by row, x-high, y-low by
        column, y-high (hx_advance_level_fluxes)."""
        kind = self._kind()
        if edge_fluxes is not None:
            _require_cuda(edge_fluxes, "edge_fluxes")
            if (tuple(edge_fluxes.shape) != (2 * level.ny + 2 * level.nx, 3)
                    or edge_fluxes.dtype != level.dtype or not edge_fluxes.is_contiguous()):
                raise ValueError("edge_fluxes must be a contiguous (2 ny + 2 nx, 3) tensor of the "
                                 "level's dtype")
        level.swap()
        p = level.patch(level.U_old)
        es = level.U_new.element_size()
        plane = level.nx * level.ny * es
        base = level.U_new.data_ptr()
        if edge_fluxes is None:
            rc = L.load().hx_advance_level(context(), kind, C.byref(p), float(dt), base,
                                           base + plane, base + 2 * plane, level.nx, _stream())
            L.check(rc, "hx_advance_level")
        else:
            rc = L.load().hx_advance_level_fluxes(context(), kind, C.byref(p), float(dt), base,
                                                  base + plane, base + 2 * plane, level.nx,
                                                  edge_fluxes.data_ptr(), _stream())
            L.check(rc, "hx_advance_level_fluxes")

    # -- compute_dt_level / BaseSolver::compute_dt -------------------------
    def compute_dt_level(self, level: Level, cfl: float) -> Tuple[float, float]:
        """(cfl * dt estimate, max |vel|) of one level; raises DtCollapse on
        an estimate below 1e-12 (Roe.cpp:88-92, HydrostaticRoe.cpp:77-80)."""
        kind = self._kind()
        p = level.patch()
        dt, vel = C.c_double(), C.c_double()
        rc = L.load().hx_level_dt(context(), kind, C.byref(p), float(cfl), C.byref(dt),
                                  C.byref(vel), _stream())
        L.check(rc, "hx_level_dt")
        return dt.value, vel.value

    def compute_dt(self, levels: List[Level], ref_ratio: Sequence[int], cfl: float) -> List[float]:
        """Multi-level protocol of include/solvers/BaseSolver.H:120-165: every
        level's constraint is projected to level 0, then divided back down."""
        dt_tmp = [self.compute_dt_level(lv, cfl)[0] for lv in levels]
        dt0, factor = dt_tmp[0], 1.0
        for lev in range(1, len(levels)):
            factor *= float(ref_ratio[lev - 1])
            dt0 = min(dt0, dt_tmp[lev] * factor)
        dt = [dt0]
        for lev in range(1, len(levels)):
            dt.append(dt[lev - 1] / float(ref_ratio[lev - 1]))
        return dt

    # -- SolverVariant::tag_cells -----------------------------------------
    def tag_cells(self, level: Level, time: float, h_grad_thresh: float,
                  umag_grad_thresh: float, z_grad_thresh: float) -> torch.Tensor:
        kind = self._kind()
        p = level.patch()
        tags = torch.empty((level.ny, level.nx), dtype=torch.uint8, device=level.U_new.device)
        th = (C.c_double * 3)(h_grad_thresh, umag_grad_thresh, z_grad_thresh)
        rc = L.load().hx_tag_level(context(), kind, C.byref(p), float(time), th, tags.data_ptr(),
                                   _stream())
        L.check(rc, "hx_tag_level")
        return tags

    # -- step-driven loop (HydroEXA::Compute for one level) ----------------
    def run(self, level: Level, cfl: float, nsteps: int, resume: bool = False):
        """nsteps x (ComputeDt; AdvanceLevel) with the CFL reduction fused into
        the advance.  Leaves the final state in U_new; returns numpy arrays of
        per-step dt and max|vel|.  resume: continue the previous run of this
        level (U_new untouched since), whose last step already reduced the
        first dt (hx_run_level_resume; raises if that run is not the
        context's last)."""
        kind = self._kind()
        p = level.patch(level.U_new)
        es = level.U_new.element_size()
        plane = level.nx * level.ny * es
        alt = level.U_old.data_ptr()
        dth = np.zeros(max(nsteps, 1))
        vh = np.zeros(max(nsteps, 1))
        done = C.c_int32()
        fn = L.load().hx_run_level_resume if resume else L.load().hx_run_level
        rc = fn(context(), kind, C.byref(p), alt, alt + plane, alt + 2 * plane, float(cfl),
                int(nsteps), dth.ctypes.data_as(C.POINTER(C.c_double)),
                vh.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done), _stream())
        if done.value % 2 == 1:
            level.swap()
        L.check(rc, f"{fn.__name__} (completed {done.value} of {nsteps} steps)")
        return dth[:nsteps], vh[:nsteps]


def set_tuning(model: str, dtype: torch.dtype, rows: int = 0, occupancy: int = 0) -> None:
    """Launch shape of the fused advance: rows marched per wavefront and the
    minimum waves per SIMD (0 restores the defaults)."""
    L.check(L.load().hx_set_tuning(context(), MODELS[model], _DT[dtype], int(rows), int(occupancy)),
            "hx_set_tuning")


def set_arith(mode: str) -> None:
    """'exact' (default; bit-identical to the reference's operation order in
    FP64) or 'fast' (shared reciprocals and FMA contraction)."""
    code = {"exact": L.HX_ARITH_EXACT, "fast": L.HX_ARITH_FAST}[mode]
    L.check(L.load().hx_set_arith(context(), code), "hx_set_arith")


def masked_average_down(fine: torch.Tensor, ratio: Sequence[int], dx_fine: float,
                        dy_fine: float, nodata: float = -9999.0):
    """NODATA-aware restriction of (ncomp, nyf, nxf) -> (ncomp, nyf/ry, nxf/rx);
    returns (crse, nvalid)."""
    _require_cuda(fine, "fine")
    f = fine.contiguous()
    if f.dim() == 2:
        f = f.unsqueeze(0)
    nc, nyf, nxf = f.shape
    rx, ry = int(ratio[0]), int(ratio[1])
    if nxf % rx or nyf % ry:
        raise ValueError("fine extent must be divisible by the ratio")
    crse = torch.empty((nc, nyf // ry, nxf // rx), dtype=f.dtype, device=f.device)
    nval = torch.empty((nc, nyf // ry, nxf // rx), dtype=torch.int32, device=f.device)
    rc = L.load().hx_masked_average_down(context(), _DT[f.dtype], nxf, nyf, nc, f.data_ptr(), rx,
                                         ry, float(dx_fine), float(dy_fine), float(nodata),
                                         crse.data_ptr(), nval.data_ptr(), _stream())
    L.check(rc, "hx_masked_average_down")
    return crse, nval


def post_process_base_grids(boxes, static_z: torch.Tensor, terrain_ref_lev: int):
    """Prune level-0 boxes that hold no valid static terrain.  Returns
    (kept_boxes, pure_fluid_indices, boundary_indices) like
    AmrMeshState::PostProcessBaseGrids (indices into the kept list)."""
    _require_cuda(static_z, "static_z")
    b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 4))
    z = static_z.contiguous()
    nys, nxs = z.shape
    keep = np.zeros(len(b), dtype=np.uint8)
    nd = np.zeros(len(b), dtype=np.uint8)
    nk = C.c_int32()
    rc = L.load().hx_prune_boxes(context(), _DT[z.dtype], len(b),
                                 b.ctypes.data_as(C.POINTER(C.c_int32)), nxs, nys, z.data_ptr(),
                                 int(terrain_ref_lev), keep.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 nd.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(nk), _stream())
    L.check(rc, "hx_prune_boxes")
    kept, pure, boundary = [], [], []
    for i in range(len(b)):
        if keep[i]:
            kept.append(tuple(int(v) for v in b[i]))
            (boundary if nd[i] else pure).append(len(kept) - 1)
    return kept, pure, boundary, keep.astype(bool), nd.astype(bool)
