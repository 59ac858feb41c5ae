"""Two-level coupling for one fixed fine box at refinement ratio 2.

The reference couples AMR levels with AMReX (src/state/AmrMeshState_Physics.cpp:13-147,
src/state/AmrMeshState_AMR.cpp:667-844); this module drives the same
sequence on the library's kernels for one coarse level and one fine box:

* dt: per-level CFL estimates, dt0 = min(dt_cfl[0], 2 dt_cfl[1]), dt1 = dt0/2
  (include/solvers/BaseSolver.H:120-165); lockstep takes min(dt0, dt1)
  for both levels (AmrMeshState_Physics.cpp:13-45);
* coarse step: the register of each coarse face around the box starts as
  -dt dA times the coarse cell's effective flux (FluxRegister::CrseInit,
  COPY), then the coarse level advances;
* fine substeps at coarse-time blends 0 and 1/2 (one substep at 0 in
  lockstep): the ghost ring is filled from the coarse level (eta and
  momenta by conservative linear interpolation, h = max(0, eta - z) on the
  fine bed), the register gains dt_f dA_f times the fine effective fluxes
  (FineAdd), and the fine level advances with its stored ghosts;
* reflux, then average_down of the state into the covered coarse cells
  (AmrMeshState_Physics.cpp:133-141).

Restrictions of this fixed-box coupling: one fine box, ratio 2, the box at
least one coarse cell away from the domain edge, no regridding.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from .solver import MODELS, _DT, Level, context

NG = 2
Box = Tuple[int, int, int, int]


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class FineBox:
    """A fine patch with 2 stored ghost layers: U (3, ny+4, nx+4) double
    buffered, terrain (2, ny+4, nx+4) including its ghosts."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, terrain: torch.Tensor,
                 dtype: torch.dtype):
        self.nx, self.ny, self.dx, self.dy, self.dtype = nx, ny, dx, dy, dtype
        shape = (3, ny + 2 * NG, nx + 2 * NG)
        self.U_new = torch.zeros(shape, dtype=dtype, device="cuda")
        self.U_old = torch.zeros(shape, dtype=dtype, device="cuda")
        self.terrain = terrain.to(device="cuda", dtype=dtype).contiguous()
        if tuple(self.terrain.shape) != (2, ny + 2 * NG, nx + 2 * NG):
            raise ValueError("fine terrain must include the 2 ghost layers")

    def _ptr(self, t: torch.Tensor, comp: int) -> int:
        W = self.nx + 2 * NG
        plane = (self.ny + 2 * NG) * W * t.element_size()
        return t.data_ptr() + comp * plane + (NG * W + NG) * t.element_size()

    def patch(self, U: Optional[torch.Tensor] = None) -> L.HxPatch:
        U = self.U_new if U is None else U
        p = L.HxPatch()
        p.nx, p.ny, p.ng, p.dtype = self.nx, self.ny, NG, _DT[self.dtype]
        p.pitch = self.nx + 2 * NG
        p.h, p.hu, p.hv = self._ptr(U, 0), self._ptr(U, 1), self._ptr(U, 2)
        p.eta = None
        p.z, p.n = self._ptr(self.terrain, 0), self._ptr(self.terrain, 1)
        p.dx, p.dy = self.dx, self.dy
        p.bc_lo[0] = p.bc_lo[1] = p.bc_hi[0] = p.bc_hi[1] = L.HX_BC_GHOST
        return p

    def valid(self, U: Optional[torch.Tensor] = None) -> torch.Tensor:
        U = self.U_new if U is None else U
        return U[:, NG:NG + self.ny, NG:NG + self.nx]

    def swap(self):
        self.U_new, self.U_old = self.U_old, self.U_new


def fine_terrain_from_coarse(crse: Level, box: Tuple[int, int, int, int]) -> torch.Tensor:
    """Piecewise-constant fine terrain with 2 ghost layers (amrex pc_interp;
    the reference fills terrain ghosts with it, AmrMeshState_AMR.cpp:702-733)."""
    ci0, cj0, ncx, ncy = box
    nxf, nyf = 2 * ncx, 2 * ncy
    fi = np.arange(-NG, nxf + NG)
    fj = np.arange(-NG, nyf + NG)
    pi = np.clip(ci0 + np.floor_divide(fi, 2), 0, crse.nx - 1)     # terrain is foextrap
    pj = np.clip(cj0 + np.floor_divide(fj, 2), 0, crse.ny - 1)
    t = crse.terrain[:, torch.as_tensor(pj, device="cuda")][:, :, torch.as_tensor(pi, device="cuda")]
    return t.contiguous()


class TwoLevel:
    """One coarse level and one fine box (ratio 2) advanced together."""

    def __init__(self, model: str, crse: Level, box: Tuple[int, int, int, int],
                 fine_terrain: Optional[torch.Tensor] = None, subcycle: bool = True):
        if model not in MODELS:
            raise ValueError(f"unknown model {model!r}")
        ci0, cj0, ncx, ncy = box
        if ci0 < 1 or cj0 < 1 or ci0 + ncx > crse.nx - 1 or cj0 + ncy > crse.ny - 1:
            raise ValueError("the fine box must stay one coarse cell away from the domain edge")
        self.model, self.solver = model, MODELS[model]
        self.crse, self.box, self.subcycle = crse, box, subcycle
        ter = fine_terrain if fine_terrain is not None else fine_terrain_from_coarse(crse, box)
        self.fine = FineBox(2 * ncx, 2 * ncy, crse.dx / 2, crse.dy / 2, ter, crse.dtype)
        self.reg = torch.zeros((2 * (ncx + ncy), 3), dtype=crse.dtype, device="cuda")
        self._faces()
        self.time = 0.0
        self.steps = 0
        # initial fine state from the coarse level (FillCoarsePatch)
        cp = crse.patch(crse.U_new)
        f = self.fine
        L.check(L.load().hx_fill_fine_ghosts(context(), C.byref(cp), C.byref(cp), 0.0, 2, ci0, cj0,
                                             C.byref(f.patch()), 1, f._ptr(f.U_new, 0),
                                             f._ptr(f.U_new, 1), f._ptr(f.U_new, 2), _stream()),
                "hx_fill_fine_ghosts")

    def _faces(self):
        ci0, cj0, ncx, ncy = self.box
        nxf, nyf = 2 * ncx, 2 * ncy
        cx = [(0, ci0, cj0 + m, 0) for m in range(ncy)] + \
             [(0, ci0 + ncx, cj0 + m, 1) for m in range(ncy)]
        cy = [(1, ci0 + m, cj0, 0) for m in range(ncx)] + \
             [(1, ci0 + m, cj0 + ncy, 1) for m in range(ncx)]
        fx = [(0, 0, 2 * m + s, 1) for m in range(ncy) for s in (0, 1)] + \
             [(0, nxf, 2 * m + s, 0) for m in range(ncy) for s in (0, 1)]
        fy = [(1, 2 * m + s, 0, 1) for m in range(ncx) for s in (0, 1)] + \
             [(1, 2 * m + s, nyf, 0) for m in range(ncx) for s in (0, 1)]
        T = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int32).reshape(-1, 4), device="cuda")
        self.cfaces = T(cx + cy)
        self.ffaces = T(fx + fy)

    def _flux(self, patch, faces, pair, scale_x, scale_y, accumulate):
        n = faces.shape[0]
        L.check(L.load().hx_face_fluxes(context(), self.solver, C.byref(patch), n,
                                        C.c_void_p(faces.data_ptr()), pair, float(scale_x),
                                        float(scale_y), int(accumulate),
                                        C.c_void_p(self.reg.data_ptr()), _stream()),
                "hx_face_fluxes")

    def level_dts(self, cfl: float) -> Tuple[float, float]:
        lib = L.load()
        out = []
        for p in (self.crse.patch(), self.fine.patch()):
            dt, vmax = C.c_double(), C.c_double()
            L.check(lib.hx_level_dt(context(), self.solver, C.byref(p), float(cfl), C.byref(dt),
                                    C.byref(vmax), _stream()), "hx_level_dt")
            out.append(dt.value)
        return out[0], out[1]

    def dt(self, cfl: float) -> Tuple[float, float]:
        d0, d1 = self.level_dts(cfl)
        dt0 = min(d0, d1 * 2.0)
        return (dt0, dt0 / 2.0) if self.subcycle else (dt0 / 2.0, dt0 / 2.0)

    def step(self, cfl: Optional[float] = None,
             dts: Optional[Tuple[float, float]] = None) -> Tuple[float, float]:
        """One coarse step with its fine substeps; dt from the CFL rule at
        ``cfl``, or the given (dt_coarse, dt_fine)."""
        if (cfl is None) == (dts is None):
            raise ValueError("give exactly one of cfl and dts")
        lib = L.load()
        ci0, cj0, ncx, ncy = self.box
        c, f = self.crse, self.fine
        dt0, dt1 = self.dt(cfl) if dts is None else (float(dts[0]), float(dts[1]))
        # coarse: register from the step's input, then the advance
        cp = c.patch(c.U_new)
        self._flux(cp, self.cfaces, 1, -dt0 * c.dy, -dt0 * c.dx, 0)
        c.swap()
        L.check(lib.hx_advance_level(context(), self.solver, C.byref(c.patch(c.U_old)), dt0,
                                     c.U_new[0].data_ptr(), c.U_new[1].data_ptr(),
                                     c.U_new[2].data_ptr(), c.nx, _stream()), "hx_advance_level")
        old_p, new_p = c.patch(c.U_old), c.patch(c.U_new)
        nsub = 2 if self.subcycle else 1
        for k in range(nsub):
            alpha = k / 2.0
            L.check(lib.hx_fill_fine_ghosts(context(), C.byref(old_p), C.byref(new_p), alpha, 2,
                                            ci0, cj0, C.byref(f.patch()), 0, f._ptr(f.U_new, 0),
                                            f._ptr(f.U_new, 1), f._ptr(f.U_new, 2), _stream()),
                    "hx_fill_fine_ghosts")
            fp = f.patch(f.U_new)
            self._flux(fp, self.ffaces, 2, dt1 * f.dy, dt1 * f.dx, 1)
            f.swap()
            out = f.patch(f.U_new)
            L.check(lib.hx_advance_level(context(), self.solver, C.byref(f.patch(f.U_old)), dt1,
                                         C.c_void_p(out.h), C.c_void_p(out.hu), C.c_void_p(out.hv),
                                         f.nx + 2 * NG, _stream()), "hx_advance_level")
        np_ = c.patch(c.U_new)
        L.check(lib.hx_reflux(context(), C.byref(np_), c.U_new[0].data_ptr(), c.U_new[1].data_ptr(),
                              c.U_new[2].data_ptr(), ci0, cj0, ncx, ncy,
                              C.c_void_p(self.reg.data_ptr()), _stream()), "hx_reflux")
        L.check(lib.hx_average_down_state(context(), C.byref(f.patch()), 2, ci0, cj0,
                                          C.byref(np_), c.U_new[0].data_ptr(),
                                          c.U_new[1].data_ptr(), c.U_new[2].data_ptr(), _stream()),
                "hx_average_down_state")
        self.time += dt0
        self.steps += 1
        return dt0, dt1

    def run(self, nsteps: int, cfl: float) -> np.ndarray:
        """nsteps coupled steps with dt from the CFL rule, on the device without
        host synchronisation between steps (hx_two_level_run).  Returns the
        (nsteps, 2) coarse and fine dt history; raises DtCollapse when an
        estimate falls below 1e-12, leaving the state of the last good step."""
        c, f = self.crse, self.fine
        hist = np.zeros((max(nsteps, 1), 2))
        done = C.c_int32()
        ca, fa = c.U_old, f.U_old
        plane = c.nx * c.ny * ca.element_size()
        rc = L.load().hx_two_level_run(
            context(), self.solver, C.byref(c.patch(c.U_new)), ca.data_ptr(),
            ca.data_ptr() + plane, ca.data_ptr() + 2 * plane, C.byref(f.patch(f.U_new)),
            f._ptr(fa, 0), f._ptr(fa, 1), f._ptr(fa, 2), self.box[0], self.box[1],
            int(self.subcycle), float(cfl), int(nsteps), C.c_void_p(self.reg.data_ptr()),
            hist.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done), _stream())
        k = done.value
        if k % 2 == 1:
            c.swap()
        if (k * (2 if self.subcycle else 1)) % 2 == 1:
            f.swap()
        self.time += float(hist[:k, 0].sum())
        self.steps += k
        if rc == L.HX_E_DT_COLLAPSE:
            raise L.DtCollapse(rc, f"hx_two_level_run (completed {k} of {nsteps} steps)")
        L.check(rc, "hx_two_level_run")
        return hist[:nsteps]

    def enqueue(self, nsteps: int, cfl: float, dt_hist: torch.Tensor,
                status: torch.Tensor) -> None:
        """nsteps coupled steps as stream work only (hx_two_level_enqueue):
        no host copy or synchronisation, so the call may be captured in a
        HIP graph (torch.cuda.graph) and replayed.  nsteps must be even, so
        that every replay starts from and ends in the same buffers.
        dt_hist: float64 device tensor of >= 2 nsteps (coarse dt, fine dt);
        status: int32 device tensor, set to 0 or to k + 1 when step k's
        estimate collapsed.  time and steps are left to the caller, who
        reads them from dt_hist."""
        if nsteps < 2 or nsteps % 2:
            raise ValueError("enqueue: nsteps must be even and positive")
        if (dt_hist.dtype != torch.float64 or not dt_hist.is_cuda or
                dt_hist.numel() < 2 * nsteps or not dt_hist.is_contiguous()):
            raise ValueError("enqueue: dt_hist must be a contiguous float64 device tensor "
                             "of at least 2 * nsteps values")
        if status.dtype != torch.int32 or not status.is_cuda or status.numel() < 1:
            raise ValueError("enqueue: status must be an int32 device tensor")
        c, f = self.crse, self.fine
        ca, fa = c.U_old, f.U_old
        plane = c.nx * c.ny * ca.element_size()
        L.check(L.load().hx_two_level_enqueue(
            context(), self.solver, C.byref(c.patch(c.U_new)), ca.data_ptr(),
            ca.data_ptr() + plane, ca.data_ptr() + 2 * plane, C.byref(f.patch(f.U_new)),
            f._ptr(fa, 0), f._ptr(fa, 1), f._ptr(fa, 2), self.box[0], self.box[1],
            int(self.subcycle), float(cfl), int(nsteps), C.c_void_p(self.reg.data_ptr()),
            C.c_void_p(dt_hist.data_ptr()), C.c_void_p(status.data_ptr()), _stream()),
            "hx_two_level_enqueue")

    def composite_mass(self) -> float:
        """Sum of h dA over the coarse level (covered cells hold the average
        of their fine children after each step)."""
        return float(self.crse.U_new[0].double().sum().item()) * self.crse.dx * self.crse.dy


# ===========================================================================
# Multi-level, multi-box hierarchy
# ===========================================================================
class DenseLevel:
    """One level held as a full-domain array with 2 cells of padding on each
    side: U_new / U_old (3, ny + 4, nx + 4), terrain (2, ny + 4, nx + 4), and
    a byte mask of the cells the level's boxes own.  Boxes read their ghost
    rings straight from the array: another box's valid cells where the level
    has them, values interpolated from the coarser level elsewhere, physical
    boundary rules outside the domain."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, bc_lo, bc_hi, dtype):
        self.nx, self.ny, self.dx, self.dy, self.dtype = nx, ny, dx, dy, dtype
        self.bc_lo, self.bc_hi = tuple(bc_lo), tuple(bc_hi)
        shape = (ny + 2 * NG, nx + 2 * NG)
        self.U_new = torch.zeros((3,) + shape, dtype=dtype, device="cuda")
        self.U_old = torch.zeros((3,) + shape, dtype=dtype, device="cuda")
        self.terrain = torch.zeros((2,) + shape, dtype=dtype, device="cuda")
        self.mask = torch.zeros(shape, dtype=torch.uint8, device="cuda")
        self.boxes: List[Box] = []
        self.patches: List[Box] = []
        self.t_new = 0.0
        self.t_old = -1.0e200
        self._faces: Optional[List[Tuple[torch.Tensor, torch.Tensor]]] = []

    @property
    def pitch(self) -> int:
        return self.nx + 2 * NG

    def addr(self, t: torch.Tensor, comp: int, i: int, j: int) -> int:
        """Address of element (comp, j, i), i and j in level index space."""
        plane = (self.ny + 2 * NG) * self.pitch
        return t.data_ptr() + (comp * plane + (j + NG) * self.pitch + (i + NG)) * t.element_size()

    def mask_addr(self, m: torch.Tensor, i: int, j: int) -> int:
        return m.data_ptr() + (j + NG) * self.pitch + (i + NG)

    def patch(self, U: torch.Tensor, box: Optional[Box] = None) -> L.HxPatch:
        """The whole level (physical sides) or one box: sides on the domain
        boundary keep the physical rule, the others read stored ghosts."""
        i0, j0, w, h = box if box is not None else (0, 0, self.nx, self.ny)
        p = L.HxPatch()
        p.nx, p.ny, p.ng, p.dtype = w, h, NG, _DT[self.dtype]
        p.pitch = self.pitch
        p.h, p.hu, p.hv = (self.addr(U, k, i0, j0) for k in range(3))
        p.eta = None
        p.z, p.n = self.addr(self.terrain, 0, i0, j0), self.addr(self.terrain, 1, i0, j0)
        p.dx, p.dy = self.dx, self.dy
        p.bc_lo[0] = self.bc_lo[0] if i0 == 0 else L.HX_BC_GHOST
        p.bc_lo[1] = self.bc_lo[1] if j0 == 0 else L.HX_BC_GHOST
        p.bc_hi[0] = self.bc_hi[0] if i0 + w == self.nx else L.HX_BC_GHOST
        p.bc_hi[1] = self.bc_hi[1] if j0 + h == self.ny else L.HX_BC_GHOST
        return p

    def interior(self, t: torch.Tensor) -> torch.Tensor:
        return t[..., NG:NG + self.ny, NG:NG + self.nx]

    def set_boxes(self, boxes: List[Box]) -> None:
        """The level's grids (bookkeeping, as AMReX's BoxArray) and the patches
        it executes on: maximal rectangles of their union, aligned to 2.
        Splitting a region into boxes changes no result (every box reads its
        neighbours' cells as ghosts and their shared faces carry the same
        fluxes), so fewer, larger patches give the same answer with fewer
        launches."""
        from .regrid import _rects
        new = [tuple(map(int, b)) for b in boxes]
        if new == self.boxes:
            return False
        self.boxes = new
        # the union painted on the host at the boxes' common alignment g (a
        # small array): expanded by g on the device into the mask in one copy,
        # and split into the patches
        g = max(int(np.gcd.reduce(np.asarray(self.boxes, np.int64).ravel())), 1) if new else 1
        nyb, nxb = -(-self.ny // g), -(-self.nx // g)
        mb = np.zeros((nyb, nxb), np.uint8)
        for (i0, j0, w, h) in self.boxes:
            mb[j0 // g:(j0 + h) // g, i0 // g:(i0 + w) // g] = 1
        t = torch.from_numpy(mb).to(self.mask.device)
        full = t[:, None, :, None].expand(nyb, g, nxb, g).reshape(nyb * g, nxb * g)
        self.mask.zero_()
        self.mask[NG:NG + self.ny, NG:NG + self.nx].copy_(full[:self.ny, :self.nx])
        if len(self.boxes) == 1:
            self.patches = list(self.boxes)
        else:
            self.patches = [(g * a, g * b, g * w, g * h) for (a, b, w, h) in _rects(mb.astype(bool))]
        self._faces = None            # built on first use (the host-driven step)
        nreg = sum(6 * (w // 2 + h // 2) for (_, _, w, h) in self.patches)
        self.regbuf = torch.zeros(max(nreg, 1), dtype=self.dtype, device="cuda")
        self.patch_arr = np.ascontiguousarray(np.asarray(self.patches, np.int32).reshape(-1, 4))
        return True

    @property
    def faces(self) -> List[Tuple[torch.Tensor, torch.Tensor]]:
        """Per patch, the device face lists of its flux registers: the coarse
        faces around it (coarse indices, coarse side's view) and its own fine
        faces (its indices, fine side's view).  hx_amr_run generates the same
        faces on the device and does not need them."""
        if self._faces is not None:
            return self._faces
        self._faces = []
        for (i0, j0, w, h) in self.patches:
            ci0, cj0, ncx, ncy = i0 // 2, j0 // 2, w // 2, h // 2
            cx = [(0, ci0, cj0 + m, 0) for m in range(ncy)] + \
                 [(0, ci0 + ncx, cj0 + m, 1) for m in range(ncy)]
            cy = [(1, ci0 + m, cj0, 0) for m in range(ncx)] + \
                 [(1, ci0 + m, cj0 + ncy, 1) for m in range(ncx)]
            fx = [(0, 0, 2 * m + s, 1) for m in range(ncy) for s in (0, 1)] + \
                 [(0, w, 2 * m + s, 0) for m in range(ncy) for s in (0, 1)]
            fy = [(1, 2 * m + s, 0, 1) for m in range(ncx) for s in (0, 1)] + \
                 [(1, 2 * m + s, h, 0) for m in range(ncx) for s in (0, 1)]
            T = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int32).reshape(-1, 4), device="cuda")
            self._faces.append((T(cx + cy), T(fx + fy)))
        return self._faces

    def swap(self) -> None:
        self.U_new, self.U_old = self.U_old, self.U_new


def _pad_edges(a: np.ndarray) -> np.ndarray:
    """(ny, nx) -> (ny + 4, nx + 4), edge values replicated (foextrap)."""
    return np.pad(a, NG, mode="edge")


def _pc_refine(a: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(a, 2, axis=0), 2, axis=1)


def _almost_equal(x: float, y: float) -> bool:
    return abs(x - y) <= 1e-12 * max(abs(x), abs(y), 1e-300)


class AmrHierarchy:
    """Levels 0..max_level at refinement ratio 2, each a list of boxes,
    advanced by the reference's recursive step (src/state/AmrMeshState_Physics.cpp:50-147):
    dt from every level's CFL (include/solvers/BaseSolver.H:120-165; lockstep
    shares the finest dt), regrid every regrid_int steps of a level from its
    gradient tags (hydroexa_amd.regrid), ghost rings from the coarser level
    at the substep's time, flux registers, reflux in AMReX's side order, and
    average_down.

    static(level) may return the level's initial (h, hu, hv, z, n) on the
    whole domain at that level's resolution (the reference's static input
    layer); otherwise new levels take their terrain piecewise-constant and
    their state by interpolation from the coarser level.  boxes maps a level
    to a fixed box list (the level is then not regridded)."""

    def __init__(self, model: str, h, hu, hv, z, n, dx: float, dy: float,
                 bc_lo=(0, 0), bc_hi=(0, 0), max_level: int = 1, regrid_int: int = 0,
                 blocking_factor: int = 8, max_grid_size: int = 128, n_error_buf: int = 1,
                 grid_eff: float = 0.7, nest: int = 2,
                 tag_thresh: Sequence[Tuple[float, float, float]] = ((0.0, 0.0, 0.0),),
                 subcycle: bool = True, do_reflux: bool = True, dtype=torch.float64,
                 static=None, boxes: Optional[dict] = None):
        if model not in MODELS:
            raise ValueError(f"unknown model {model!r}")
        self.model, self.solver = model, MODELS[model]
        self.max_level, self.regrid_int = max_level, regrid_int
        self.bf, self.mgs, self.nbuf = blocking_factor, max_grid_size, n_error_buf
        self.eff, self.nest = grid_eff, max(nest, 2)
        self.tag_thresh = list(tag_thresh)
        self.subcycle, self.do_reflux, self.dtype = subcycle, do_reflux, dtype
        self.static, self.fixed = static, dict(boxes or {})
        self.bc_lo, self.bc_hi = tuple(bc_lo), tuple(bc_hi)
        h = np.asarray(h, np.float64)
        ny, nx = h.shape
        lv0 = DenseLevel(nx, ny, dx, dy, bc_lo, bc_hi, dtype)
        self._load(lv0, h, hu, hv, z, n)
        lv0.set_boxes([(0, 0, nx, ny)])
        self.levels: List[DenseLevel] = [lv0]
        self.dt = [0.0] * (max_level + 1)
        self.istep = [0] * (max_level + 1)
        self.last_regrid = [0] * (max_level + 1)
        self.regs: List[List[torch.Tensor]] = [[] for _ in range(max_level + 1)]
        self.time = 0.0
        self.iteration = 0
        self.ctx = context()
        self._make_grids(0, 0.0, from_scratch=True)

    # -- construction ----------------------------------------------------
    def _load(self, lv: DenseLevel, h, hu, hv, z, n) -> None:
        T = lambda a: torch.as_tensor(_pad_edges(np.asarray(a, np.float64)), dtype=self.dtype,
                                      device="cuda")
        lv.terrain[0].copy_(T(z))
        lv.terrain[1].copy_(T(n))
        for k, a in enumerate((h, hu, hv)):
            lv.U_new[k].copy_(T(a))

    @property
    def finest(self) -> int:
        return len(self.levels) - 1

    def _new_level(self, lev: int) -> DenseLevel:
        c = self.levels[lev - 1]
        lv = DenseLevel(2 * c.nx, 2 * c.ny, c.dx / 2, c.dy / 2, self.bc_lo, self.bc_hi, self.dtype)
        st = self.static(lev) if self.static is not None else None
        if st is not None:
            self._load(lv, *st)
            lv.from_static = True
        else:
            zc = c.interior(c.terrain).double().cpu().numpy()
            lv.terrain.copy_(torch.as_tensor(np.stack([_pad_edges(_pc_refine(zc[k])) for k in range(2)]),
                                             dtype=self.dtype, device="cuda"))
            lv.from_static = False
        return lv

    def _make_grids(self, lbase: int, time: float, from_scratch: bool = False) -> None:
        """New grids for levels lbase + 1 .. max_level (AmrMesh::MakeNewGrids +
        MakeNewLevelFromScratch / MakeNewLevelFromCoarse / RemakeLevel)."""
        lib = L.load()
        for k in range(lbase, self.max_level):
            if k > self.finest:
                break
            lv = self.levels[k]
            if (k + 1) in self.fixed:
                new = list(self.fixed[k + 1])
                if not from_scratch and k + 1 <= self.finest:
                    continue
            else:
                if k > 0:
                    self._fill_ghosts(k, time, whole_level=True)
                th = self.tag_thresh[min(k, len(self.tag_thresh) - 1)]
                tags = torch.empty((lv.ny, lv.nx), dtype=torch.uint8, device="cuda")
                thr = (C.c_double * 3)(*th)
                full = lv.patch(lv.U_new)
                full.ng = 0
                for d in range(2):
                    full.bc_lo[d], full.bc_hi[d] = self.bc_lo[d], self.bc_hi[d]
                L.check(lib.hx_tag_level(self.ctx, self.solver, C.byref(full), float(time), thr,
                                         tags.data_ptr(), _stream()), "hx_tag_level")
                from . import regrid as RG
                b = max(self.bf // 2, 1)
                aligned = (lv.nx % b == 0 and lv.ny % b == 0 and self.nest <= b and
                           not (np.asarray(lv.boxes, np.int64) % b).any())
                if aligned:
                    # tags grown, masked and blocked on the device; only the
                    # block array comes to the host
                    blk = torch.empty((lv.ny // b, lv.nx // b), dtype=torch.uint8, device="cuda")
                    L.check(lib.hx_tag_blocks(self.ctx, tags.data_ptr(), lv.nx,
                                              lv.mask_addr(lv.mask, 0, 0), lv.pitch, lv.nx, lv.ny,
                                              self.nbuf, b, blk.data_ptr(), _stream()),
                            "hx_tag_blocks")
                    allowed = RG.allowed_blocks(lv.boxes, lv.ny, lv.nx, b, self.nest)
                    new = RG.grids_from_blocks(blk.bool().cpu().numpy(), allowed, b, 2, self.eff,
                                               self.mgs)
                else:
                    valid = lv.interior(lv.mask).bool().cpu().numpy()
                    new = RG.new_grids(tags.bool().cpu().numpy(), valid, 2, self.bf, self.mgs,
                                       self.nbuf, self.eff, self.nest)
            if not new:
                del self.levels[k + 1:]
                break
            if k + 1 > self.finest:
                fl = self._new_level(k + 1)
                self.levels.append(fl)
                old_mask = None
                fl.t_new, fl.t_old = time, time - 1.0e200
                self.dt[k + 1] = self.dt[k] / (2.0 if self.subcycle else 1.0)
            else:
                fl = self.levels[k + 1]
                old_mask = fl.mask.clone()
            if not fl.set_boxes(new) and old_mask is not None:
                continue                        # the same grids: nothing to remake
            if from_scratch and getattr(fl, "from_static", False):
                continue                        # MakeNewLevelFromScratch: the static layer
            cp = lv.patch(lv.U_new)
            cp.ng = 0
            for d in range(2):
                cp.bc_lo[d], cp.bc_hi[d] = self.bc_lo[d], self.bc_hi[d]
            # one fill over the whole level: every cell the level did not hold
            # takes the interpolation from the coarser level (each box's own
            # fill gives its cells the same values; cells outside the new
            # boxes are refilled as ghosts before any use)
            bp = fl.patch(fl.U_new)
            skip = fl.mask_addr(old_mask, 0, 0) if old_mask is not None else None
            L.check(lib.hx_fill_fine_region(
                self.ctx, C.byref(cp), C.byref(cp), 0.0, 2, 0, 0, C.byref(bp), 1, skip, fl.pitch,
                fl.addr(fl.U_new, 0, 0, 0), fl.addr(fl.U_new, 1, 0, 0), fl.addr(fl.U_new, 2, 0, 0),
                _stream()), "hx_fill_fine_region")

    # -- one level's pieces ------------------------------------------------
    def _crse_patch(self, lv: DenseLevel, U: torch.Tensor) -> L.HxPatch:
        p = lv.patch(U)
        p.ng = 0
        for d in range(2):
            p.bc_lo[d], p.bc_hi[d] = self.bc_lo[d], self.bc_hi[d]
        return p

    def _fill_ghosts(self, lev: int, t: float, whole_level: bool = False) -> None:
        """Ring cells of every box the level does not own, from the coarser
        level at time t (GetData's time selection: the new state, the old
        state, or their blend).  whole_level: every cell the level does not
        own, in one launch (the rings take the same values)."""
        c, f = self.levels[lev - 1], self.levels[lev]
        if _almost_equal(t, c.t_new):
            co = cn = c.U_new
            alpha = 0.0
        elif _almost_equal(t, c.t_old):
            co = cn = c.U_old
            alpha = 0.0
        else:
            co, cn = c.U_old, c.U_new
            alpha = (t - c.t_old) / (c.t_new - c.t_old)
        po, pn = self._crse_patch(c, co), self._crse_patch(c, cn)
        lib = L.load()
        st = _stream()
        if whole_level:
            bp = f.patch(f.U_new)
            L.check(lib.hx_fill_fine_region(
                self.ctx, C.byref(po), C.byref(pn), float(alpha), 2, 0, 0, C.byref(bp), 1,
                f.mask_addr(f.mask, 0, 0), f.pitch, f.addr(f.U_new, 0, 0, 0),
                f.addr(f.U_new, 1, 0, 0), f.addr(f.U_new, 2, 0, 0), st), "hx_fill_fine_region")
            return
        for b in f.patches:
            bp = f.patch(f.U_new, b)
            L.check(lib.hx_fill_fine_region(
                self.ctx, C.byref(po), C.byref(pn), float(alpha), 2, b[0] // 2, b[1] // 2,
                C.byref(bp), 0, f.mask_addr(f.mask, b[0], b[1]), f.pitch,
                f.addr(f.U_new, 0, b[0], b[1]), f.addr(f.U_new, 1, b[0], b[1]),
                f.addr(f.U_new, 2, b[0], b[1]), st), "hx_fill_fine_region")

    def _flux(self, patch, faces, pair, sx, sy, accumulate, reg) -> None:
        L.check(L.load().hx_face_fluxes(self.ctx, self.solver, C.byref(patch), faces.shape[0],
                                        C.c_void_p(faces.data_ptr()), pair, float(sx), float(sy),
                                        int(accumulate), C.c_void_p(reg.data_ptr()), _stream()),
                "hx_face_fluxes")

    def _advance_level(self, lev: int, t: float, dt: float) -> None:
        lv = self.levels[lev]
        lib = L.load()
        if lev > 0:
            self._fill_ghosts(lev, t)
        if self.do_reflux and lev < self.finest:
            fl = self.levels[lev + 1]
            cp = self._crse_patch(lv, lv.U_new)
            self.regs[lev + 1] = []
            for b, (cf, _) in zip(fl.patches, fl.faces):
                reg = torch.zeros((cf.shape[0], 3), dtype=self.dtype, device="cuda")
                self._flux(cp, cf, 1, -dt * lv.dy, -dt * lv.dx, 0, reg)
                self.regs[lev + 1].append(reg)
        if self.do_reflux and lev > 0:
            for b, (_, ff), reg in zip(lv.patches, lv.faces, self.regs[lev]):
                self._flux(lv.patch(lv.U_new, b), ff, 2, dt * lv.dy, dt * lv.dx, 1, reg)
        lv.swap()
        for b in lv.patches:
            L.check(lib.hx_advance_level(self.ctx, self.solver, C.byref(lv.patch(lv.U_old, b)),
                                         float(dt), lv.addr(lv.U_new, 0, b[0], b[1]),
                                         lv.addr(lv.U_new, 1, b[0], b[1]),
                                         lv.addr(lv.U_new, 2, b[0], b[1]), lv.pitch, _stream()),
                    "hx_advance_level")

    def _reflux(self, lev: int) -> None:
        lv, fl = self.levels[lev], self.levels[lev + 1]
        cp = self._crse_patch(lv, lv.U_new)
        lib = L.load()
        for orient in range(4):            # AMReX Orientation order: x-lo, y-lo, x-hi, y-hi
            for b, reg in zip(fl.patches, self.regs[lev + 1]):
                L.check(lib.hx_reflux_side(self.ctx, C.byref(cp), lv.addr(lv.U_new, 0, 0, 0),
                                           lv.addr(lv.U_new, 1, 0, 0), lv.addr(lv.U_new, 2, 0, 0),
                                           b[0] // 2, b[1] // 2, b[2] // 2, b[3] // 2,
                                           C.c_void_p(reg.data_ptr()), orient, _stream()),
                        "hx_reflux_side")

    def _average_down(self, lev: int) -> None:
        lv, fl = self.levels[lev], self.levels[lev + 1]
        cp = self._crse_patch(lv, lv.U_new)
        lib = L.load()
        for b in fl.patches:
            L.check(lib.hx_average_down_state(self.ctx, C.byref(fl.patch(fl.U_new, b)), 2,
                                              b[0] // 2, b[1] // 2, C.byref(cp),
                                              lv.addr(lv.U_new, 0, 0, 0), lv.addr(lv.U_new, 1, 0, 0),
                                              lv.addr(lv.U_new, 2, 0, 0), _stream()),
                    "hx_average_down_state")

    # -- dt and the recursive step -----------------------------------------
    def level_dt(self, lev: int, cfl: float) -> float:
        lv = self.levels[lev]
        arr = (L.HxPatch * len(lv.patches))(*[lv.patch(lv.U_new, b) for b in lv.patches])
        dt, vel = C.c_double(), C.c_double()
        L.check(L.load().hx_boxes_dt(self.ctx, self.solver, len(lv.patches), arr, float(cfl),
                                     C.byref(dt), C.byref(vel), _stream()), "hx_boxes_dt")
        return dt.value

    def compute_dt(self, cfl: float) -> List[float]:
        """BaseSolver::compute_dt at ratio 2, then the lockstep rule."""
        tmp = [self.level_dt(l, cfl) for l in range(self.finest + 1)]
        dt0, factor = tmp[0], 1.0
        for l in range(1, self.finest + 1):
            factor *= 2.0
            dt0 = min(dt0, tmp[l] * factor)
        self.dt[0] = dt0
        for l in range(1, self.finest + 1):
            self.dt[l] = self.dt[l - 1] / 2.0
        if not self.subcycle:
            m = min(self.dt[:self.finest + 1])
            for l in range(self.finest + 1):
                self.dt[l] = m
        return self.dt[:self.finest + 1]

    def _time_step(self, lev: int, time: float, iteration: int) -> None:
        if (self.regrid_int > 0 and lev < self.max_level and self.istep[lev] > self.last_regrid[lev]
                and self.istep[lev] % self.regrid_int == 0):
            old_finest = self.finest
            self._make_grids(lev, time)
            for k in range(lev, self.finest + 1):
                self.last_regrid[k] = self.istep[k]
            for k in range(old_finest + 1, self.finest + 1):
                self.dt[k] = self.dt[k - 1] / (2.0 if self.subcycle else 1.0)
        lv = self.levels[lev]
        lv.t_old = lv.t_new
        lv.t_new += self.dt[lev]
        self._advance_level(lev, lv.t_old, self.dt[lev])
        self.istep[lev] += 1
        if lev < self.finest:
            nsub = 2 if self.subcycle else 1
            for i in range(1, nsub + 1):
                self._time_step(lev + 1, time + (i - 1) * self.dt[lev + 1], i)
            if self.do_reflux:
                self._reflux(lev)
            self._average_down(lev)

    def step(self, cfl: Optional[float] = None, dts: Optional[Sequence[float]] = None) -> float:
        """ComputeDt then one coarse step (HydroEXA::Compute's loop body); dts
        gives every level's dt instead."""
        if dts is None:
            self.compute_dt(cfl)
        else:
            for l, d in enumerate(dts):
                self.dt[l] = float(d)
        self._time_step(0, self.time, self.iteration)
        self.time = self.levels[0].t_new
        self.iteration += 1
        return self.dt[0]

    def run_until(self, t_end: float, cfl: float, max_steps: int = 10 ** 9) -> int:
        k = 0
        while self.time < t_end and k < max_steps:
            self.step(cfl)
            k += 1
        return k

    # -- native run (hx_amr_run) -------------------------------------------
    def _publish(self) -> None:
        """Write the level table hx_amr_run reads (current state in u[0])."""
        for l, lv in enumerate(self.levels):
            t = self._table[l]
            t.nx, t.ny, t.npatch, t.cur = lv.nx, lv.ny, len(lv.patches), 0
            t.patches = lv.patch_arr.ctypes.data_as(C.POINTER(C.c_int32))
            t.u[0], t.u[1] = lv.U_new.data_ptr(), lv.U_old.data_ptr()
            t.terrain, t.mask, t.regs = lv.terrain.data_ptr(), lv.mask.data_ptr(), lv.regbuf.data_ptr()
            t.dx, t.dy, t.t_new, t.t_old = lv.dx, lv.dy, lv.t_new, lv.t_old
            t.istep, t.last_regrid = self.istep[l], self.last_regrid[l]
        self._nlev.value = len(self.levels)

    def _absorb(self) -> None:
        """Take back buffer roles, clocks and step counts from the table."""
        for l, lv in enumerate(self.levels):
            t = self._table[l]
            if t.cur == 1:
                lv.swap()
            lv.t_new, lv.t_old = t.t_new, t.t_old
            self.istep[l], self.last_regrid[l] = t.istep, t.last_regrid

    def run_native(self, t_end: float, cfl: float, max_steps: int = 2 ** 31 - 1) -> np.ndarray:
        """Coarse steps until t_end (or max_steps) in C++ (hx_amr_run): the same
        kernels and order as step(); the host is called back only to regrid.
        Returns the coarse dt of each step."""
        self._table = (L.HxAmrLevel * (self.max_level + 1))()
        self._nlev = C.c_int(len(self.levels))
        err = []

        def regrid(user, lev, time):
            try:
                self._absorb()
                self._make_grids(lev, time)
                self._publish()
                return 0
            except Exception as e:          # reported after the run
                err.append(e)
                return 1

        cb = L.HX_AMR_REGRID_FN(regrid)
        self._publish()
        n = min(max_steps, 10 ** 7)
        hist = np.zeros(max(n, 1) if n < 100000 else 100000)
        tm = C.c_double(self.time)
        done = C.c_int32()
        lo = (C.c_int32 * 2)(*self.bc_lo)
        hi = (C.c_int32 * 2)(*self.bc_hi)
        rc = L.load().hx_amr_run(self.ctx, self.solver, _DT[self.dtype], self._table,
                                 C.byref(self._nlev), self.max_level, lo, hi, int(self.subcycle),
                                 int(self.do_reflux), int(self.regrid_int), cb, None, float(cfl),
                                 float(t_end), int(min(n, len(hist))), C.byref(tm),
                                 hist.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done),
                                 _stream())
        self._absorb()
        self.time = tm.value
        self.iteration += done.value
        if err:
            raise err[0]
        if rc == L.HX_E_DT_COLLAPSE:
            raise L.DtCollapse(rc, f"hx_amr_run (completed {done.value} steps)")
        L.check(rc, "hx_amr_run")
        return hist[:done.value]

    # -- diagnostics and output ----------------------------------------------
    def level_stats(self, lev: int) -> np.ndarray:
        """(mass, px, py, KE, min h) over the valid cells of a level: the
        per-level reduction of AmrMeshState::ComputeMassDiagnostics
        (src/state/AmrMeshState_Utils.cpp:213-299) summed over every patch
        (the reference assigns each box's result, keeping its last box)."""
        lv = self.levels[lev]
        out = np.zeros(5)
        out[4] = 1.7976931348623157e308
        r = np.zeros(5)
        for b in lv.patches:
            L.check(L.load().hx_conservation(self.ctx, C.byref(lv.patch(lv.U_new, b)),
                                             r.ctypes.data_as(C.POINTER(C.c_double)), _stream()),
                    "hx_conservation")
            out[:4] += r[:4]
            out[4] = min(out[4], r[4])
        return out

    def conservation(self) -> dict:
        """The reference's totals (sums over levels, so a refined region counts
        once per level that holds it; src/state/AmrMeshState_Utils.cpp:311-321)
        and the composite mass and momenta (level 0 after average_down)."""
        per = [self.level_stats(l) for l in range(self.finest + 1)]
        tot = np.sum([p[:4] for p in per], axis=0)
        return {"levels": per, "mass": tot[0], "px": tot[1], "py": tot[2], "KE": tot[3],
                "min_h": min(p[4] for p in per), "composite_mass": per[0][0],
                "composite_px": per[0][1], "composite_py": per[0][2]}

    def write_plotfile(self, prefix: str, iteration: int) -> str:
        """Native AMReX plotfile of every level on its grids (WriteNativePlotfile,
        src/io/NativePlotFile.cpp:19-97, through hydroexa_amd.plotfile)."""
        from types import SimpleNamespace
        from .plotfile import write_native_plotfile
        views, boxes = [], []
        for lv in self.levels:
            views.append(SimpleNamespace(nx=lv.nx, ny=lv.ny, dx=lv.dx, dy=lv.dy,
                                         U_new=lv.interior(lv.U_new), terrain=lv.interior(lv.terrain)))
            boxes.append([(i0, j0, i0 + w - 1, j0 + h - 1) for (i0, j0, w, h) in lv.boxes])
        return write_native_plotfile(prefix, iteration, self.time, views,
                                     ref_ratio=[2] * self.finest, boxes=boxes)

    # -- views -------------------------------------------------------------
    def state(self, lev: int) -> torch.Tensor:
        lv = self.levels[lev]
        return lv.interior(lv.U_new)

    def composite_mass(self) -> float:
        lv = self.levels[0]
        return float(lv.interior(lv.U_new)[0].double().sum().item()) * lv.dx * lv.dy

    def centerline(self, row_frac: float = 0.5):
        """x and h along a row of the finest data (each coarse cell replaced by
        its finer cells where a finer level holds them)."""
        xs, hs = [], []
        for l in range(self.finest + 1):
            lv = self.levels[l]
            j = min(int(row_frac * lv.ny), lv.ny - 1)
            m = lv.interior(lv.mask)[j].bool().cpu().numpy()
            if l < self.finest:
                fm = self.levels[l + 1]
                jf = min(int(row_frac * fm.ny), fm.ny - 1)
                covered = fm.interior(fm.mask)[jf].bool().cpu().numpy().reshape(-1, 2).any(axis=1)
                m = m & ~covered
            h = lv.interior(lv.U_new)[0, j].double().cpu().numpy()
            x = (np.arange(lv.nx) + 0.5) * lv.dx
            xs.append(x[m])
            hs.append(h[m])
        x, h = np.concatenate(xs), np.concatenate(hs)
        o = np.argsort(x)
        return x[o], h[o]
