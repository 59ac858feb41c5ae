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
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib as L
from .solver import MODELS, _DT, Level, context

NG = 2


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

    def composite_mass(self) -> float:
        """Sum of h dA over the coarse level (covered cells hold the average
        of their fine children after each step)."""
        return float(self.crse.U_new[0].double().sum().item()) * self.crse.dx * self.crse.dy
