"""Differentiable single-level rollouts: a ``torch.autograd.Function`` whose
backward is the library's hand-written adjoint (``hx_adjoint_*``).

No counterpart exists in the reference (HydroEXA computes no derivatives);
the forward is the same step as ``SolverVariant.compute_fluxes``
(src/solvers/Roe.cpp:250-434, src/solvers/HydrostaticRoe.cpp:400-606) driven
step by step like ``SolverVariant.run``.

A :class:`Rollout` fixes the geometry, solver, step count, checkpoint
interval and gauge cells, and owns the device workspace (checkpoints,
segment states, adjoint buffers).  Calling it runs the forward through the
autograd Function and returns the gauge series (depth at each gauge after
every step) and the final state.  The backward recomputes each segment of
``ckpt_every`` steps from its checkpoint and runs the adjoint backwards
through it, with the forward's dt schedule held fixed (dt is not
differentiated).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from .. import _lib as L
from ..solver import MODELS, context

_dp = C.POINTER(C.c_double)


def _vp(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


class Rollout:
    """A configured FP64 rollout with its adjoint workspace.

    model: "Roe" or "HydrostaticRoe"; gauges: flat cell indices j*nx + i.
    cfl > 0: each forward picks dt by the CFL rule (as ``SolverVariant.run``);
    with ``freeze_dt`` the first forward's schedule is reused by later ones,
    so that repeated evaluations (calibration) differentiate one smooth map.
    dt_schedule: a fixed schedule instead (cfl is then ignored)."""

    def __init__(self, model: str, nx: int, ny: int, dx: float, dy: float,
                 bc_lo: Sequence[int], bc_hi: Sequence[int], nsteps: int,
                 gauges: Sequence[int], ckpt_every: int = 20, cfl: float = 0.45,
                 dt_schedule: Optional[Sequence[float]] = None, freeze_dt: bool = False,
                 device: Optional[int] = None):
        if model not in MODELS:
            raise ValueError(f"unknown model {model!r}")
        self.model, self.nx, self.ny = model, int(nx), int(ny)
        self.dx, self.dy = float(dx), float(dy)
        self.bc_lo, self.bc_hi = tuple(bc_lo), tuple(bc_hi)
        self.nsteps, self.ckpt_every = int(nsteps), int(ckpt_every)
        self.gauges = np.ascontiguousarray(np.asarray(gauges, dtype=np.int64).reshape(-1))
        self.cfl = float(cfl)
        self.freeze_dt = bool(freeze_dt)
        self.dt_schedule = None if dt_schedule is None else np.ascontiguousarray(
            np.asarray(dt_schedule, dtype=np.float64))
        if self.dt_schedule is not None and self.dt_schedule.shape != (self.nsteps,):
            raise ValueError("dt_schedule must hold one dt per step")
        self.device = torch.cuda.current_device() if device is None else int(device)
        geom = L.HxPatch()
        geom.nx, geom.ny, geom.ng, geom.dtype = self.nx, self.ny, 0, L.HX_F64
        geom.pitch = self.nx
        geom.dx, geom.dy = self.dx, self.dy
        geom.bc_lo[0], geom.bc_lo[1] = self.bc_lo
        geom.bc_hi[0], geom.bc_hi[1] = self.bc_hi
        self._h = C.c_void_p()
        gp = self.gauges.ctypes.data_as(C.POINTER(C.c_int64)) if self.gauges.size else None
        L.check(L.load().hx_adjoint_create(context(self.device), MODELS[model], C.byref(geom),
                                           self.nsteps, self.ckpt_every, int(self.gauges.size),
                                           gp, C.byref(self._h)), "hx_adjoint_create")
        self._generation = 0
        self.last_dt = None

    # -- raw passes (no autograd) -------------------------------------------
    def _patch(self, U0: torch.Tensor, terrain: torch.Tensor) -> L.HxPatch:
        plane = self.nx * self.ny * 8
        p = L.HxPatch()
        p.nx, p.ny, p.ng, p.dtype = self.nx, self.ny, 0, L.HX_F64
        p.pitch = self.nx
        b = U0.data_ptr()
        p.h, p.hu, p.hv = b, b + plane, b + 2 * plane
        p.eta = None
        t = terrain.data_ptr()
        p.z, p.n = t, t + plane
        p.dx, p.dy = self.dx, self.dy
        p.bc_lo[0], p.bc_lo[1] = self.bc_lo
        p.bc_hi[0], p.bc_hi[1] = self.bc_hi
        return p

    def forward_raw(self, U0: torch.Tensor, terrain: torch.Tensor):
        """U0 (3, ny, nx) and terrain (2, ny, nx), FP64 device tensors.
        Returns (series [nsteps, ngauge], final state (3, ny, nx))."""
        for name, t, k in (("state", U0, 3), ("terrain", terrain, 2)):
            if (t.dtype != torch.float64 or not t.is_cuda or tuple(t.shape) != (k, self.ny, self.nx)
                    or not t.is_contiguous()):
                raise ValueError(f"{name} must be a contiguous float64 device tensor "
                                 f"of shape ({k}, {self.ny}, {self.nx})")
        series = torch.empty((self.nsteps, self.gauges.size), dtype=torch.float64, device=U0.device)
        out = torch.empty_like(U0)
        if self.dt_schedule is not None:
            sched, cfl = self.dt_schedule.copy(), -1.0
        else:
            sched, cfl = np.zeros(self.nsteps), self.cfl
        done = C.c_int32(0)
        plane = self.nx * self.ny * 8
        ob = out.data_ptr()
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = L.load().hx_adjoint_forward(self._h, C.byref(self._patch(U0, terrain)), cfl,
                                         sched.ctypes.data_as(_dp), _vp(series), C.c_void_p(ob),
                                         C.c_void_p(ob + plane), C.c_void_p(ob + 2 * plane),
                                         C.byref(done), st)
        L.check(rc, "hx_adjoint_forward")
        self.last_dt = sched
        if self.freeze_dt and self.dt_schedule is None:
            self.dt_schedule = sched.copy()
        self._generation += 1
        return series, out

    def backward_raw(self, g_series: Optional[torch.Tensor], g_final: Optional[torch.Tensor]):
        """Adjoint of the last forward.  Returns (g_state (3, ny, nx),
        g_terrain (2, ny, nx))."""
        dev = torch.device("cuda", self.device)
        gs = None if g_series is None else g_series.to(dev, torch.float64).contiguous()
        gf = None if g_final is None else g_final.to(dev, torch.float64).contiguous()
        g_state = torch.empty((3, self.ny, self.nx), dtype=torch.float64, device=dev)
        g_ter = torch.empty((2, self.ny, self.nx), dtype=torch.float64, device=dev)
        plane = self.nx * self.ny * 8
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = L.load().hx_adjoint_backward(self._h, _vp(gs), _vp(gf), _vp(g_state),
                                          C.c_void_p(g_ter.data_ptr()),
                                          C.c_void_p(g_ter.data_ptr() + plane), st)
        L.check(rc, "hx_adjoint_backward")
        return g_state, g_ter

    # -- autograd -------------------------------------------------------------
    def __call__(self, h0: torch.Tensor, hu0: torch.Tensor, hv0: torch.Tensor, z: torch.Tensor,
                 n: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
        """Differentiable rollout: returns (series, h, hu, hv)."""
        return ShallowWaterRollout.apply(self, h0, hu0, hv0, z, n)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            L.load().hx_adjoint_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RolloutBatch:
    """Many FP64 rollouts (scenarios of uneven sizes, each with its own
    terrain, gauges and CFL dt) advanced and differentiated together:
    hx_adjoint_batch_* makes one launch per step and per adjoint kernel for
    the whole batch.  Each scenario's results equal a :class:`Rollout` of
    that scenario alone, bit for bit.

    scenarios: a list of dicts with nx, ny, dx, dy, bc_lo, bc_hi, gauges; all
    share model, nsteps, ckpt_every and cfl.  freeze_dt as in Rollout."""

    def __init__(self, model: str, scenarios: Sequence[dict], nsteps: int, ckpt_every: int = 20,
                 cfl: float = 0.45, freeze_dt: bool = False, device: Optional[int] = None):
        if model not in MODELS:
            raise ValueError(f"unknown model {model!r}")
        if not scenarios:
            raise ValueError("a batch needs at least one scenario")
        self.model = model
        self.nsteps, self.ckpt_every, self.cfl = int(nsteps), int(ckpt_every), float(cfl)
        self.freeze_dt = bool(freeze_dt)
        self.dt_schedule = None
        self.device = torch.cuda.current_device() if device is None else int(device)
        self.sc = [dict(s) for s in scenarios]
        n = len(self.sc)
        geoms = (L.HxPatch * n)()
        self._gauges = []
        for k, s in enumerate(self.sc):
            g = geoms[k]
            g.nx, g.ny, g.ng, g.dtype = int(s["nx"]), int(s["ny"]), 0, L.HX_F64
            g.pitch = g.nx
            g.dx, g.dy = float(s["dx"]), float(s["dy"])
            g.bc_lo[0], g.bc_lo[1] = s["bc_lo"]
            g.bc_hi[0], g.bc_hi[1] = s["bc_hi"]
            self._gauges.append(np.ascontiguousarray(np.asarray(s["gauges"], np.int64).reshape(-1)))
        ng = (C.c_int * n)(*[a.size for a in self._gauges])
        gp = (C.POINTER(C.c_int64) * n)(*[a.ctypes.data_as(C.POINTER(C.c_int64)) for a in self._gauges])
        self._geoms = geoms
        self._h = C.c_void_p()
        L.check(L.load().hx_adjoint_batch_create(context(self.device), MODELS[model], n, geoms,
                                                 self.nsteps, self.ckpt_every, ng, gp,
                                                 C.byref(self._h)), "hx_adjoint_batch_create")
        self.last_dt = None

    def forward_raw(self, U0s: Sequence[torch.Tensor], terrains: Sequence[torch.Tensor]):
        """Per scenario U0 (3, ny, nx) and terrain (2, ny, nx), contiguous FP64
        device tensors.  Returns (series list [nsteps, ngauge], finals list)."""
        n = len(self.sc)
        if len(U0s) != n or len(terrains) != n:
            raise ValueError("one state and one terrain per scenario")
        states = (L.HxPatch * n)()
        series, finals = [], []
        ser = (C.c_void_p * n)()
        fin = (C.c_void_p * (3 * n))()
        for k, (s, U, T) in enumerate(zip(self.sc, U0s, terrains)):
            ny, nx = int(s["ny"]), int(s["nx"])
            for name, t, m in (("state", U, 3), ("terrain", T, 2)):
                if (t.dtype != torch.float64 or not t.is_cuda or tuple(t.shape) != (m, ny, nx)
                        or not t.is_contiguous()):
                    raise ValueError(f"scenario {k}: {name} must be a contiguous float64 device "
                                     f"tensor of shape ({m}, {ny}, {nx})")
            plane = nx * ny * 8
            p = states[k]
            p.nx, p.ny, p.ng, p.dtype, p.pitch = nx, ny, 0, L.HX_F64, nx
            b, t = U.data_ptr(), T.data_ptr()
            p.h, p.hu, p.hv, p.eta, p.z, p.n = b, b + plane, b + 2 * plane, None, t, t + plane
            p.dx, p.dy = float(s["dx"]), float(s["dy"])
            p.bc_lo[0], p.bc_lo[1] = s["bc_lo"]
            p.bc_hi[0], p.bc_hi[1] = s["bc_hi"]
            se = torch.empty((self.nsteps, self._gauges[k].size), dtype=torch.float64, device=U.device)
            out = torch.empty_like(U)
            series.append(se)
            finals.append(out)
            ser[k] = se.data_ptr() if se.numel() else None
            o = out.data_ptr()
            fin[3 * k], fin[3 * k + 1], fin[3 * k + 2] = o, o + plane, o + 2 * plane
        if self.dt_schedule is not None:
            sched, cfl = self.dt_schedule.copy(), -1.0
        else:
            sched, cfl = np.zeros((n, self.nsteps)), self.cfl
        done = (C.c_int32 * n)()
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = L.load().hx_adjoint_batch_forward(self._h, states, cfl, sched.ctypes.data_as(_dp), ser,
                                               fin, done, st)
        if rc == L.HX_E_DT_COLLAPSE:
            e = L.DtCollapse(rc, "hx_adjoint_batch_forward")
            e.done = np.frombuffer(done, dtype=np.int32).copy()
            raise e
        L.check(rc, "hx_adjoint_batch_forward")
        self.last_dt = sched
        if self.freeze_dt and self.dt_schedule is None:
            self.dt_schedule = sched.copy()
        return series, finals

    def backward_raw(self, g_series: Optional[Sequence[Optional[torch.Tensor]]],
                     g_finals: Optional[Sequence[Optional[torch.Tensor]]]):
        """Adjoint of the last forward.  Returns (g_state list, g_terrain list)."""
        n = len(self.sc)
        dev = torch.device("cuda", self.device)
        keep = []

        def arr(ts):
            if ts is None:
                return None
            a = (C.c_void_p * n)()
            for k, t in enumerate(ts):
                if t is not None:
                    t = t.to(dev, torch.float64).contiguous()
                    keep.append(t)
                    a[k] = t.data_ptr()
            return a
        gs, gf = arr(g_series), arr(g_finals)
        g_state, g_ter = [], []
        a_st, a_z, a_n = (C.c_void_p * n)(), (C.c_void_p * n)(), (C.c_void_p * n)()
        for k, s in enumerate(self.sc):
            ny, nx = int(s["ny"]), int(s["nx"])
            gst = torch.empty((3, ny, nx), dtype=torch.float64, device=dev)
            gt = torch.empty((2, ny, nx), dtype=torch.float64, device=dev)
            g_state.append(gst)
            g_ter.append(gt)
            a_st[k] = gst.data_ptr()
            a_z[k] = gt.data_ptr()
            a_n[k] = gt.data_ptr() + nx * ny * 8
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = L.load().hx_adjoint_batch_backward(self._h, gs, gf, a_st, a_z, a_n, st)
        L.check(rc, "hx_adjoint_batch_backward")
        return g_state, g_ter

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            L.load().hx_adjoint_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ShallowWaterRollout(torch.autograd.Function):
    """(h0, hu0, hv0, z, n) -> (gauge series, h, hu, hv) through
    ``rollout.nsteps`` steps; backward = hand-written adjoint with
    checkpointing.  Only the most recent forward of a Rollout can be
    differentiated (its checkpoints are the ones kept)."""

    @staticmethod
    def forward(ctx, rollout: Rollout, h0, hu0, hv0, z, n):
        U0 = torch.stack([h0, hu0, hv0]).to(torch.float64).contiguous()
        terrain = torch.stack([z, n]).to(torch.float64).contiguous()
        series, out = rollout.forward_raw(U0, terrain)
        ctx.rollout = rollout
        ctx.generation = rollout._generation
        ctx.terrain = terrain               # the adjoint reads z and n in place
        return series, out[0], out[1], out[2]

    @staticmethod
    def backward(ctx, g_series, g_h, g_hu, g_hv):
        rollout: Rollout = ctx.rollout
        if rollout._generation != ctx.generation:
            raise RuntimeError("Rollout: another forward ran after this one; only the most "
                               "recent forward of a Rollout can be differentiated")
        if ctx.terrain is None:
            raise RuntimeError("Rollout: backward called twice on one forward")
        if g_h is None and g_hu is None and g_hv is None:
            g_final = None
        else:
            like = g_h if g_h is not None else (g_hu if g_hu is not None else g_hv)
            z0 = torch.zeros_like(like, dtype=torch.float64)
            g_final = torch.stack([z0 if g is None else g.to(torch.float64)
                                   for g in (g_h, g_hu, g_hv)])
        g_state, g_ter = rollout.backward_raw(g_series, g_final)
        ctx.terrain = None
        need = ctx.needs_input_grad
        out = [None]
        for k, t in enumerate((g_state[0], g_state[1], g_state[2], g_ter[0], g_ter[1])):
            out.append(t if need[k + 1] else None)
        return tuple(out)
