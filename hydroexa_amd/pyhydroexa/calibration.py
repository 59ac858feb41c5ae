"""Calibration of Manning roughness and bathymetry against gauge depths, by
gradient descent through the differentiable rollout.

The synthetic problem: a square level of n x n cells of size dx (default
2048 x 2048 at 10 m) with a bed of 64 seeded Gaussian hills, a NODATA
margin (z = n = -9999, h = 0, the reference's preprocessing convention,
PyHydroEXA/src/pyhydroexa/preprocessing/hdf5.py:167-181) and a flood wave:
the free surface starts at eta_hi over the low-x strip and eta_lo elsewhere.
Roughness is piecewise constant on a ctrl x ctrl control grid; the hidden
field n_true makes the observed series at seeded wet gauges; calibration
starts from a uniform guess.  Solver: hydrostatic Roe with implicit Manning
friction, reflective sides, CFL-chosen dt frozen from the first forward.
Gauges sit within the reach of the waves from the dam (upstream
rarefaction, downstream bore): in still water the well-balanced scheme
keeps the velocity exactly zero, so roughness has no effect there.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Sequence

import numpy as np
import torch

from ..workloads import NODATA, SEED, splitmix64_uniform
from .differentiable import Rollout


@dataclass
class CalibrationProblem:
    n: int
    dx: float
    ctrl: int
    block: int
    z: torch.Tensor            # (n, n) bed, NODATA in the margin
    nodata: torch.Tensor       # (n, n) bool
    h0: torch.Tensor
    hu0: torch.Tensor
    hv0: torch.Tensor
    n_true: torch.Tensor       # (ctrl, ctrl)
    gauges: np.ndarray         # flat cell indices
    nsteps: int
    ckpt_every: int
    rollout: Rollout
    observed: torch.Tensor     # (nsteps, ngauge)
    meta: Dict = field(default_factory=dict)

    def roughness(self, n_ctrl: torch.Tensor) -> torch.Tensor:
        """Per-cell roughness from the control grid (NODATA kept)."""
        cell = n_ctrl.repeat_interleave(self.block, 0).repeat_interleave(self.block, 1)
        return torch.where(self.nodata, torch.full_like(cell, NODATA), cell)

    def bed(self, dz_ctrl: torch.Tensor = None) -> torch.Tensor:
        """Bed with an optional control-grid correction (NODATA kept)."""
        if dz_ctrl is None:
            return self.z
        cell = dz_ctrl.repeat_interleave(self.block, 0).repeat_interleave(self.block, 1)
        return torch.where(self.nodata, self.z, self.z + cell)

    def simulate(self, n_ctrl: torch.Tensor, dz_ctrl: torch.Tensor = None) -> torch.Tensor:
        series, _, _, _ = self.rollout(self.h0, self.hu0, self.hv0, self.bed(dz_ctrl),
                                       self.roughness(n_ctrl))
        return series

    def loss(self, n_ctrl: torch.Tensor, dz_ctrl: torch.Tensor = None) -> torch.Tensor:
        """Mean squared depth misfit over gauges and steps."""
        return ((self.simulate(n_ctrl, dz_ctrl) - self.observed) ** 2).mean()


def synthetic_problem(n: int = 2048, dx: float = 10.0, ctrl: int = 64, hills: int = 64,
                      margin: int = 32, ngauge: int = 32, nsteps: int = 400, ckpt_every: int = 20,
                      eta_hi: float = 2.0, eta_lo: float = 1.0, wave_frac: float = 0.25,
                      seed: int = SEED, cfl: float = 0.45, device: str = "cuda") -> CalibrationProblem:
    if n % ctrl:
        raise ValueError("the control grid must divide the level")
    block = n // ctrl
    L = n * dx
    u = splitmix64_uniform(seed + 11, 4 * hills).reshape(hills, 4)
    x = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    z = np.zeros((n, n))
    for A, cx, cy, s in u:
        amp, sig = 0.2 + 0.6 * A, (0.02 + 0.06 * s) * L
        z += amp * np.exp(-((X - cx * L) ** 2 + (Y - cy * L) ** 2) / (2.0 * sig * sig))
    nod = np.zeros((n, n), dtype=bool)
    nod[:margin, :] = nod[-margin:, :] = True
    nod[:, :margin] = nod[:, -margin:] = True
    eta = np.where(X < wave_frac * L, eta_hi, eta_lo)
    h = np.where(nod, 0.0, np.maximum(eta - z, 0.0))
    z = np.where(nod, NODATA, z)
    # hidden roughness: 0.02 + 0.04 * smooth seeded field on the control grid
    r = splitmix64_uniform(seed + 13, ctrl * ctrl).reshape(ctrl, ctrl)
    k = np.array([1.0, 2.0, 1.0]) / 4.0
    for _ in range(3):
        r = np.apply_along_axis(lambda a: np.convolve(np.pad(a, 1, mode="edge"), k, "valid"), 0, r)
        r = np.apply_along_axis(lambda a: np.convolve(np.pad(a, 1, mode="edge"), k, "valid"), 1, r)
    r = (r - r.min()) / max(r.max() - r.min(), 1e-12)
    n_true = 0.02 + 0.04 * r
    # gauges: seeded wet interior cells within the waves' reach of the dam.
    # With dt = cfl / ((|u| + c)/dx + (|v| + c)/dy) the fastest signal covers
    # about cfl/2 cells per step, whatever dx is.
    reach = max(4, int(0.5 * cfl * nsteps))
    dam = int(wave_frac * n)
    g = splitmix64_uniform(seed + 17, 4 * ngauge * 8).reshape(-1, 2)
    gauges = []
    for a, b in g:
        i = max(margin, min(n - margin - 1, dam + int((-0.6 + 1.0 * a) * reach)))
        j = margin + int(b * (n - 2 * margin))
        if not nod[j, i] and h[j, i] > 0.05 and j * n + i not in gauges:
            gauges.append(j * n + i)
        if len(gauges) == ngauge:
            break
    gauges = np.array(gauges, dtype=np.int64)
    dev = torch.device(device)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)
    ro = Rollout("HydrostaticRoe", n, n, dx, dx, (0, 0), (0, 0), nsteps, gauges,
                 ckpt_every=ckpt_every, cfl=cfl, freeze_dt=True)
    prob = CalibrationProblem(n, dx, ctrl, block, T(z), torch.as_tensor(nod, device=dev), T(h),
                              T(np.zeros((n, n))), T(np.zeros((n, n))), T(n_true), gauges, nsteps,
                              ckpt_every, ro, torch.empty(0, device=dev),
                              meta={"hills": hills, "margin": margin, "eta": (eta_hi, eta_lo)})
    with torch.no_grad():
        prob.observed = prob.simulate(prob.n_true).clone()
    return prob


def calibrate(prob: CalibrationProblem, iters: int = 20, lr: float = 2e-3, n_init: float = 0.04,
              fit_bed: bool = False, lr_bed: float = 0.02) -> List[Dict]:
    """Adam on the control-grid roughness (and optionally a bed correction).
    Returns the per-iteration history: loss, mean roughness error over all
    control cells and over the cells the gauges see (non-zero gradient at
    the start), after each update."""
    n_ctrl = torch.full((prob.ctrl, prob.ctrl), n_init, dtype=torch.float64,
                        device=prob.z.device, requires_grad=True)
    params = [{"params": [n_ctrl], "lr": lr}]
    dz = None
    if fit_bed:
        dz = torch.zeros_like(n_ctrl, requires_grad=True)
        params.append({"params": [dz], "lr": lr_bed})
    opt = torch.optim.Adam(params)
    hist = []
    seen = None
    for it in range(iters):
        opt.zero_grad()
        loss = prob.loss(n_ctrl, dz)
        loss.backward()
        if seen is None:
            # control cells the gauges can see (the others get no gradient)
            seen = n_ctrl.grad != 0
        opt.step()
        with torch.no_grad():
            n_ctrl.clamp_(min=1e-3)
            err = (n_ctrl - prob.n_true).abs()
            hist.append({"iter": it, "loss": loss.item(), "n_mae": err.mean().item(),
                         "n_mae_seen": err[seen].mean().item() if seen.any() else float("nan"),
                         "seen_cells": int(seen.sum().item())})
    return hist
