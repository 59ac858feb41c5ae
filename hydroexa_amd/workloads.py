"""Synthetic inputs for the benchmarked and parity-tested workloads.

All grids live on the unit square with cell-centred coordinates
x_i = (i + 0.5) dx.  Random fields use SplitMix64 (seeded, reproducible in
C, numpy and torch alike): u_k = (mix(seed + (k+1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple

import numpy as np

NODATA = -9999.0
SEED = 20261203


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) from SplitMix64 started at ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class Workload:
    name: str
    solver: str                 # "Roe" or "HydrostaticRoe"
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    z: np.ndarray
    n: np.ndarray
    dx: float
    dy: float
    bc_lo: Tuple[int, int]
    bc_hi: Tuple[int, int]
    cfl: float
    steps: int
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.h.shape

    @property
    def cells(self) -> int:
        return int(self.h.size)


def _centres(nx: int, ny: int):
    dx, dy = 1.0 / nx, 1.0 / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    return X, Y, dx, dy


def dam_break(n: int = 256, steps: int = 500, cfl: float = 0.45, h_left: float = 1.0,
              h_right: float = 0.5) -> Workload:
    """Plain dam break on a flat, frictionless bed: h = h_left where the cell
    centre has x < 0.5, h_right elsewhere; transmissive sides; first-order
    plain Roe (the reference's Roe has no reconstruction)."""
    X, _, dx, dy = _centres(n, n)
    h = np.where(X < 0.5, h_left, h_right)
    zero = np.zeros_like(h)
    return Workload(f"dam_break_{n}x{n}_fp64_roe", "Roe", h, zero.copy(), zero.copy(),
                    zero.copy(), zero.copy(), dx, dy, (1, 1), (1, 1), cfl, steps)


def lake_at_rest(n: int = 1024, steps: int = 2000, cfl: float = 0.45, seed: int = SEED,
                 nodata_corner: int = None, eta: float = 1.0) -> Workload:
    """Lake at rest over a Gaussian bump plus seeded uniform noise of 0.01,
    with a NODATA corner (i, j < corner; z = roughness = -9999, h = 0), a
    reflective box and the hydrostatic-reconstruction solver."""
    if nodata_corner is None:
        nodata_corner = int(round(290 * n / 1024))
    X, Y, dx, dy = _centres(n, n)
    z = 0.8 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / 0.01)
    z = z + 0.01 * splitmix64_uniform(seed, n * n).reshape(n, n)
    h = np.maximum(eta - z, 0.0)
    rough = np.zeros_like(z)
    nd = np.zeros_like(z, dtype=bool)
    nd[:nodata_corner, :nodata_corner] = True
    z = np.where(nd, NODATA, z)
    rough = np.where(nd, NODATA, rough)
    h = np.where(nd, 0.0, h)
    zero = np.zeros_like(h)
    return Workload(f"lake_at_rest_{n}x{n}_hydrostatic", "HydrostaticRoe", h, zero.copy(),
                    zero.copy(), z, rough, dx, dy, (0, 0), (0, 0), cfl, steps,
                    meta={"nodata_corner": nodata_corner, "eta": eta})


def circular_dam_break(n: int = 512, steps: int = 1000, cfl: float = 0.45, seed: int = SEED,
                       manning: float = 0.03, bed_amp: float = 0.05,
                       solver: str = "HydrostaticRoe", bc: Tuple[int, int] = (1, 1)) -> Workload:
    """Cylinder of radius 0.2 at the centre of the unit square: depth h = 2.0
    inside (r < 0.2 at the cell centre) and 0.5 outside, at rest, over a
    seeded random bed z = bed_amp U[0, 1) (SplitMix64, seed + 1) with a
    uniform Manning coefficient.  Sides transmissive by default (the
    reference's circular dam break uses outflow sides,
    tests/dambreakcircle/inputs); bc = (0, 0) closes the box."""
    X, Y, dx, dy = _centres(n, n)
    z = bed_amp * splitmix64_uniform(seed + 1, n * n).reshape(n, n)
    r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
    h = np.where(r2 < 0.04, 2.0, 0.5)
    zero = np.zeros_like(h)
    return Workload(f"circular_dam_break_{n}x{n}_{solver}", solver, h, zero.copy(), zero.copy(),
                    z, np.full_like(z, manning), dx, dy, tuple(bc), tuple(bc), cfl, steps)


def boxes_of(nx: int, ny: int, box: int) -> np.ndarray:
    """Uniform box decomposition {ilo, jlo, ihi, jhi} (AMReX max_grid_size)."""
    out = []
    for j0 in range(0, ny, box):
        for i0 in range(0, nx, box):
            out.append((i0, j0, min(i0 + box, nx) - 1, min(j0 + box, ny) - 1))
    return np.array(out, dtype=np.int32)


def ensemble(nscen: int = 256, side_min: int = 128, side_max: int = 1024, seed: int = SEED,
             steps: int = 1000, cfl: float = 0.45, solver: str = "Roe"):
    """Batched dam-break ensemble: side ~ U{side_min..side_max} per scenario
    (seeded), dam at x_d ~ U[0.25, 0.75], h_L = 1, h_R = r h_L with
    r ~ U[0.1, 0.9], bed A exp(-|x - c|^2 / (2 s^2)) with A ~ U[0, 0.2],
    s ~ U[0.05, 0.15], c ~ U[0.3, 0.7]^2, no friction, transmissive sides.
    The free surface is set, h = max(eta - z, 0)."""
    u = splitmix64_uniform(seed + 7, 8 * nscen).reshape(nscen, 8)
    out = []
    for s in range(nscen):
        n = side_min + int(u[s, 0] * (side_max - side_min + 1))
        n = min(n, side_max)
        xd = 0.25 + 0.5 * u[s, 1]
        r = 0.1 + 0.8 * u[s, 2]
        A = 0.2 * u[s, 3]
        sg = 0.05 + 0.1 * u[s, 4]
        cx, cy = 0.3 + 0.4 * u[s, 5], 0.3 + 0.4 * u[s, 6]
        X, Y, dx, dy = _centres(n, n)
        z = A * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg * sg))
        eta = np.where(X < xd, 1.0, r)
        h = np.maximum(eta - z, 0.0)
        zero = np.zeros_like(h)
        out.append(Workload(f"ensemble_{s}_{n}x{n}", solver, h, zero.copy(), zero.copy(), z,
                            zero.copy(), dx, dy, (1, 1), (1, 1), cfl, steps,
                            meta={"dam": xd, "ratio": r, "bump": (A, sg, cx, cy)}))
    return out


def adjoint_ensemble(nscen: int = 128, side_min: int = 128, side_max: int = 512, seed: int = SEED,
                     steps: int = 100, cfl: float = 0.45, ngauge: int = 8, manning: float = 0.03):
    """Scenarios for the batched forward + adjoint: the dam breaks of
    ensemble() (seeded side, dam position, depth ratio and Gaussian bump)
    on the hydrostatic solver with a seeded Manning field
    (manning (1 + 0.5 U[0, 1))), and ngauge seeded gauges per scenario at
    wet cells within the waves' reach of the dam (cfl / 2 cells per step).
    Returns (workloads, gauges): gauges[s] holds flat cell indices."""
    ws = ensemble(nscen, side_min, side_max, seed=seed, steps=steps, cfl=cfl,
                  solver="HydrostaticRoe")
    gauges = []
    for s, w in enumerate(ws):
        n = w.h.shape[0]
        u = splitmix64_uniform(seed + 1000 + s, n * n + 2 * 64 * ngauge)
        w.n = manning * (1.0 + 0.5 * u[:n * n].reshape(n, n))
        reach = max(4, int(0.5 * cfl * steps))
        dam = int(w.meta["dam"] * n)
        g = []
        k = n * n
        while len(g) < ngauge and k + 1 < u.size:
            i = min(n - 1, max(0, dam + int((-1.0 + 2.0 * u[k]) * reach)))
            j = min(n - 1, int(u[k + 1] * n))
            k += 2
            if w.h[j, i] > 0.05 and j * n + i not in g:
                g.append(j * n + i)
        gauges.append(np.asarray(g, dtype=np.int64))
    return ws, gauges
