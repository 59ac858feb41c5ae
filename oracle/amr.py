"""Two-level coupling (one fine box, refinement ratio 2) restated on the CPU.

Test infrastructure only: the checker for hydroexa_amd.amr and the coupling
entry points of libhydroexa.  FP64 numpy around the C oracle's face and step
functions.  Arrays are (ny, nx), element (i, j) at a[j, i]; a fine level is
held padded with its 2 ghost layers, (nyf+4, nxf+4), valid cell (0, 0) at
[2, 2].

What is restated, and from where:

* ghost halo of a level with physical sides: include/boundaries/BCFill.H:49-132
  with the BCRec codes of src/state/AmrMeshState_Init.cpp:52-70 (the same
  rule as the C oracle's fill_ghosts);
* fine ghost fill: AmrMeshState::FillPatch / FillPatchEta
  (src/state/AmrMeshState_AMR.cpp:667-767) through FillPatchFrom
  (include/utils/FillPatchUtils.H:117-140) = amrex::FillPatchTwoLevels with
  cell_cons_interp; eta derived as h + z per time level
  (src/state/AmrMeshState_Utils.cpp:15-50), time blend (1 - a) old + a new,
  then h = max(0, eta - z) on the fine bed (src/solvers/Roe.cpp:315-339);
* flux registers: side-selected effective fluxes D+ = F + S+ / D- = F - S-
  (src/solvers/Roe.cpp:355-392, 440-592), hydrostatic 0.5 FL + 0.5 FR
  (src/solvers/HydrostaticRoe.cpp:609-648); CrseInit COPY with -dt dA,
  FineAdd with dt_f dA_f added face by face;
* Reflux with scale 1 over volume dx dy, then average_down
  (src/state/AmrMeshState_Physics.cpp:130-141, AmrMeshState_AMR.cpp:828-844);
* dt protocol: include/solvers/BaseSolver.H:120-165 and
  src/state/AmrMeshState_Physics.cpp:13-45.

AMReX (absent from the reference tree) supplies CellConservativeLinear,
FluxRegister and average_down; their published algorithms are restated here:
cell_cons_interp is CellConservativeLinear without linear limiting -- per
component monotonized-central slopes (central 0.5 (u+ - u-), bounded by
2 |u+ - u| and 2 |u - u-|, zero at an extremum) and fine value
u_c + xoff sx + yoff sy with offsets +-1/4.  Its min/max limiter over the 3x3
coarse neighbourhood never binds for such slopes at ratio 2 (each offset term
is at most half a one-sided difference), so it is not restated.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import oracle as O

G = 2


# ---------------------------------------------------------------------------
# halo of a level with physical sides
# ---------------------------------------------------------------------------
def pad_physical(a: np.ndarray, kind: int, bc_lo: Sequence[int], bc_hi: Sequence[int]) -> np.ndarray:
    """kind: 0 scalar (eta), 1 x-momentum, 2 y-momentum, 3 foextrap (terrain).
    x sides for valid rows first, then y sides over all columns."""
    ny, nx = a.shape
    P = np.zeros((ny + 2 * G, nx + 2 * G))
    P[G:G + ny, G:G + nx] = a
    for g in (1, 2):
        refl = kind != 3 and bc_lo[0] == 0
        sgn = -1.0 if (refl and kind == 1) else 1.0
        src = g - 1 if refl else 0
        P[G:G + ny, G - g] = sgn * P[G:G + ny, G + src]
        refl = kind != 3 and bc_hi[0] == 0
        sgn = -1.0 if (refl and kind == 1) else 1.0
        src = nx - g if refl else nx - 1
        P[G:G + ny, G + nx - 1 + g] = sgn * P[G:G + ny, G + src]
    for g in (1, 2):
        refl = kind != 3 and bc_lo[1] == 0
        sgn = -1.0 if (refl and kind == 2) else 1.0
        src = g - 1 if refl else 0
        P[G - g, :] = sgn * P[G + src, :]
        refl = kind != 3 and bc_hi[1] == 0
        sgn = -1.0 if (refl and kind == 2) else 1.0
        src = ny - g if refl else ny - 1
        P[G + ny - 1 + g, :] = sgn * P[G + src, :]
    return P


def working_state(h, hu, hv, z, n, bc_lo=None, bc_hi=None) -> Dict[str, np.ndarray]:
    """Padded working state the faces read: he = max(0, eta - z) with eta and
    z carried across the sides (physical BC when bc_lo/bc_hi are given, else
    the arrays are already padded with stored ghosts)."""
    if bc_lo is not None:
        eta = pad_physical(h + z, 0, bc_lo, bc_hi)
        hu, hv = pad_physical(hu, 1, bc_lo, bc_hi), pad_physical(hv, 2, bc_lo, bc_hi)
        z, n = pad_physical(z, 3, bc_lo, bc_hi), pad_physical(n, 3, bc_lo, bc_hi)
    else:
        eta = h + z
    he = np.where(eta - z > 0.0, eta - z, 0.0)
    return {"he": he, "hu": hu, "hv": hv, "z": z, "n": n}


# ---------------------------------------------------------------------------
# effective face fluxes and the flux register
# ---------------------------------------------------------------------------
def face_flux(solver: int, W: Dict[str, np.ndarray], face: Sequence[int], dx: float,
              dy: float) -> np.ndarray:
    """Effective flux at face {dir, i, j, side} of a padded working state:
    x face (i, j) lies between cells (i-1, j) and (i, j)."""
    d, i, j, side = face
    di, dj = (1, 0) if d == 0 else (0, 1)
    nx, ny = (1.0, 0.0) if d == 0 else (0.0, 1.0)
    at = lambda ii, jj, k: W[k][jj + G, ii + G]
    if solver == O.ROE:
        L = [at(i - di, j - dj, k) for k in ("he", "hu", "hv", "z", "n")]
        R = [at(i, j, k) for k in ("he", "hu", "hv", "z", "n")]
        F, Sm, Sp = O.roe_face(L, R, 0.0, dx if d == 0 else dy, nx, ny)
        return F + Sp if side == 0 else F - Sm
    cells = [[at(i + (t - 2) * di, j + (t - 2) * dj, k) for k in ("he", "hu", "hv", "z")]
             for t in range(4)]
    FL, FR = O.hydro_face(np.asarray(cells), nx, ny)
    return 0.5 * FL + 0.5 * FR


def box_faces(box: Tuple[int, int, int, int]):
    """Coarse faces around the box (x-low, x-high, y-low, y-high) with the
    coarse side's view, and the fine faces (two per coarse face) with the
    fine side's view; the fine indices are the fine box's own."""
    ci0, cj0, ncx, ncy = box
    nxf, nyf = 2 * ncx, 2 * ncy
    cx = [(0, ci0, cj0 + m, 0) for m in range(ncy)] + [(0, ci0 + ncx, cj0 + m, 1) for m in range(ncy)]
    cy = [(1, ci0 + m, cj0, 0) for m in range(ncx)] + [(1, ci0 + m, cj0 + ncy, 1) for m in range(ncx)]
    fx = [(0, 0, 2 * m + s, 1) for m in range(ncy) for s in (0, 1)] + \
         [(0, nxf, 2 * m + s, 0) for m in range(ncy) for s in (0, 1)]
    fy = [(1, 2 * m + s, 0, 1) for m in range(ncx) for s in (0, 1)] + \
         [(1, 2 * m + s, nyf, 0) for m in range(ncx) for s in (0, 1)]
    return cx, cy, fx, fy


def crse_init(solver, W, box, dt, dx, dy) -> np.ndarray:
    """FluxRegister::CrseInit(COPY) with scale -dt dA: (2 (ncx + ncy), 3)."""
    cx, cy, _, _ = box_faces(box)
    reg = [(-dt * dy) * face_flux(solver, W, f, dx, dy) for f in cx]
    reg += [(-dt * dx) * face_flux(solver, W, f, dx, dy) for f in cy]
    return np.asarray(reg)


def fine_add(reg: np.ndarray, solver, Wf, box, dt_f, dxf, dyf) -> None:
    """FluxRegister::FineAdd with scale dt_f dA_f, one fine face at a time."""
    _, _, fx, fy = box_faces(box)
    nq_x = len(fx) // 2
    for q in range(reg.shape[0]):
        if q < nq_x:
            pair, s = fx[2 * q:2 * q + 2], dt_f * dyf
        else:
            k = q - nq_x
            pair, s = fy[2 * k:2 * k + 2], dt_f * dxf
        for f in pair:
            reg[q] = reg[q] + s * face_flux(solver, Wf, f, dxf, dyf)


def reflux(U: List[np.ndarray], reg: np.ndarray, box, dx: float, dy: float) -> None:
    """FluxRegister::Reflux, scale 1: the uncovered coarse cell beside each
    face gains (-reg)/vol on the low sides and reg/vol on the high sides."""
    ci0, cj0, ncx, ncy = box
    vol = dx * dy
    for q in range(reg.shape[0]):
        if q < ncy:
            i, j, sg = ci0 - 1, cj0 + q, -1.0
        elif q < 2 * ncy:
            i, j, sg = ci0 + ncx, cj0 + q - ncy, 1.0
        elif q < 2 * ncy + ncx:
            i, j, sg = ci0 + q - 2 * ncy, cj0 - 1, -1.0
        else:
            i, j, sg = ci0 + q - 2 * ncy - ncx, cj0 + ncy, 1.0
        for k in range(3):
            U[k][j, i] = U[k][j, i] + (sg * reg[q, k]) / vol


def average_down(Uf_valid: List[np.ndarray], U: List[np.ndarray], box) -> None:
    """amrex_avgdown: children summed j-outer, i-inner from 0, times 1/4."""
    ci0, cj0, ncx, ncy = box
    for k in range(3):
        f = Uf_valid[k]
        c = np.zeros((ncy, ncx))
        for jr in range(2):
            for ir in range(2):
                c = c + f[jr::2, ir::2]
        U[k][cj0:cj0 + ncy, ci0:ci0 + ncx] = 0.25 * c


# ---------------------------------------------------------------------------
# fine ghost fill
# ---------------------------------------------------------------------------
def mc_slope(m, c, p):
    cen = 0.5 * (p - m)
    fw, bw = p - c, c - m
    lim = np.where(fw * bw > 0.0, 2.0 * np.minimum(np.abs(fw), np.abs(bw)), 0.0)
    return np.copysign(np.minimum(np.abs(cen), lim), cen)


def _coarse_sample(C: Dict[str, np.ndarray], bc_lo, bc_hi, i, j):
    """eta, qx, qy of coarse index (i, j) that may lie one cell outside the
    level (physical rule: mirror for a reflective side with the normal
    momentum negated, edge cell for a transmissive side)."""
    ny, nx = C["h"].shape

    def src(a, n, lo, hi):
        a = np.asarray(a)
        s, flip = a.copy(), np.zeros(a.shape, bool)
        low, high = a < 0, a >= n
        s = np.where(low, np.where(lo == 0, -1 - a, 0), s)
        s = np.where(high, np.where(hi == 0, 2 * n - 1 - a, n - 1), s)
        flip = (low & (lo == 0)) | (high & (hi == 0))
        return s, flip

    si, fx = src(i, nx, bc_lo[0], bc_hi[0])
    sj, fy = src(j, ny, bc_lo[1], bc_hi[1])
    eta = C["h"][sj, si] + C["z"][sj, si]
    qx = np.where(fx, -C["hu"][sj, si], C["hu"][sj, si])
    qy = np.where(fy, -C["hv"][sj, si], C["hv"][sj, si])
    return [eta, qx, qy]


def fill_fine(C_old, C_new, alpha: float, box, bc_lo, bc_hi, zf: np.ndarray, Uf: np.ndarray,
              whole: bool = False) -> None:
    """Fill the fine ghost ring (or the whole padded box) of Uf (3, nyf+4,
    nxf+4) in place from the coarse level at time blend alpha."""
    ci0, cj0, ncx, ncy = box
    nyf, nxf = 2 * ncy, 2 * ncx
    fj, fi = np.meshgrid(np.arange(-G, nyf + G), np.arange(-G, nxf + G), indexing="ij")
    sel = np.ones(fi.shape, bool) if whole else ~((fi >= 0) & (fi < nxf) & (fj >= 0) & (fj < nyf))
    fi, fj = fi[sel], fj[sel]
    pi, pj = ci0 + np.floor_divide(fi, 2), cj0 + np.floor_divide(fj, 2)
    ox = np.where(fi % 2 == 0, -0.25, 0.25)
    oy = np.where(fj % 2 == 0, -0.25, 0.25)

    def q(ii, jj):
        a = _coarse_sample(C_old, bc_lo, bc_hi, ii, jj)
        if alpha == 0.0:
            return a
        b = _coarse_sample(C_new, bc_lo, bc_hi, ii, jj)
        if alpha == 1.0:
            return b
        return [(1.0 - alpha) * x + alpha * y for x, y in zip(a, b)]

    qc, qw, qe = q(pi, pj), q(pi - 1, pj), q(pi + 1, pj)
    qs, qn = q(pi, pj - 1), q(pi, pj + 1)
    v = [qc[k] + mc_slope(qw[k], qc[k], qe[k]) * ox + mc_slope(qs[k], qc[k], qn[k]) * oy
         for k in range(3)]
    x = v[0] - zf[fj + G, fi + G]
    Uf[0][fj + G, fi + G] = np.where(0.0 < x, x, 0.0)
    Uf[1][fj + G, fi + G] = v[1]
    Uf[2][fj + G, fi + G] = v[2]


# ---------------------------------------------------------------------------
# the coupled step
# ---------------------------------------------------------------------------
def fine_terrain(z: np.ndarray, n: np.ndarray, box) -> Tuple[np.ndarray, np.ndarray]:
    """Piecewise-constant fine terrain with 2 ghost layers (pc_interp of the
    foextrap-filled coarse terrain)."""
    ci0, cj0, ncx, ncy = box
    ny, nx = z.shape
    fi = np.arange(-G, 2 * ncx + G)
    fj = np.arange(-G, 2 * ncy + G)
    pi = np.clip(ci0 + np.floor_divide(fi, 2), 0, nx - 1)
    pj = np.clip(cj0 + np.floor_divide(fj, 2), 0, ny - 1)
    return z[np.ix_(pj, pi)].copy(), n[np.ix_(pj, pi)].copy()


def level_dts(solver, C, Ff, dx, dy) -> Tuple[float, float]:
    """CFL estimates (before the cfl factor) of the coarse level and the fine
    valid cells."""
    d0, _ = O.level_dt(C["h"], C["hu"], C["hv"], C["z"], dx, dy)
    v = (slice(G, -G), slice(G, -G))
    d1, _ = O.level_dt(Ff["U"][0][v], Ff["U"][1][v], Ff["U"][2][v], Ff["z"][v], dx / 2, dy / 2)
    return d0, d1


def dt_rule(d0: float, d1: float, subcycle: bool) -> Tuple[float, float]:
    """BaseSolver::compute_dt for two levels at ratio 2; lockstep shares the
    finer dt."""
    dt0 = min(d0, d1 * 2.0)
    dt1 = dt0 / 2.0
    return (dt0, dt1) if subcycle else (dt1, dt1)


class TwoLevelOracle:
    def __init__(self, solver, h, hu, hv, z, n, dx, dy, bc_lo, bc_hi, box, zf=None, nf=None,
                 subcycle=True):
        self.solver, self.dx, self.dy = solver, dx, dy
        self.bc_lo, self.bc_hi, self.box, self.subcycle = tuple(bc_lo), tuple(bc_hi), box, subcycle
        f64 = lambda a: np.array(a, dtype=np.float64)
        self.C = {"h": f64(h), "hu": f64(hu), "hv": f64(hv), "z": f64(z), "n": f64(n)}
        if zf is None:
            zf, nf = fine_terrain(self.C["z"], self.C["n"], box)
        ncx, ncy = box[2], box[3]
        self.F = {"U": np.zeros((3, 2 * ncy + 2 * G, 2 * ncx + 2 * G)), "z": f64(zf), "n": f64(nf)}
        fill_fine(self.C, self.C, 0.0, box, self.bc_lo, self.bc_hi, self.F["z"], self.F["U"],
                  whole=True)

    def fine_valid(self) -> np.ndarray:
        return self.F["U"][:, G:-G, G:-G]

    def step(self, dt0: float, dt1: float) -> None:
        s, C, F, box = self.solver, self.C, self.F, self.box
        dx, dy = self.dx, self.dy
        W = working_state(C["h"], C["hu"], C["hv"], C["z"], C["n"], self.bc_lo, self.bc_hi)
        reg = crse_init(s, W, box, dt0, dx, dy)
        old = {k: v.copy() for k, v in C.items()}
        nh, nhu, nhv = O.step(s, C["h"], C["hu"], C["hv"], C["z"], C["n"], dt0, dx, dy,
                              self.bc_lo, self.bc_hi)
        C["h"], C["hu"], C["hv"] = nh, nhu, nhv
        nsub = 2 if self.subcycle else 1
        for k in range(nsub):
            fill_fine(old, C, k / 2.0, box, self.bc_lo, self.bc_hi, F["z"], F["U"])
            U = F["U"]
            Wf = working_state(U[0], U[1], U[2], F["z"], F["n"])
            fine_add(reg, s, Wf, box, dt1, dx / 2, dy / 2)
            out = O.step_padded(s, U[0], U[1], U[2], F["z"], F["n"], dt1, dx / 2, dy / 2)
            for c in range(3):
                U[c][G:-G, G:-G] = out[c]
        Uc = [C["h"], C["hu"], C["hv"]]
        reflux(Uc, reg, box, dx, dy)
        average_down([F["U"][c][G:-G, G:-G] for c in range(3)], Uc, box)


# ---------------------------------------------------------------------------
# multi-box hierarchy (dense levels with 2 cells of padding, fixed boxes)
# ---------------------------------------------------------------------------
def _axis_map(idx: np.ndarray, n: int, phys_lo: bool, phys_hi: bool, blo: int, bhi: int):
    """Source cell, edge cell and normal flip of level indices idx along one
    axis: a side on the domain boundary maps with the physical rule, any
    other index is read as stored."""
    src, edge, flip = idx.copy(), idx.copy(), np.zeros(idx.shape, bool)
    if phys_lo:
        low = idx < 0
        edge[low] = 0
        if blo == 0:
            src[low] = -1 - idx[low]
            flip[low] = True
        else:
            src[low] = 0
    if phys_hi:
        high = idx >= n
        edge[high] = n - 1
        if bhi == 0:
            src[high] = 2 * n - 1 - idx[high]
            flip[high] = True
        else:
            src[high] = n - 1
    return src, edge, flip


class DenseLevelOracle:
    def __init__(self, nx, ny, dx, dy, bc_lo, bc_hi):
        self.nx, self.ny, self.dx, self.dy = nx, ny, dx, dy
        self.bc_lo, self.bc_hi = tuple(bc_lo), tuple(bc_hi)
        shp = (ny + 2 * G, nx + 2 * G)
        self.U = np.zeros((3,) + shp)
        self.Uo = np.zeros((3,) + shp)
        self.T = np.zeros((2,) + shp)
        self.mask = np.zeros(shp, bool)
        self.boxes = []
        self.t_new, self.t_old = 0.0, -1.0e200

    def set_boxes(self, boxes):
        self.boxes = [tuple(b) for b in boxes]
        self.mask[:] = False
        for (i0, j0, w, h) in self.boxes:
            self.mask[G + j0:G + j0 + h, G + i0:G + i0 + w] = True

    def interior(self, a):
        return a[..., G:G + self.ny, G:G + self.nx]

    def full_working(self, U):
        """Working state of the whole level with the physical halo."""
        I = lambda a: self.interior(a)
        return working_state(I(U[0]), I(U[1]), I(U[2]), I(self.T[0]), I(self.T[1]),
                             self.bc_lo, self.bc_hi)

    def gather_box(self, U, box):
        """Padded (h, hu, hv, z, n, eta) of a box with its 2-cell ring, as the
        advance's loader sees them."""
        i0, j0, w, h = box
        ii = np.arange(i0 - G, i0 + w + G)
        jj = np.arange(j0 - G, j0 + h + G)
        si, ei, fi = _axis_map(ii, self.nx, i0 == 0, i0 + w == self.nx, self.bc_lo[0], self.bc_hi[0])
        sj, ej, fj = _axis_map(jj, self.ny, j0 == 0, j0 + h == self.ny, self.bc_lo[1], self.bc_hi[1])
        S = np.ix_(sj + G, si + G)
        E = np.ix_(ej + G, ei + G)
        hs = U[0][S]
        eta = hs + self.T[0][S]
        hu = np.where(fi[None, :], -U[1][S], U[1][S])
        hv = np.where(fj[:, None], -U[2][S], U[2][S])
        return hs, hu, hv, self.T[0][E], self.T[1][E], eta


class AmrOracle:
    """Restatement of hydroexa_amd.amr.AmrHierarchy on fixed boxes: the same
    recursion, ghost fills (mask-skipped), registers, reflux order and
    average_down, with the C oracle's ghost-fed step per box."""

    def __init__(self, solver, h, hu, hv, z, n, dx, dy, bc_lo, bc_hi, boxes, terrain=None,
                 subcycle=True, do_reflux=True, regrid_int=0, max_level=None, events=None):
        """boxes: level -> box list at time 0.  With regrid_int > 0 the grids
        change where the hierarchy's step regrids (AmrMeshState's step,
        src/state/AmrMeshState_Physics.cpp:50-147, with AMReX's regrid_int
        cadence): level l remakes the levels above it at the start of its
        step when its step count is a positive multiple of regrid_int not yet
        regridded.  The new boxes are not derived here: events is a list,
        consumed in order, of (l, time, {level: boxes}) recorded from the run
        under test; the oracle checks that each event happens where its
        cadence says and remakes the levels as the hierarchy does."""
        self.solver, self.subcycle, self.do_reflux = solver, subcycle, do_reflux
        self.regrid_int = regrid_int
        self.max_level = max(boxes) if max_level is None else max_level
        self.events = events if events is not None else []
        self.bc_lo, self.bc_hi = tuple(bc_lo), tuple(bc_hi)
        ny, nx = np.asarray(h).shape
        lv = DenseLevelOracle(nx, ny, dx, dy, bc_lo, bc_hi)
        pad = lambda a: np.pad(np.asarray(a, np.float64), G, mode="edge")
        lv.U[:] = [pad(h), pad(hu), pad(hv)]
        lv.T[:] = [pad(z), pad(n)]
        lv.set_boxes([(0, 0, nx, ny)])
        self.levels = [lv]
        for lev in sorted(boxes):
            c = self.levels[-1]
            f = DenseLevelOracle(2 * c.nx, 2 * c.ny, c.dx / 2, c.dy / 2, bc_lo, bc_hi)
            if terrain is not None and lev in terrain:
                zf, nf = terrain[lev]
                f.T[:] = [pad(zf), pad(nf)]
            else:
                ref = lambda a: np.repeat(np.repeat(a, 2, 0), 2, 1)
                f.T[:] = [pad(ref(c.interior(c.T[0]))), pad(ref(c.interior(c.T[1])))]
            f.set_boxes(boxes[lev])
            self.levels.append(f)
            self._fill(lev, 0.0, whole=True, skip=None, crse=(c.U, c.U, 0.0))
        self.dt = [0.0] * (self.max_level + 1)
        self.time = 0.0
        self.regs = [[] for _ in range(self.max_level + 1)]
        self.istep = [0] * (self.max_level + 1)
        self.last_regrid = [0] * (self.max_level + 1)

    @property
    def finest(self):
        return len(self.levels) - 1

    def _crse_dict(self, lv, U):
        I = lambda a: lv.interior(a)
        return {"h": I(U[0]), "hu": I(U[1]), "hv": I(U[2]), "z": I(lv.T[0]), "n": I(lv.T[1])}

    def _fill(self, lev, t, whole, skip, crse=None):
        c, f = self.levels[lev - 1], self.levels[lev]
        if crse is None:
            if abs(t - c.t_new) <= 1e-12 * max(abs(t), abs(c.t_new), 1e-300):
                co, cn, alpha = c.U, c.U, 0.0
            elif abs(t - c.t_old) <= 1e-12 * max(abs(t), abs(c.t_old), 1e-300):
                co, cn, alpha = c.Uo, c.Uo, 0.0
            else:
                co, cn, alpha = c.Uo, c.U, (t - c.t_old) / (c.t_new - c.t_old)
        else:
            co, cn, alpha = crse
        Co, Cn = self._crse_dict(c, co), self._crse_dict(c, cn)
        for (i0, j0, w, h) in f.boxes:
            sub = f.U[:, j0:j0 + h + 2 * G, i0:i0 + w + 2 * G].copy()
            zf = f.T[0][j0:j0 + h + 2 * G, i0:i0 + w + 2 * G]
            before = sub.copy()
            fill_fine(Co, Cn, alpha, (i0 // 2, j0 // 2, w // 2, h // 2), self.bc_lo, self.bc_hi,
                      zf, sub, whole=whole)
            keep = (f.mask if skip is None else skip)[j0:j0 + h + 2 * G, i0:i0 + w + 2 * G]
            if skip is None and whole:
                keep = np.zeros_like(keep)
            sub[:, keep] = before[:, keep]
            f.U[:, j0:j0 + h + 2 * G, i0:i0 + w + 2 * G] = sub

    def _face_list(self, box):
        i0, j0, w, h = box
        cx, cy, fx, fy = box_faces((i0 // 2, j0 // 2, w // 2, h // 2))
        return cx + cy, fx + fy

    def _advance(self, lev, t, dt):
        lv = self.levels[lev]
        if lev > 0:
            self._fill(lev, t, whole=False, skip=None)
        if self.do_reflux and lev < self.finest:
            fl = self.levels[lev + 1]
            W = lv.full_working(lv.U)
            self.regs[lev + 1] = []
            for b in fl.boxes:
                cf, _ = self._face_list(b)
                reg = np.asarray([(-dt * (lv.dy if f[0] == 0 else lv.dx)) *
                                  face_flux(self.solver, W, f, lv.dx, lv.dy) for f in cf])
                self.regs[lev + 1].append(reg)
        if self.do_reflux and lev > 0:
            for b, reg in zip(lv.boxes, self.regs[lev]):
                hs, hu, hv, z, n, eta = lv.gather_box(lv.U, b)
                he = np.where(eta - z > 0.0, eta - z, 0.0)
                Wb = {"he": he, "hu": hu, "hv": hv, "z": z, "n": n}
                _, ff = self._face_list(b)
                for q in range(reg.shape[0]):
                    for f in ff[2 * q:2 * q + 2]:
                        s = dt * (lv.dy if f[0] == 0 else lv.dx)
                        reg[q] = reg[q] + s * face_flux(self.solver, Wb, f, lv.dx, lv.dy)
        lv.U, lv.Uo = lv.Uo, lv.U
        lv.U[:] = lv.Uo      # the invalid cells keep the last values, as on the device
        for b in lv.boxes:
            i0, j0, w, h = b
            hs, hu, hv, z, n, eta = lv.gather_box(lv.Uo, b)
            out = O.step_padded(self.solver, hs, hu, hv, z, n, dt, lv.dx, lv.dy, eta=eta)
            for k in range(3):
                lv.U[k, G + j0:G + j0 + h, G + i0:G + i0 + w] = out[k]

    def _reflux(self, lev):
        lv, fl = self.levels[lev], self.levels[lev + 1]
        I = [lv.interior(lv.U[k]) for k in range(3)]
        for orient in range(4):
            for b, reg in zip(fl.boxes, self.regs[lev + 1]):
                ci0, cj0, ncx, ncy = b[0] // 2, b[1] // 2, b[2] // 2, b[3] // 2
                for q in range(reg.shape[0]):
                    if q < ncy:
                        i, j, sg, side = ci0 - 1, cj0 + q, -1.0, 0
                    elif q < 2 * ncy:
                        i, j, sg, side = ci0 + ncx, cj0 + q - ncy, 1.0, 2
                    elif q < 2 * ncy + ncx:
                        i, j, sg, side = ci0 + q - 2 * ncy, cj0 - 1, -1.0, 1
                    else:
                        i, j, sg, side = ci0 + q - 2 * ncy - ncx, cj0 + ncy, 1.0, 3
                    if side != orient or i < 0 or j < 0 or i >= lv.nx or j >= lv.ny:
                        continue
                    for k in range(3):
                        I[k][j, i] = I[k][j, i] + (sg * reg[q, k]) / (lv.dx * lv.dy)

    def _average_down(self, lev):
        lv, fl = self.levels[lev], self.levels[lev + 1]
        for (i0, j0, w, h) in fl.boxes:
            src = [fl.U[k, G + j0:G + j0 + h, G + i0:G + i0 + w] for k in range(3)]
            dst = [lv.interior(lv.U[k]) for k in range(3)]
            average_down(src, dst, (i0 // 2, j0 // 2, w // 2, h // 2))

    def level_dt(self, lev):
        lv = self.levels[lev]
        est = np.inf
        for (i0, j0, w, h) in lv.boxes:
            S = (slice(G + j0, G + j0 + h), slice(G + i0, G + i0 + w))
            d, _ = O.level_dt(lv.U[0][S], lv.U[1][S], lv.U[2][S], lv.T[0][S], lv.dx, lv.dy)
            est = min(est, d)
        return est

    def compute_dt(self, cfl):
        tmp = [cfl * self.level_dt(l) for l in range(self.finest + 1)]
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
        return list(self.dt[:self.finest + 1])

    def _remake(self, lbase, time, boxes):
        """New grids for the levels above lbase (AmrMesh::MakeNewGrids with the
        boxes given): an existing level keeps its data where the new boxes
        cover cells it held and takes the conservative interpolation of the
        level below elsewhere (RemakeLevel), a new level takes piecewise-
        constant terrain and the interpolation everywhere
        (MakeNewLevelFromCoarse); a level with no boxes is removed with the
        levels above it."""
        for k in range(lbase, self.max_level):
            if k > self.finest:
                break
            new = [tuple(b) for b in boxes.get(k + 1, [])]
            if not new:
                del self.levels[k + 1:]
                break
            c = self.levels[k]
            if k + 1 > self.finest:
                f = DenseLevelOracle(2 * c.nx, 2 * c.ny, c.dx / 2, c.dy / 2, self.bc_lo, self.bc_hi)
                pad = lambda a: np.pad(a, G, mode="edge")
                ref = lambda a: np.repeat(np.repeat(a, 2, 0), 2, 1)
                f.T[:] = [pad(ref(c.interior(c.T[0]))), pad(ref(c.interior(c.T[1])))]
                f.t_new, f.t_old = time, time - 1.0e200
                self.levels.append(f)
                self.dt[k + 1] = self.dt[k] / (2.0 if self.subcycle else 1.0)
                old = None
            else:
                f = self.levels[k + 1]
                if new == f.boxes:
                    continue                   # the same grids: nothing to remake
                old = f.mask.copy()
            f.set_boxes(new)
            if old is None:
                self._fill(k + 1, time, whole=True, skip=None, crse=(c.U, c.U, 0.0))
            else:
                self._fill(k + 1, time, whole=True, skip=old, crse=(c.U, c.U, 0.0))

    def _time_step(self, lev, time):
        if (self.regrid_int > 0 and lev < self.max_level and self.istep[lev] > self.last_regrid[lev]
                and self.istep[lev] % self.regrid_int == 0):
            if not self.events:
                raise AssertionError(f"level {lev} regrids at t = {time!r} but no event is left")
            el, et, eb = self.events.pop(0)
            if el != lev or et != time:
                raise AssertionError(f"regrid of level {lev} at t = {time!r}; next event {el} at {et!r}")
            old_finest = self.finest
            self._remake(lev, time, eb)
            for k in range(lev, self.finest + 1):
                self.last_regrid[k] = self.istep[k]
            for k in range(old_finest + 1, self.finest + 1):
                self.dt[k] = self.dt[k - 1] / (2.0 if self.subcycle else 1.0)
        lv = self.levels[lev]
        lv.t_old = lv.t_new
        lv.t_new += self.dt[lev]
        self._advance(lev, lv.t_old, self.dt[lev])
        self.istep[lev] += 1
        if lev < self.finest:
            nsub = 2 if self.subcycle else 1
            for i in range(1, nsub + 1):
                self._time_step(lev + 1, time + (i - 1) * self.dt[lev + 1])
            if self.do_reflux:
                self._reflux(lev)
            self._average_down(lev)

    def step(self, dts=None, cfl=None):
        if dts is None:
            self.compute_dt(cfl)
        else:
            for l, d in enumerate(dts):
                self.dt[l] = float(d)
        self._time_step(0, self.time)
        self.time = self.levels[0].t_new
