"""Host box layer of the AMR hierarchy: new fine grids from cell tags.

The reference delegates grid generation to AMReX (AmrMesh::MakeNewGrids,
called from AmrMeshState's regrid, src/state/AmrMeshState_Physics.cpp:50-105,
and from the initial MakeNewLevelFromScratch sequence).  AMReX is absent from
the reference tree; this module restates its published algorithm:

1. tags of level k (valid cells only) grown by n_error_buf cells;
2. coarsened by bf = blocking_factor / ref_ratio (a block is tagged when any
   of its cells is), so that the new boxes align to blocking_factor at level
   k + 1;
3. tags outside the proper-nesting domain dropped: blocks whose cells all lie
   in level k's valid region shrunk by ``nest`` cells (the domain boundary
   does not shrink it);
4. Berger-Rigoutsos clustering (ClusterList::chop with grid_eff: a cluster
   below the efficiency is cut at the centre-most signature hole, else at
   the steepest Laplacian sign change when its jump exceeds 2, else bisected;
   the direction with the best cut kind and the most balanced cut wins);
5. clusters intersected with the nesting domain, refined back to level k,
   then by ref_ratio, and chopped to max_grid_size (BoxList::maxSize:
   pieces of nearly equal size in the largest power-of-2 granule the box
   and the limit share).

Boxes are (i0, j0, nx, ny) in the index space of their level.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

Box = Tuple[int, int, int, int]

MINOFF = 2
CUT_THRESH = 2
HOLE, STEEP, BISECT, INVALID = 0, 1, 2, 3


def grow_tags(tags: np.ndarray, nbuf: int) -> np.ndarray:
    """Each tag marks its (2 nbuf + 1)^2 neighbourhood (clipped to the level)."""
    if nbuf <= 0:
        return tags.copy()
    ny, nx = tags.shape
    out = np.zeros_like(tags)
    for dj in range(-nbuf, nbuf + 1):
        for di in range(-nbuf, nbuf + 1):
            out[max(0, dj):ny + min(0, dj), max(0, di):nx + min(0, di)] |= \
                tags[max(0, -dj):ny - max(0, dj), max(0, -di):nx - max(0, di)]
    return out


def coarsen_tags(tags: np.ndarray, b: int) -> np.ndarray:
    """Block b x b: tagged when any cell is (the level extents are multiples of b)."""
    ny, nx = tags.shape
    return tags.reshape(ny // b, b, nx // b, b).any(axis=(1, 3))


def erode(valid: np.ndarray, m: int) -> np.ndarray:
    """Cells whose (2m+1)^2 neighbourhood lies in valid; outside the level
    counts as valid (the physical boundary does not shrink the region)."""
    if m <= 0:
        return valid.copy()
    ny, nx = valid.shape
    pad = np.ones((ny + 2 * m, nx + 2 * m), bool)
    pad[m:m + ny, m:m + nx] = valid
    out = np.ones_like(valid)
    for dj in range(2 * m + 1):
        for di in range(2 * m + 1):
            out &= pad[dj:dj + ny, di:di + nx]
    return out


def find_cut(hist: np.ndarray) -> Tuple[int, int]:
    """AMReX FindCut on one signature: (offset of the cut, status)."""
    n = len(hist)
    if n <= 1:
        return 0, INVALID
    mid = n // 2
    cut, status = -1, INVALID
    for i in range(n):
        if hist[i] == 0:
            status = HOLE
            if abs(cut - mid) > abs(i - mid):
                cut = i
                if i > mid:
                    break
    if status == HOLE:
        return cut, HOLE
    d = np.zeros(n, dtype=np.int64)
    d[1:n - 1] = hist[2:] - 2 * hist[1:n - 1] + hist[:n - 2]
    locmax = -1
    for i in range(MINOFF, n - MINOFF):
        prev, cur = int(d[i - 1]), int(d[i])
        dif = abs(prev - cur)
        if prev * cur < 0 and dif >= locmax:
            if dif > locmax:
                status, cut, locmax = STEEP, i, dif
            elif abs(i - mid) < abs(cut - mid):
                cut = i
    if locmax <= CUT_THRESH:
        return mid, BISECT
    return cut, status


class _Cluster:
    def __init__(self, pts: np.ndarray):
        self.pts = pts                                   # (n, 2) rows of (i, j)
        self.lo = pts.min(axis=0)
        self.hi = pts.max(axis=0)

    def eff(self) -> float:
        ext = self.hi - self.lo + 1
        return len(self.pts) / float(ext[0] * ext[1])

    def chop(self) -> "_Cluster":
        ext = self.hi - self.lo + 1
        best = None
        cuts = []
        for d in range(2):
            hist = np.bincount(self.pts[:, d] - self.lo[d], minlength=ext[d])
            off, st = find_cut(hist)
            cuts.append((self.lo[d] + off, st))
        mincut = min(st for _, st in cuts)
        minlen = -1
        for d in range(2):
            c, st = cuts[d]
            if st == mincut:
                ml = min(c - self.lo[d], self.hi[d] - c)
                if ml >= minlen:
                    best, minlen = d, ml
        d = best
        c = cuts[d][0]
        low = self.pts[:, d] < c
        if not low.any() or low.all():           # degenerate: bisect the box
            c = self.lo[d] + (ext[d] + 1) // 2
            low = self.pts[:, d] < c
        hi_pts = self.pts[~low]
        self.__init__(self.pts[low])
        return _Cluster(hi_pts)


def cluster(tags: np.ndarray, grid_eff: float) -> List[Box]:
    """Berger-Rigoutsos boxes (i0, j0, nx, ny) covering every tag."""
    jj, ii = np.nonzero(tags)
    if len(ii) == 0:
        return []
    work = [_Cluster(np.stack([ii, jj], axis=1))]
    k = 0
    while k < len(work):
        c = work[k]
        if c.eff() < grid_eff and len(c.pts) > 1:
            work.append(c.chop())
        else:
            k += 1
    return [(int(c.lo[0]), int(c.lo[1]), int(c.hi[0] - c.lo[0] + 1), int(c.hi[1] - c.lo[1] + 1))
            for c in work]


def _row_runs_all(mask: np.ndarray) -> dict:
    """Row j -> its runs of true cells [(start, end)), ascending, for every
    row of mask at once."""
    ny, nx = mask.shape
    m = np.zeros((ny, nx + 2), np.int8)
    m[:, 1:-1] = mask
    d = np.diff(m, axis=1)
    sj, si = np.nonzero(d == 1)
    _, ei = np.nonzero(d == -1)
    rows = {}
    for j, a, b in zip(sj.tolist(), si.tolist(), ei.tolist()):
        rows.setdefault(j, []).append((a, b))
    return rows


def _rects(mask: np.ndarray) -> List[Box]:
    """A partition of mask's true cells into rectangles (row runs merged
    down while they repeat)."""
    ny, nx = mask.shape
    out = []
    open_runs = {}
    all_runs = _row_runs_all(mask) if ny and nx else {}
    for j in range(ny + 1):
        runs = set(all_runs.get(j, ())) if j < ny else set()
        nxt = {}
        for r, j0 in open_runs.items():
            if r in runs:
                nxt[r] = j0
            else:
                out.append((r[0], j0, r[1] - r[0], j - j0))
        for r in runs:
            if r not in nxt:
                nxt[r] = j
        open_runs = nxt
    return out


def intersect_nesting(boxes: List[Box], allowed: np.ndarray, tags: np.ndarray) -> List[Box]:
    """ClusterList::intersect with the nesting domain: boxes reaching outside
    the allowed blocks are replaced by the allowed rectangles inside them that
    hold tags."""
    out = []
    for (i0, j0, nx, ny) in boxes:
        sub = allowed[j0:j0 + ny, i0:i0 + nx]
        if sub.all():
            out.append((i0, j0, nx, ny))
            continue
        for (a, b, w, h) in _rects(sub):
            if tags[j0 + b:j0 + b + h, i0 + a:i0 + a + w].any():
                out.append((i0 + a, j0 + b, w, h))
    return out


def max_size(boxes: List[Box], chunk: int) -> List[Box]:
    """BoxList::maxSize: no side longer than chunk; pieces of nearly equal
    size in units of the largest power of 2 dividing both the side and
    chunk, chopped from the high end."""
    work = list(boxes)
    out = []
    while work:
        b = list(work.pop(0))
        for d in range(2):
            n = b[2 + d]
            if n <= chunk:
                continue
            ratio, bs, nlen = 1, chunk, n
            while bs % 2 == 0 and nlen % 2 == 0:
                ratio, bs, nlen = ratio * 2, bs // 2, nlen // 2
            numblk = nlen // bs + (1 if nlen % bs else 0)
            sz, ex = nlen // numblk, nlen % numblk
            hi = b[d] + b[2 + d]                 # exclusive end
            for k in range(numblk - 1):
                ksize = (sz + 1 if k < ex else sz) * ratio
                piece = list(b)
                piece[d] = hi - ksize
                piece[2 + d] = ksize
                work.append(tuple(piece))
                hi -= ksize
            b[2 + d] = hi - b[d]
        out.append(tuple(b))
    return out


def new_grids(tags: np.ndarray, valid: np.ndarray, ratio: int, blocking_factor: int,
              max_grid_size: int, n_error_buf: int, grid_eff: float, nest: int) -> List[Box]:
    """Boxes of level k + 1 (its index space) from level k's tags and valid
    mask, both (ny_k, nx_k) booleans."""
    t = grow_tags(tags & valid, n_error_buf) & valid
    b = max(blocking_factor // ratio, 1)
    ny, nx = t.shape
    if nx % b or ny % b:
        raise ValueError(f"level extents {nx}x{ny} are not multiples of blocking_factor/ratio = {b}")
    vb_all = valid.reshape(ny // b, b, nx // b, b).all(axis=(1, 3))
    if 1 <= nest <= b and (vb_all == coarsen_tags(valid, b)).all():
        allowed = erode(vb_all, 1)
    else:
        allowed = erode(valid, nest).reshape(ny // b, b, nx // b, b).all(axis=(1, 3))
    return grids_from_blocks(coarsen_tags(t, b), allowed, b, ratio, grid_eff, max_grid_size)


def allowed_blocks(boxes: Sequence[Box], ny: int, nx: int, b: int, nest: int) -> np.ndarray:
    """Nesting domain on the block grid of a level whose valid region is the
    union of boxes aligned to b: a block is allowed when it and its eight
    neighbours are valid (outside the level counts as valid); requires
    1 <= nest <= b."""
    if not 1 <= nest <= b:
        raise ValueError("nest must lie in [1, b] for the block-grid nesting test")
    vb = np.zeros((ny // b, nx // b), bool)
    for (i0, j0, w, h) in boxes:
        if i0 % b or j0 % b or w % b or h % b:
            raise ValueError("boxes are not aligned to the block size")
        vb[j0 // b:(j0 + h) // b, i0 // b:(i0 + w) // b] = True
    return erode(vb, 1)


def _order(boxes: List[Box]) -> List[Box]:
    return sorted(boxes, key=lambda x: (x[1], x[0], x[2], x[3]))


def grids_from_blocks_reference(tb: np.ndarray, allowed: np.ndarray, b: int, ratio: int,
                                grid_eff: float, max_grid_size: int) -> List[Box]:
    """Steps 3-5 of the module docstring from tags already grown and blocked,
    in Python (the restatement the native hx_cluster_blocks is checked
    against); boxes sorted by (j0, i0, w, h)."""
    tb = tb & allowed
    boxes = intersect_nesting(cluster(tb, grid_eff), allowed, tb)
    fine = [((i0 * b) * ratio, (j0 * b) * ratio, (w * b) * ratio, (h * b) * ratio)
            for (i0, j0, w, h) in boxes]
    return _order(max_size(fine, max_grid_size))


def grids_from_blocks(tb: np.ndarray, allowed: np.ndarray, b: int, ratio: int, grid_eff: float,
                      max_grid_size: int) -> List[Box]:
    """Steps 3-5 of the module docstring from tags already grown and blocked
    (nby x nbx booleans), by the library's hx_cluster_blocks; boxes sorted by
    (j0, i0, w, h), as grids_from_blocks_reference gives them."""
    import ctypes as C
    from . import _lib as L
    t = np.ascontiguousarray(tb, np.uint8)
    a = np.ascontiguousarray(allowed, np.uint8)
    if t.shape != a.shape or t.ndim != 2:
        raise ValueError("tags and allowed must be equal 2-D block arrays")
    n = C.c_int32()
    cap = 256
    while True:
        out = np.zeros(4 * cap, np.int32)
        rc = L.load().hx_cluster_blocks(t.ctypes.data, a.ctypes.data, t.shape[0], t.shape[1], int(b),
                                        int(ratio), float(grid_eff), int(max_grid_size),
                                        out.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(n))
        if rc == L.HX_E_BADARG and n.value > cap:
            cap = n.value
            continue
        L.check(rc, "hx_cluster_blocks")
        return [tuple(r) for r in out[:4 * n.value].reshape(-1, 4).tolist()]


def covers(boxes: Sequence[Box], ny: int, nx: int) -> np.ndarray:
    m = np.zeros((ny, nx), bool)
    for (i0, j0, w, h) in boxes:
        m[j0:j0 + h, i0:i0 + w] = True
    return m
