"""Native AMReX plotfiles of a run, as HydroEXA writes them.

``write_native_plotfile`` mirrors WriteNativePlotfile
(src/io/NativePlotFile.cpp:19-97): directory ``prefix`` + the iteration in
5 digits, components U (h, hu, hv) then terrain (z, roughness), variable
names h_fluid, hu_momentum, hv_momentum, z_bathymetry, extra_comp_4, every
level's step set to the iteration.  The bytes are written by
``hxio_write_plotfile`` (libhydroexa_io.so) in the layout of
amrex::WriteMultiLevelPlotfile.  ``read_native_plotfile`` parses that layout
back (Header, Level_<l>/Cell_H, FAB data) for checks and analysis.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Dict, List, Optional, Sequence

import numpy as np

from .h5input import _check, load_io

VARNAMES = ("h_fluid", "hu_momentum", "hv_momentum", "z_bathymetry")


class HxioPlotLevel(C.Structure):
    _fields_ = [("nboxes", C.c_int32), ("domain", C.c_int32 * 4), ("step", C.c_int32),
                ("boxes", C.POINTER(C.c_int32)), ("data", C.POINTER(C.POINTER(C.c_double))),
                ("dx", C.c_double), ("dy", C.c_double)]


def _bind():
    lib = load_io()
    if not getattr(lib, "_plot_bound", False):
        lib.hxio_write_plotfile.argtypes = [C.c_char_p, C.c_int, C.POINTER(HxioPlotLevel), C.c_int,
                                            C.POINTER(C.c_char_p), C.c_double,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_int32)]
        lib._plot_bound = True
    return lib


def plotfile_name(prefix: str, iteration: int) -> str:
    """amrex::Concatenate(prefix, iteration, 5)."""
    return f"{prefix}{iteration:05d}"


def write_native_plotfile(prefix: str, iteration: int, time: float, levels: Sequence,
                          prob_lo: Sequence[float] = (0.0, 0.0),
                          ref_ratio: Sequence[int] = (), boxes: Optional[Sequence] = None) -> str:
    """Write a plotfile of ``levels`` (hydroexa_amd.Level objects, one box
    covering each level unless ``boxes`` gives per level a list of
    {ilo, jlo, ihi, jhi} boxes that partition it).  Returns the directory."""
    if not levels:
        raise ValueError("no levels to write")
    name = plotfile_name(prefix, iteration)
    keep = []
    arr = (HxioPlotLevel * len(levels))()
    ncomp = 5
    for l, lv in enumerate(levels):
        full = np.concatenate([lv.U_new.double().cpu().numpy(),
                               lv.terrain.double().cpu().numpy()])      # (5, ny, nx)
        bl = boxes[l] if boxes is not None else [(0, 0, lv.nx - 1, lv.ny - 1)]
        bx = np.ascontiguousarray(np.asarray(bl, dtype=np.int32).reshape(-1, 4))
        datas = []
        for ilo, jlo, ihi, jhi in bx:
            datas.append(np.ascontiguousarray(full[:, jlo:jhi + 1, ilo:ihi + 1]))
        ptrs = (C.POINTER(C.c_double) * len(datas))(*[d.ctypes.data_as(C.POINTER(C.c_double))
                                                       for d in datas])
        keep += [bx, datas, ptrs]
        L = arr[l]
        L.nboxes = len(bx)
        L.domain[:] = [0, 0, lv.nx - 1, lv.ny - 1]
        L.step = iteration
        L.boxes = bx.ctypes.data_as(C.POINTER(C.c_int32))
        L.data = ptrs
        L.dx, L.dy = lv.dx, lv.dy
    lv0 = levels[0]
    names = list(VARNAMES) + [f"extra_comp_{k}" for k in range(len(VARNAMES), ncomp)]
    cnames = (C.c_char_p * ncomp)(*[n.encode() for n in names])
    lo = (C.c_double * 2)(*prob_lo)
    hi = (C.c_double * 2)(prob_lo[0] + lv0.nx * lv0.dx, prob_lo[1] + lv0.ny * lv0.dy)
    rr = (C.c_int32 * max(1, len(ref_ratio)))(*(list(ref_ratio) or [0]))
    _check(_bind().hxio_write_plotfile(os.fsencode(name), len(levels), arr, ncomp, cnames,
                                       float(time), lo, hi, rr), f"write_native_plotfile({name!r})")
    return name


def _box(text: str):
    m = re.match(r"\(\((-?\d+),(-?\d+)\) \((-?\d+),(-?\d+)\) \((\d+),(\d+)\)\)", text)
    return tuple(int(m.group(k)) for k in range(1, 5))


def read_native_plotfile(path: str) -> Dict:
    """Parse a native plotfile: header fields and, per level, the boxes and
    their data (ncomp, ny, nx) arrays."""
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    out = {"version": next(it)}
    nvar = int(next(it))
    out["varnames"] = [next(it) for _ in range(nvar)]
    out["dim"] = int(next(it))
    out["time"] = float(next(it))
    finest = int(next(it))
    out["finest_level"] = finest
    out["prob_lo"] = [float(x) for x in next(it).split()]
    out["prob_hi"] = [float(x) for x in next(it).split()]
    out["ref_ratio"] = [int(x) for x in next(it).split()]
    out["domains"] = [_box(m) for m in
                      re.findall(r"\(\(-?\d+,-?\d+\) \(-?\d+,-?\d+\) \(\d+,\d+\)\)", next(it))]
    out["steps"] = [int(x) for x in next(it).split()]
    out["cell_size"] = [[float(x) for x in next(it).split()] for _ in range(finest + 1)]
    out["coord"] = int(next(it))
    out["bwidth"] = int(next(it))
    levels = []
    for _ in range(finest + 1):
        lev, nb, t = next(it).split()
        step = int(next(it))
        extents = []
        for _b in range(int(nb)):
            x = [float(v) for v in next(it).split()]
            y = [float(v) for v in next(it).split()]
            extents.append((x[0], x[1], y[0], y[1]))
        rel = next(it)
        levels.append(_read_level(os.path.join(path, rel), int(lev), step, extents))
    out["levels"] = levels
    return out


def _read_level(cell_path: str, lev: int, step: int, extents) -> Dict:
    with open(cell_path + "_H") as f:
        h = f.read().split("\n")
    vers, how, ncomp, ngrow = (int(h[k]) for k in range(4))
    nb = int(h[4].strip("(").split()[0])
    boxes = [_box(h[5 + b]) for b in range(nb)]
    k = 5 + nb + 1
    nfod = int(h[k])
    fods = []
    for b in range(nfod):
        _, fname, off = h[k + 1 + b].split()
        fods.append((fname, int(off)))
    k = k + 1 + nfod + 1
    mins = [[float(v) for v in h[k + 1 + b].rstrip(",").split(",")] for b in range(nb)]
    k = k + 1 + nb + 1
    maxs = [[float(v) for v in h[k + 1 + b].rstrip(",").split(",")] for b in range(nb)]
    data = []
    base = os.path.dirname(cell_path)
    for (fname, off), (ilo, jlo, ihi, jhi) in zip(fods, boxes):
        with open(os.path.join(base, fname), "rb") as f:
            f.seek(off)
            header = f.readline().decode()
            nc = int(header.rsplit(")", 1)[1])
            ny, nx = jhi - jlo + 1, ihi - ilo + 1
            a = np.frombuffer(f.read(8 * nc * ny * nx), dtype="<f8").reshape(nc, ny, nx)
        data.append({"box": (ilo, jlo, ihi, jhi), "fab_header": header.rstrip("\n"), "data": a})
    return {"level": lev, "step": step, "extents": extents, "vismf_version": vers, "how": how,
            "ncomp": ncomp, "ngrow": ngrow, "boxes": boxes, "min": mins, "max": maxs,
            "fabs": data}
