"""HydroEXA's HDF5 input schema: read, write, geometry contract, level 0.

Thin ctypes layer over ``libhydroexa_io.so`` (include/hydroexa_io.h); the
schema is PyHydroEXA's (PyHydroEXA/src/pyhydroexa/preprocessing/hdf5.py:1-135):
dataset "terrain" (2, ny, nx) = z, roughness; dataset "fluid" (3, ny, nx) =
h, hu, hv; attributes dx, dy, x_ll, y_ll; NODATA -9999.

``level0_from_file`` builds level 0 the way the reference initialises it:
the whole file is read as the static layer (terrain NODATA-filled, fluid
0-filled outside the file; src/state/AmrMeshState_Init.cpp:150-170), then

* terrain_ref_lev = 0: level 0 is the static layer (a ratio-1 interpolation
  is a copy), with h rebuilt as max(0, eta - z), eta = z at NODATA
  (src/state/AmrMeshState_AMR.cpp:75-107);
* terrain_ref_lev > 0: terrain by masked_average_down (the library's
  ``hx_masked_average_down``), momenta and eta_s = h + z by AMReX's plain
  average_down, h = max(0, eta - z) on the level's own bed
  (src/state/AmrMeshState_AMR.cpp:18-31, 109-146).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Sequence

import numpy as np
import torch

from . import _lib as L
from .solver import Level, masked_average_down

NODATA = -9999.0
_HERE = os.path.dirname(os.path.abspath(__file__))
_IO_PATH = os.environ.get("HYDROEXA_IO_LIB", os.path.join(_HERE, "lib", "libhydroexa_io.so"))
_io = None

HXIO_OK, HXIO_E_BADARG, HXIO_E_OPEN, HXIO_E_DATASET = 0, 2, 10, 11
HXIO_E_SHAPE, HXIO_E_READ, HXIO_E_WRITE, HXIO_E_GEOMETRY = 12, 13, 14, 15
_MSG = {HXIO_E_BADARG: "invalid argument", HXIO_E_OPEN: "cannot open file",
        HXIO_E_DATASET: "dataset missing", HXIO_E_SHAPE: "dataset shape mismatch",
        HXIO_E_READ: "read failed", HXIO_E_WRITE: "write failed",
        HXIO_E_GEOMETRY: "geometry contract violated"}


class InputError(RuntimeError):
    pass


class GeometryError(InputError):
    pass


class HxioMeta(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("ndims", C.c_int32), ("ncomp", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("x_ll", C.c_double), ("y_ll", C.c_double),
                ("x_hi", C.c_double), ("y_hi", C.c_double)]


class HxioGeometry(C.Structure):
    _fields_ = [("base_nx", C.c_int32), ("base_ny", C.c_int32), ("refinement_scale", C.c_int32),
                ("pad", C.c_int32), ("file_cell_size", C.c_double), ("level0_dx", C.c_double),
                ("level0_dy", C.c_double), ("prob_lo", C.c_double * 2),
                ("prob_hi", C.c_double * 2)]


IO_EXPORTS = ("hxio_read_metadata", "hxio_read_components", "hxio_write_input",
              "hxio_check_geometry", "hxio_write_plotfile")


def load_io():
    """The HDF5 input library (built by hydroexa_amd.build.build_io)."""
    global _io
    if _io is None:
        if not os.path.exists(_IO_PATH):
            raise ImportError(f"{_IO_PATH} is missing: build it with hydroexa_amd.build.build_io() "
                              "(needs the HDF5 C library)")
        lib = C.CDLL(_IO_PATH)
        dp = C.POINTER(C.c_double)
        lib.hxio_read_metadata.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(HxioMeta)]
        lib.hxio_read_components.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_int, C.c_int, C.c_double, dp]
        lib.hxio_write_input.argtypes = [C.c_char_p, dp, dp, C.c_int, C.c_int, C.c_double,
                                         C.c_double, C.c_double, C.c_double]
        lib.hxio_check_geometry.argtypes = [C.POINTER(HxioMeta), C.c_int, C.c_int,
                                            C.POINTER(HxioGeometry), C.c_char_p, C.c_int]
        _io = lib
    return _io


def _check(rc: int, what: str) -> None:
    if rc != HXIO_OK:
        raise InputError(f"{what}: {_MSG.get(rc, 'error')} (status {rc})")


def read_metadata(path: str, dataset: str = "terrain") -> Dict:
    """IOHandler::ReadHDF5Metadata (src/io/IOHandler.cpp:230-307)."""
    m = HxioMeta()
    _check(load_io().hxio_read_metadata(os.fsencode(path), dataset.encode(), C.byref(m)),
           f"read_metadata({path!r}, {dataset!r})")
    return {k: getattr(m, k) for k, _ in HxioMeta._fields_}


def read_components(path: str, dataset: str, first_comp: int, ncomp: int, i_lo: int, j_lo: int,
                    nx: int, ny: int, nodata: float) -> np.ndarray:
    """IOHandler::ReadHDF5HyperslabComponents (src/io/IOHandler.cpp:127-227):
    (ncomp, ny, nx) over the box, nodata outside the file."""
    out = np.empty((ncomp, ny, nx))
    _check(load_io().hxio_read_components(os.fsencode(path), dataset.encode(), first_comp, ncomp,
                                          i_lo, j_lo, nx, ny, nodata,
                                          out.ctypes.data_as(C.POINTER(C.c_double))),
           f"read_components({path!r}, {dataset!r})")
    return out


def write_input(path: str, terrain: np.ndarray, fluid: np.ndarray, dx: float, dy: float,
                x_ll: float = 0.0, y_ll: float = 0.0) -> None:
    """PyHydroEXA's write_hdf5 (preprocessing/hdf5.py:51-135)."""
    t = np.ascontiguousarray(terrain, dtype=np.float64)
    f = np.ascontiguousarray(fluid, dtype=np.float64)
    if t.ndim != 3 or t.shape[0] != 2:
        raise ValueError(f"terrain must have shape (2, ny, nx), got {t.shape}")
    if f.ndim != 3 or f.shape[0] != 3:
        raise ValueError(f"fluid must have shape (3, ny, nx), got {f.shape}")
    if t.shape[1:] != f.shape[1:]:
        raise ValueError("terrain and fluid dimensions must match")
    ny, nx = t.shape[1:]
    dp = C.POINTER(C.c_double)
    _check(load_io().hxio_write_input(os.fsencode(path), t.ctypes.data_as(dp), f.ctypes.data_as(dp),
                                      nx, ny, dx, dy, x_ll, y_ll), f"write_input({path!r})")


def check_geometry(meta: Dict, terrain_ref_lev: int = 0, blocking_factor: int = 8) -> Dict:
    """The geometry contract of HydroEXA::Initialize (src/HydroEXA.cpp:95-205);
    raises GeometryError with the reference's assertion text."""
    m = HxioMeta(**{k: meta[k] for k, _ in HxioMeta._fields_})
    g = HxioGeometry()
    msg = C.create_string_buffer(256)
    rc = load_io().hxio_check_geometry(C.byref(m), terrain_ref_lev, blocking_factor, C.byref(g),
                                       msg, 256)
    if rc == HXIO_E_GEOMETRY:
        raise GeometryError(msg.value.decode())
    _check(rc, "check_geometry")
    return {"base_nx": g.base_nx, "base_ny": g.base_ny, "refinement_scale": g.refinement_scale,
            "file_cell_size": g.file_cell_size, "level0_dx": g.level0_dx,
            "level0_dy": g.level0_dy, "prob_lo": tuple(g.prob_lo), "prob_hi": tuple(g.prob_hi)}


def _average_down(fine: torch.Tensor, r: int) -> torch.Tensor:
    """AMReX's plain average_down for cell data: c = sum over jref (outer)
    and iref (inner) of the r x r children, times 1/r^2, in that order."""
    vol = 1.0 / (r * r)
    c = torch.zeros_like(fine[..., ::r, ::r])
    for jref in range(r):
        for iref in range(r):
            c = c + fine[..., jref::r, iref::r]
    return vol * c


def _static_layer(path: str):
    meta = read_metadata(path, "terrain")
    nx, ny = meta["nx"], meta["ny"]
    terr = read_components(path, "terrain", 0, 2, 0, 0, nx, ny, NODATA)
    fluid = read_components(path, "fluid", 0, 3, 0, 0, nx, ny, 0.0)
    return meta, terr, fluid


def level_from_static(meta, terr, fluid, terrain_ref_lev: int, lev: int,
                      bc_lo: Sequence[int] = (0, 0), bc_hi: Sequence[int] = (0, 0),
                      device: str = "cuda"):
    """(h, hu, hv, z, n) of level lev (FP64 device tensors, (ny, nx)) from the
    static layer of a file read at level terrain_ref_lev:
    TerrainMapStaticToDynamic + FluidMapStaticToDynamic
    (src/state/AmrMeshState_AMR.cpp:8-195).  At or below the file's level
    the terrain is masked-averaged and eta_s = h + z (z at NODATA) and the
    momenta are averaged; one level finer than the file the terrain is
    piecewise constant and eta_s and the momenta are interpolated by
    cell_cons_interp (the library's ghost-fill kernel over the whole level).
    Either way h = max(0, eta - z) on the level's own bed.  Two or more
    levels finer than the file (ratios of 4 and up, where AMReX's min/max
    limiter can bind) are not supported."""
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=device)
    zs, ns = T(terr[0]), T(terr[1])
    hs, qxs, qys = T(fluid[0]), T(fluid[1]), T(fluid[2])
    eta_s = torch.where(ns == NODATA, zs, hs + zs)       # eta = z at NODATA
    if lev <= terrain_ref_lev:
        r = 2 ** (terrain_ref_lev - lev)
        if r == 1:
            z, n, eta, qx, qy = zs, ns, eta_s, qxs, qys
        else:
            t, _ = masked_average_down(torch.stack([zs, ns]).contiguous(), (r, r), meta["dx"],
                                       meta["dy"], NODATA)
            z, n = t[0], t[1]
            qx, qy = _average_down(qxs, r), _average_down(qys, r)
            eta = _average_down(eta_s, r)
        x = eta - z
        h = torch.where(x > 0, x, torch.zeros_like(x))     # max(0, eta - z)
        return h, qx, qy, z, n
    if lev != terrain_ref_lev + 1:
        raise ValueError("levels more than one above the file's resolution are not supported")
    import ctypes as C
    from .amr import DenseLevel, _pad_edges, _pc_refine
    from .solver import Level, context
    ny, nx = zs.shape
    # the interpolation reads h + z as the free surface: h = 0 at NODATA
    # makes that eta_s exactly (z there), the file's h elsewhere
    hsrc = torch.where(ns == NODATA, torch.zeros_like(hs), hs)
    src = Level(hsrc, qxs, qys, zs, ns, meta["dx"], meta["dy"], bc_lo, bc_hi)
    fl = DenseLevel(2 * nx, 2 * ny, meta["dx"] / 2, meta["dy"] / 2, bc_lo, bc_hi, torch.float64)
    zc = torch.stack([zs, ns]).double().cpu().numpy()
    fl.terrain.copy_(torch.as_tensor(np.stack([_pad_edges(_pc_refine(zc[k])) for k in range(2)]),
                                     device=device))
    fl.set_boxes([(0, 0, 2 * nx, 2 * ny)])
    cp = src.patch()
    bp = fl.patch(fl.U_new, (0, 0, 2 * nx, 2 * ny))
    L.check(L.load().hx_fill_fine_region(
        context(), C.byref(cp), C.byref(cp), 0.0, 2, 0, 0, C.byref(bp), 1, None, fl.pitch,
        fl.addr(fl.U_new, 0, 0, 0), fl.addr(fl.U_new, 1, 0, 0), fl.addr(fl.U_new, 2, 0, 0),
        C.c_void_p(torch.cuda.current_stream().cuda_stream)), "hx_fill_fine_region")
    U, Tr = fl.interior(fl.U_new), fl.interior(fl.terrain)
    return U[0].clone(), U[1].clone(), U[2].clone(), Tr[0].clone(), Tr[1].clone()


def hierarchy_from_file(path: str, model: str, terrain_ref_lev: int = 0, max_level: int = 1,
                        blocking_factor: int = 8, bc_lo: Sequence[int] = (0, 0),
                        bc_hi: Sequence[int] = (0, 0), **kw):
    """An AmrHierarchy initialised from an input file the way the reference
    initialises a run: level 0 and every new level from the static layer
    (MakeNewLevelFromScratch), the geometry contract checked first.  kw goes
    to AmrHierarchy (regrid_int, tag_thresh, subcycle, ...)."""
    from .amr import AmrHierarchy
    meta, terr, fluid = _static_layer(path)
    geo = check_geometry(meta, terrain_ref_lev, blocking_factor)
    lv0 = [a.double().cpu().numpy() for a in level_from_static(meta, terr, fluid,
                                                                 terrain_ref_lev, 0, bc_lo, bc_hi)]
    cache = {}

    def static(lev):
        if lev not in cache:
            cache[lev] = [a.double().cpu().numpy() for a in
                          level_from_static(meta, terr, fluid, terrain_ref_lev, lev, bc_lo, bc_hi)]
        return cache[lev]

    H = AmrHierarchy(model, *lv0, geo["level0_dx"], geo["level0_dy"], bc_lo, bc_hi,
                     max_level=max_level, blocking_factor=blocking_factor, static=static, **kw)
    return H, geo


def level0_from_file(path: str, terrain_ref_lev: int = 0, blocking_factor: int = 8,
                     bc_lo: Sequence[int] = (0, 0), bc_hi: Sequence[int] = (0, 0),
                     dtype: torch.dtype = torch.float64, device: str = "cuda"):
    """Level 0 of a run from an input file.  Returns (Level, geometry)."""
    meta = read_metadata(path, "terrain")
    geo = check_geometry(meta, terrain_ref_lev, blocking_factor)
    nx, ny = meta["nx"], meta["ny"]
    terr = read_components(path, "terrain", 0, 2, 0, 0, nx, ny, NODATA)
    fluid = read_components(path, "fluid", 0, 3, 0, 0, nx, ny, 0.0)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=device)
    zs, ns = T(terr[0]), T(terr[1])
    hs, qxs, qys = T(fluid[0]), T(fluid[1]), T(fluid[2])
    eta_s = torch.where(ns == NODATA, zs, hs + zs)       # eta = z at NODATA
    r = geo["refinement_scale"]
    if r == 1:
        z, n, eta, qx, qy = zs, ns, eta_s, qxs, qys
    else:
        t, _ = masked_average_down(torch.stack([zs, ns]).contiguous(), (r, r), meta["dx"],
                                   meta["dy"], NODATA)
        z, n = t[0], t[1]
        qx, qy = _average_down(qxs, r), _average_down(qys, r)
        eta = _average_down(eta_s, r)
    x = eta - z
    h = torch.where(x > 0, x, torch.zeros_like(x))     # max(0, eta - z)
    lv = Level(h.to(dtype), qx.to(dtype), qy.to(dtype), z.to(dtype), n.to(dtype),
               geo["level0_dx"], geo["level0_dy"], bc_lo, bc_hi, dtype=dtype, device=device)
    return lv, geo
