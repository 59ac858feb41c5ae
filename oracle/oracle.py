"""ctypes wrapper around ``libhx_oracle.so`` (the FP64 CPU restatement).

Test infrastructure only.  Arrays are numpy float64, shape (ny, nx), C order,
so element (i, j) of the reference's Array4 is ``a[j, i]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhx_oracle.so")
_LIB32_PATH = os.path.join(_HERE, "libhx_oracle_f32.so")
_lib = None
_lib32 = None

ROE = 0
HYDROSTATIC_ROE = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no FP contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.hxo_roe_face.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double,
                                   C.c_double, _dp, _dp, _dp]
        L.hxo_hroe_face.argtypes = [C.c_double] * 6 + [C.c_double, C.c_double, _dp]
        L.hxo_vanleer_slope.argtypes = [C.c_double] * 3
        L.hxo_vanleer_slope.restype = C.c_double
        L.hxo_hydro_face.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp]
        step_args = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                     C.c_double, _ip, _ip, _dp, _dp, _dp]
        L.hxo_roe_step.argtypes = step_args
        L.hxo_hydro_step.argtypes = step_args
        L.hxo_step_padded.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                      C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.hxo_level_dt.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                                   C.c_double, _dp, _dp]
        L.hxo_tag.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                              C.c_double, _ip, _ip, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.POINTER(C.c_uint8)]
        L.hxo_masked_average_down.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int,
                                              C.c_int, C.c_double, C.c_double, C.c_double,
                                              _dp, C.POINTER(C.c_int32)]
        L.hxo_prune_boxes.argtypes = [C.c_int, C.POINTER(C.c_int32), C.c_int, C.c_int, _dp,
                                      C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]
        L.hxo_conservation.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                                       C.c_double, _dp]
        L.hxo_run.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double,
                              C.c_double, _ip, _ip, C.c_double, C.c_int, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _bc(bc_lo, bc_hi):
    lo = (C.c_int * 2)(*bc_lo)
    hi = (C.c_int * 2)(*bc_hi)
    return lo, hi


def roe_face(L, R, dt, ds, nx, ny):
    """Plain Roe face: returns (F, S_minus, S_plus), each length 3."""
    Lb, Rb = _f64(L), _f64(R)
    F, Sm, Sp = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().hxo_roe_face(_p(Lb), _p(Rb), dt, ds, nx, ny, _p(F), _p(Sm), _p(Sp))
    return F, Sm, Sp


def hroe_face(hi, hui, hvi, hj, huj, hvj, nx, ny):
    F = np.zeros(3)
    lib().hxo_hroe_face(hi, hui, hvi, hj, huj, hvj, nx, ny, _p(F))
    return F


def vanleer_slope(qm, q, qp):
    return lib().hxo_vanleer_slope(qm, q, qp)


def hydro_face(cells, nx, ny):
    """cells: 4x4 array, rows Lm, L, R, Rp of {h, hu, hv, z}; returns (FL, FR)."""
    c = _f64(cells).reshape(16)
    FL, FR = np.zeros(3), np.zeros(3)
    lib().hxo_hydro_face(_p(c), nx, ny, _p(FL), _p(FR))
    return FL, FR


def step(solver, h, hu, hv, z, n, dt, dx, dy, bc_lo=(1, 1), bc_hi=(1, 1)):
    """One AdvanceLevel of the chosen solver; returns (h, hu, hv) new."""
    h, hu, hv, z, n = (_f64(a) for a in (h, hu, hv, z, n))
    ny, nx = h.shape
    ho, huo, hvo = np.empty_like(h), np.empty_like(h), np.empty_like(h)
    lo, hi = _bc(bc_lo, bc_hi)
    fn = lib().hxo_roe_step if solver == ROE else lib().hxo_hydro_step
    rc = fn(nx, ny, _p(h), _p(hu), _p(hv), _p(z), _p(n), dt, dx, dy, lo, hi,
            _p(ho), _p(huo), _p(hvo))
    if rc != 0:
        raise ValueError(f"oracle step failed with status {rc}")
    return ho, huo, hvo


def step_padded(solver, h, hu, hv, z, n, dt, dx, dy, eta=None):
    """One step of a box whose 2 ghost layers are given: inputs (ny+4, nx+4),
    outputs (ny, nx).  eta (optional) is each cell's free surface in place of
    h + z."""
    a = [_f64(x) for x in (h, hu, hv, z, n)]
    e = _f64(eta) if eta is not None else None
    ny, nx = a[0].shape[0] - 4, a[0].shape[1] - 4
    out = [np.empty((ny, nx)) for _ in range(3)]
    rc = lib().hxo_step_padded(solver, nx, ny, *[_p(x) for x in a],
                               _p(e) if e is not None else None, dt, dx, dy,
                               *[_p(o) for o in out])
    if rc != 0:
        raise ValueError(f"hxo_step_padded failed ({rc})")
    return tuple(out)


def level_dt(h, hu, hv, z, dx, dy):
    """(dt_est, max_vel) before the cfl factor."""
    h, hu, hv, z = (_f64(a) for a in (h, hu, hv, z))
    ny, nx = h.shape
    d, v = C.c_double(), C.c_double()
    lib().hxo_level_dt(nx, ny, _p(h), _p(hu), _p(hv), _p(z), dx, dy, C.byref(d), C.byref(v))
    return d.value, v.value


def tag(solver, h, hu, hv, z, dx, dy, time, h_thresh, u_thresh, z_thresh,
        bc_lo=(1, 1), bc_hi=(1, 1)):
    h, hu, hv, z = (_f64(a) for a in (h, hu, hv, z))
    ny, nx = h.shape
    out = np.zeros((ny, nx), dtype=np.uint8)
    lo, hi = _bc(bc_lo, bc_hi)
    rc = lib().hxo_tag(solver, nx, ny, _p(h), _p(hu), _p(hv), _p(z), dx, dy, lo, hi, time,
                       h_thresh, u_thresh, z_thresh,
                       out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc != 0:
        raise ValueError(f"oracle tag failed with status {rc}")
    return out


def masked_average_down(fine, rx, ry, dxf, dyf, nodata=-9999.0):
    """fine: (ncomp, nyf, nxf); returns (crse, nvalid) of shape (ncomp, nyf/ry, nxf/rx)."""
    f = _f64(fine)
    if f.ndim == 2:
        f = f[None]
    nc, nyf, nxf = f.shape
    crse = np.zeros((nc, nyf // ry, nxf // rx))
    nval = np.zeros((nc, nyf // ry, nxf // rx), dtype=np.int32)
    rc = lib().hxo_masked_average_down(nxf, nyf, nc, _p(f), rx, ry, dxf, dyf, nodata,
                                       _p(crse), nval.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise ValueError(f"oracle masked_average_down failed with status {rc}")
    return crse, nval


def prune_boxes(boxes, z_static, scale_log2):
    """boxes: (nbox, 4) int {ilo, jlo, ihi, jhi}; returns (keep, has_nodata, nkeep)."""
    b = np.ascontiguousarray(boxes, dtype=np.int32)
    z = _f64(z_static)
    nys, nxs = z.shape
    keep = np.zeros(len(b), dtype=np.uint8)
    nd = np.zeros(len(b), dtype=np.uint8)
    nk = lib().hxo_prune_boxes(len(b), b.ctypes.data_as(C.POINTER(C.c_int32)), nxs, nys,
                               _p(z), scale_log2, keep.ctypes.data_as(C.POINTER(C.c_uint8)),
                               nd.ctypes.data_as(C.POINTER(C.c_uint8)))
    return keep.astype(bool), nd.astype(bool), nk


def run(solver, h, hu, hv, z, n, dx, dy, cfl, nsteps, bc_lo=(1, 1), bc_hi=(1, 1)):
    """Step-driven loop (ComputeDt + AdvanceLevel); returns (h, hu, hv, dt_hist, vel_hist)."""
    h, hu, hv = (_f64(a).copy() for a in (h, hu, hv))
    z, n = _f64(z), _f64(n)
    ny, nx = h.shape
    dth = np.zeros(nsteps)
    vh = np.zeros(nsteps)
    lo, hi = _bc(bc_lo, bc_hi)
    rc = lib().hxo_run(solver, nx, ny, _p(h), _p(hu), _p(hv), _p(z), _p(n), dx, dy, lo, hi,
                       cfl, nsteps, _p(dth), _p(vh))
    if rc != 0:
        raise RuntimeError(f"oracle run failed with status {rc}")
    return h, hu, hv, dth, vh


def run32(solver, h, hu, hv, z, n, dx, dy, cfl, nsteps, bc_lo=(1, 1), bc_hi=(1, 1)):
    """run() of the single-precision restatement (hx_oracle_f32.c): every
    operation rounded to float in the oracle's order.  Returns float32
    (h, hu, hv, dt_hist, vel_hist)."""
    global _lib32
    if _lib32 is None:
        if not os.path.exists(_LIB32_PATH):
            build(force=True)
        L = C.CDLL(_LIB32_PATH)
        fp = C.POINTER(C.c_float)
        L.hxo_run.argtypes = [C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, fp, C.c_float, C.c_float,
                              _ip, _ip, C.c_float, C.c_int, fp, fp]
        _lib32 = L
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    h, hu, hv = (f32(a).copy() for a in (h, hu, hv))
    z, n = f32(z), f32(n)
    ny, nx = h.shape
    dth = np.zeros(nsteps, np.float32)
    vh = np.zeros(nsteps, np.float32)
    lo, hi = _bc(bc_lo, bc_hi)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    rc = _lib32.hxo_run(solver, nx, ny, p(h), p(hu), p(hv), p(z), p(n), dx, dy, lo, hi, cfl,
                        nsteps, p(dth), p(vh))
    if rc != 0:
        raise RuntimeError(f"single-precision oracle run failed with status {rc}")
    return h, hu, hv, dth, vh


def conservation(h, hu, hv, n, dx, dy):
    """(mass, px, py, KE, min_h) of one level, serial cell-order sums."""
    h, hu, hv, n = (_f64(a) for a in (h, hu, hv, n))
    ny, nx = h.shape
    out = np.zeros(5)
    lib().hxo_conservation(nx, ny, _p(h), _p(hu), _p(hv), _p(n), dx, dy, _p(out))
    return out
