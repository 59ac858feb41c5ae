"""ctypes binding of the C-ABI in ``include/hydroexa.h``.

torch is imported first so that ``libhydroexa.so`` binds to the HIP runtime
torch already loaded (same soname): device pointers and streams are then
shared between torch and the library.  A missing library is an error -- there
is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libhydroexa.so")

HX_OK = 0
HX_E_BADARG = 2
HX_E_DT_COLLAPSE = 3
HX_E_HIP = 4
HX_E_NOMEM = 5
HX_E_STATE = 6

HX_SOLVER_ROE = 0
HX_SOLVER_HYDROSTATIC_ROE = 1
HX_BC_REFLECT = 0
HX_BC_TRANSMIT = 1
HX_BC_GHOST = -1
HX_F64 = 0
HX_F32 = 1
HX_ARITH_EXACT = 0
HX_ARITH_FAST = 1

# Every symbol include/hydroexa.h declares.
EXPORTS = (
    "hx_ctx_create", "hx_ctx_destroy", "hx_abi_version", "hx_status_string",
    "hx_advance_level", "hx_advance_level_fluxes", "hx_level_dt", "hx_tag_level",
    "hx_masked_average_down",
    "hx_prune_boxes", "hx_run_level", "hx_run_level_resume", "hx_run_level_noalloc", "hx_set_tuning", "hx_set_arith",
    "hx_ensemble_create", "hx_ensemble_destroy", "hx_ensemble_set_tuning", "hx_ensemble_run",
    "hx_adjoint_create", "hx_adjoint_destroy", "hx_adjoint_forward", "hx_adjoint_backward",
    "hx_adjoint_batch_create", "hx_adjoint_batch_forward", "hx_adjoint_batch_backward",
    "hx_conservation", "hx_face_fluxes", "hx_fill_fine_ghosts", "hx_reflux",
    "hx_average_down_state", "hx_two_level_run", "hx_two_level_enqueue", "hx_fill_fine_region", "hx_reflux_side",
    "hx_boxes_dt", "hx_amr_run", "hx_tag_blocks", "hx_cluster_blocks", "hx_stream_probe",
)


class HxAmrLevel(C.Structure):
    """Mirror of ``hx_amr_level``."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("npatch", C.c_int32), ("cur", C.c_int32),
                ("patches", C.POINTER(C.c_int32)), ("u", C.c_void_p * 2), ("terrain", C.c_void_p),
                ("mask", C.c_void_p), ("regs", C.c_void_p), ("dx", C.c_double), ("dy", C.c_double),
                ("t_new", C.c_double), ("t_old", C.c_double), ("istep", C.c_int32),
                ("last_regrid", C.c_int32)]


HX_AMR_REGRID_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double)


class HxPatch(C.Structure):
    """Mirror of ``hx_patch``."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("ng", C.c_int32), ("dtype", C.c_int32),
        ("pitch", C.c_int64),
        ("h", C.c_void_p), ("hu", C.c_void_p), ("hv", C.c_void_p), ("eta", C.c_void_p),
        ("z", C.c_void_p), ("n", C.c_void_p),
        ("dx", C.c_double), ("dy", C.c_double),
        ("bc_lo", C.c_int32 * 2), ("bc_hi", C.c_int32 * 2),
    ]


class HydroExaError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} (status {status})")


class DtCollapse(HydroExaError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc, gfx950). The solver has no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    vp, dp, i32p, u8p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
    pp = C.POINTER(HxPatch)
    L.hx_ctx_create.argtypes = [C.POINTER(vp), C.c_int]
    L.hx_ctx_destroy.argtypes = [vp]
    L.hx_abi_version.argtypes = []
    L.hx_status_string.argtypes = [C.c_int]
    L.hx_status_string.restype = C.c_char_p
    L.hx_advance_level.argtypes = [vp, C.c_int, pp, C.c_double, vp, vp, vp, C.c_int64, vp]
    L.hx_advance_level_fluxes.argtypes = [vp, C.c_int, pp, C.c_double, vp, vp, vp, C.c_int64, vp, vp]
    L.hx_level_dt.argtypes = [vp, C.c_int, pp, C.c_double, dp, dp, vp]
    L.hx_tag_level.argtypes = [vp, C.c_int, pp, C.c_double, dp, vp, vp]
    L.hx_masked_average_down.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int,
                                         C.c_int, C.c_double, C.c_double, C.c_double, vp, vp, vp]
    L.hx_prune_boxes.argtypes = [vp, C.c_int, C.c_int, i32p, C.c_int, C.c_int, vp, C.c_int,
                                 u8p, u8p, i32p, vp]
    L.hx_run_level.argtypes = [vp, C.c_int, pp, vp, vp, vp, C.c_double, C.c_int, dp, dp, i32p, vp]
    L.hx_run_level_resume.argtypes = L.hx_run_level.argtypes
    L.hx_run_level_noalloc.argtypes = L.hx_run_level.argtypes
    L.hx_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.hx_set_arith.argtypes = [vp, C.c_int]
    L.hx_ensemble_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, pp, C.POINTER(vp), C.POINTER(vp)]
    L.hx_ensemble_destroy.argtypes = [vp]
    L.hx_ensemble_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.hx_ensemble_run.argtypes = [vp, C.c_double, C.c_int, C.c_int, dp, i32p, vp]
    L.hx_adjoint_create.argtypes = [vp, C.c_int, pp, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_int64), C.POINTER(vp)]
    L.hx_adjoint_destroy.argtypes = [vp]
    L.hx_conservation.argtypes = [vp, pp, dp, vp]
    L.hx_stream_probe.argtypes = [vp, C.c_int64, C.POINTER(vp), C.POINTER(vp), vp]
    L.hx_face_fluxes.argtypes = [vp, C.c_int, pp, C.c_int, vp, C.c_int, C.c_double, C.c_double,
                                 C.c_int, vp, vp]
    L.hx_fill_fine_region.argtypes = [vp, pp, pp, C.c_double, C.c_int, C.c_int, C.c_int, pp,
                                      C.c_int, vp, C.c_int64, vp, vp, vp, vp]
    L.hx_reflux_side.argtypes = [vp, pp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                 C.c_int, vp]
    L.hx_boxes_dt.argtypes = [vp, C.c_int, C.c_int, pp, C.c_double, dp, dp, vp]
    L.hx_tag_blocks.argtypes = [vp, vp, C.c_int64, vp, C.c_int64, C.c_int, C.c_int, C.c_int,
                                C.c_int, vp, vp]
    L.hx_cluster_blocks.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_int, i32p, C.c_int, i32p]
    L.hx_amr_run.argtypes = [vp, C.c_int, C.c_int, C.POINTER(HxAmrLevel), C.POINTER(C.c_int),
                             C.c_int, i32p, i32p, C.c_int, C.c_int, C.c_int, HX_AMR_REGRID_FN, vp,
                             C.c_double, C.c_double, C.c_int, dp, dp, i32p, vp]
    L.hx_two_level_run.argtypes = [vp, C.c_int, pp, vp, vp, vp, pp, vp, vp, vp, C.c_int, C.c_int,
                                   C.c_int, C.c_double, C.c_int, vp, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int32), vp]
    L.hx_two_level_enqueue.argtypes = [vp, C.c_int, pp, vp, vp, vp, pp, vp, vp, vp, C.c_int,
                                       C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp, vp]
    L.hx_fill_fine_ghosts.argtypes = [vp, pp, pp, C.c_double, C.c_int, C.c_int, C.c_int, pp,
                                      C.c_int, vp, vp, vp, vp]
    L.hx_reflux.argtypes = [vp, pp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
    L.hx_average_down_state.argtypes = [vp, pp, C.c_int, C.c_int, C.c_int, pp, vp, vp, vp, vp]
    L.hx_adjoint_forward.argtypes = [vp, pp, C.c_double, dp, vp, vp, vp, vp, i32p, vp]
    L.hx_adjoint_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.hx_adjoint_batch_create.argtypes = [vp, C.c_int, C.c_int, pp, C.c_int, C.c_int,
                                          C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_int64)),
                                          C.POINTER(vp)]
    L.hx_adjoint_batch_forward.argtypes = [vp, pp, C.c_double, dp, C.POINTER(vp), C.POINTER(vp),
                                           i32p, vp]
    L.hx_adjoint_batch_backward.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                            C.POINTER(vp), C.POINTER(vp), vp]
    _lib = L
    return L


def status_string(status: int) -> str:
    try:
        return load().hx_status_string(status).decode()
    except ImportError:
        return "unknown"


def check(status: int, what: str) -> None:
    if status == HX_OK:
        return
    if status == HX_E_DT_COLLAPSE:
        raise DtCollapse(status, what)
    raise HydroExaError(status, what)
