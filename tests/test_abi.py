"""The C-ABI library loads and exports every symbol include/hydroexa.h
declares; argument checks that run before any device work behave as
documented.  No GPU needed."""
import ctypes as C
import os
import re

import pytest

from hydroexa_amd import _lib as L
from hydroexa_amd import SolverVariant

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hydroexa.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hx_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_match_binding():
    assert declared_functions() == sorted(L.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_abi_version_and_status_strings():
    lib = L.load()
    assert lib.hx_abi_version() == 1
    assert L.status_string(L.HX_E_DT_COLLAPSE) == "dt collapsed below 1e-12"
    assert L.status_string(12345) == "unknown status"


def test_calls_without_context_are_rejected():
    lib = L.load()
    p = L.HxPatch()
    assert lib.hx_advance_level(None, 0, C.byref(p), 1e-3, None, None, None, 0, None) == L.HX_E_STATE
    assert lib.hx_level_dt(None, 0, C.byref(p), 0.5, None, None, None) == L.HX_E_STATE
    assert lib.hx_ctx_destroy(None) == L.HX_OK


def test_solver_variant_requires_a_model():
    with pytest.raises(ValueError):
        SolverVariant("Exner")
    sv = SolverVariant()
    assert not sv.has_solver()
    with pytest.raises(RuntimeError, match="Solver not initialized"):
        sv._kind()


def test_patch_struct_layout():
    """hx_patch: 4 int32, int64 pitch, 6 pointers, 2 doubles, 4 int32."""
    assert C.sizeof(L.HxPatch) == 104
    assert L.HxPatch.pitch.offset == 16
    assert L.HxPatch.dx.offset == 72
    assert L.HxPatch.bc_hi.offset == 96


def test_io_header_declarations_match_binding():
    """include/hydroexa_io.h against libhydroexa_io.so and its binding."""
    from hydroexa_amd import h5input as H
    if not os.path.exists(H._IO_PATH):
        pytest.skip("libhydroexa_io.so not built (no HDF5)")
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "hydroexa_io.h")).read(),
                  flags=re.S)
    names = sorted(set(re.findall(r"\b(hxio_[a-z_0-9]+)\s*\(", text)))
    assert names == sorted(H.IO_EXPORTS)
    lib = H.load_io()
    for n in names:
        assert hasattr(lib, n), n
    assert C.sizeof(H.HxioMeta) == 4 * 4 + 6 * 8
    assert C.sizeof(H.HxioGeometry) == 4 * 4 + 7 * 8
    from hydroexa_amd.plotfile import HxioPlotLevel
    assert C.sizeof(HxioPlotLevel) == 6 * 4 + 2 * 8 + 2 * 8
