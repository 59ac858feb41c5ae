"""hx_run_level_noalloc (the allocation-free variant of hx_run_level, scratch
in the library's static per-device array) against hx_run_level, through the
C-ABI: two calls in a row, the second on a smaller level, bit for bit in
state, dt and max|vel| history; steps beyond the scratch's 8192 are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from hydroexa_amd import Level
from hydroexa_amd import _lib as L
from hydroexa_amd import workloads as W
from hydroexa_amd.solver import MODELS, _stream, context

pytestmark = pytest.mark.gpu


def call(fn, model, lv, cfl, nsteps):
    p = lv.patch(lv.U_new)
    plane = lv.nx * lv.ny * lv.U_new.element_size()
    alt = lv.U_old.data_ptr()
    dth, vh = np.zeros(max(nsteps, 1)), np.zeros(max(nsteps, 1))
    done = C.c_int32()
    rc = fn(context(), MODELS[model], C.byref(p), alt, alt + plane, alt + 2 * plane, float(cfl),
            int(nsteps), dth.ctypes.data_as(C.POINTER(C.c_double)),
            vh.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done), _stream())
    torch.cuda.synchronize()
    final = lv.U_old if nsteps % 2 else lv.U_new
    return rc, done.value, final.clone(), dth[:nsteps], vh[:nsteps]


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
def test_noalloc_run_equals_run_level(gpu, model):
    lib = L.load()
    big = W.dam_break(300, 1) if model == "Roe" else W.lake_at_rest(300, 1)
    small = W.dam_break(70, 1) if model == "Roe" else W.lake_at_rest(70, 1)
    if model == "HydrostaticRoe":
        for w in (big, small):
            w.h = w.h * np.where(np.arange(w.h.shape[1])[None, :] < w.h.shape[1] // 3, 1.4, 1.0)
            w.n = np.where(w.z < -9000, w.n, 0.03)
    for w, nsteps in ((big, 37), (small, 12)):
        print(f"noalloc {model} {w.h.shape} {nsteps} steps", flush=True)
        mk = lambda: Level(w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi)
        a, b = mk(), mk()
        ra = call(lib.hx_run_level, model, a, w.cfl, nsteps)
        rb = call(lib.hx_run_level_noalloc, model, b, w.cfl, nsteps)
        assert ra[0] == L.HX_OK and rb[0] == L.HX_OK
        assert ra[1] == rb[1] == nsteps
        assert torch.equal(ra[2], rb[2])
        np.testing.assert_array_equal(ra[3], rb[3])
        np.testing.assert_array_equal(ra[4], rb[4])
        assert np.all(rb[3] > 0)


def test_noalloc_refuses_more_steps_than_its_scratch(gpu):
    w = W.dam_break(16, 1)
    lv = Level(w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi)
    before = lv.U_new.clone()
    rc = call(L.load().hx_run_level_noalloc, "Roe", lv, w.cfl, 8193)[0]
    assert rc == L.HX_E_BADARG
    assert torch.equal(lv.U_new, before)
