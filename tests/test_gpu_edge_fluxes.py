"""Box-edge effective fluxes from the fused advance (hx_advance_level_fluxes):
what the reference hands to the fine side of a flux register
(src/solvers/Roe.cpp:544-566, D- on a box's low faces and D+ on its high
faces; src/solvers/HydrostaticRoe.cpp:612-627, 0.5 FL + 0.5 FR).

Checks, through the C-ABI:
* the advanced state is bit-identical to hx_advance_level's;
* every edge flux equals hx_face_fluxes' re-evaluation of the same face and
  side (==, so a zero's sign does not count), on physical sides and on a
  patch with stored ghosts;
* every edge flux is within 1e-12 of the oracle's face flux
  (oracle/amr.py face_flux), relative to the largest flux component;
* with no source share (flat, frictionless, wet), the box's change of mass
  equals the net edge mass flux: a telescoping identity that a wrong side,
  sign or face index breaks at O(1).
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from oracle import amr as A
from oracle import oracle as O
from hydroexa_amd import Level, SolverVariant, context, set_arith
from hydroexa_amd import _lib as L

pytestmark = pytest.mark.gpu

MODELS = {"Roe": O.ROE, "HydrostaticRoe": O.HYDROSTATIC_ROE}


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def state(rng, ny, nx, dry=False, flat=False):
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    z = np.zeros((ny, nx)) if flat else (0.2 * np.sin(0.37 * ii + 0.2) * np.cos(0.23 * jj)
                                         + 0.02 * rng.random((ny, nx)))
    h = 1.2 + 0.3 * np.exp(-((ii - 0.4 * nx) ** 2 + (jj - 0.5 * ny) ** 2) / (0.02 * nx * ny)) - z
    if dry:
        h[rng.random((ny, nx)) < 0.08] = 0.0
        h[:, : nx // 6] = rng.uniform(0, 2e-6, (ny, nx // 6))
    hu = 0.3 * h * rng.standard_normal((ny, nx))
    hv = 0.3 * h * rng.standard_normal((ny, nx))
    n = np.zeros((ny, nx)) if flat else rng.uniform(0.0, 0.04, (ny, nx))
    return h, hu, hv, z, n


def edge_faces(nx, ny):
    """{dir, i, j, side} in the output's row order; side 1 = D- (the box's
    cell is on the high side), side 0 = D+."""
    f = [(0, 0, j, 1) for j in range(ny)] + [(0, nx, j, 0) for j in range(ny)]
    f += [(1, i, 0, 1) for i in range(nx)] + [(1, i, ny, 0) for i in range(nx)]
    return f


def reeval(model, patch, faces, dtype=torch.float64):
    dev = torch.as_tensor(np.asarray(faces, np.int32), device="cuda")
    out = torch.zeros((len(faces), 3), dtype=dtype, device="cuda")
    L.check(L.load().hx_face_fluxes(context(), MODELS[model], C.byref(patch), len(faces),
                                    C.c_void_p(dev.data_ptr()), 1, 1.0, 1.0, 0,
                                    C.c_void_p(out.data_ptr()), _stream()), "hx_face_fluxes")
    return out


@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
@pytest.mark.parametrize("shape,bc,dry", [
    ((37, 130), ((0, 0), (0, 0)), False),
    ((61, 190), ((1, 0), (0, 1)), True),      # 3-4 strips across, a partly dry level
])
def test_edge_fluxes_match_reevaluation_and_oracle(gpu, arith, model, shape, bc, dry):
    ny, nx = shape
    lo, hi = bc
    dx, dy = 0.7, 0.5
    st = state(np.random.default_rng(31), ny, nx, dry)
    dt = 0.4 * O.level_dt(*st[:4], dx, dy)[0]
    set_arith(arith)
    try:
        print(f"edge fluxes: {model} {arith} {shape} bc {bc} dry {dry}", flush=True)
        lv, ref = Level(*st, dx, dy, lo, hi), Level(*st, dx, dy, lo, hi)
        ef = torch.full((2 * ny + 2 * nx, 3), np.nan, dtype=torch.float64, device="cuda")
        sv = SolverVariant(model)
        sv.compute_fluxes(lv, dt, edge_fluxes=ef)
        sv.compute_fluxes(ref, dt)
        again = reeval(model, lv.patch(lv.U_old), edge_faces(nx, ny))
        torch.cuda.synchronize()
        assert torch.equal(lv.U_new, ref.U_new)
        got = ef.cpu().numpy()
        assert np.isfinite(got).all()
        assert np.array_equal(got, again.cpu().numpy())
        W = A.working_state(*st, lo, hi)
        want = np.asarray([A.face_flux(MODELS[model], W, f, dx, dy) for f in edge_faces(nx, ny)])
        # relative to the flux vector's scale: on reflective sides the wall
        # mass flux is exactly 0 in the oracle and round-off in the fast build
        scale = np.abs(want).max()
        for k in range(3):
            err = np.abs(got[:, k] - want[:, k]).max() / scale
            assert err <= 1e-12, (k, err)
    finally:
        set_arith("exact")


@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
def test_edge_fluxes_fp32(gpu, arith, model):
    """The single-precision build's variant: edge fluxes equal hx_face_fluxes'
    FP32 re-evaluation, the state equals hx_advance_level's, and the fluxes
    lie within 2e-6 of the FP64 oracle relative to the largest component
    (measured 2.05e-7 to 2.19e-7 in both builds and solvers)."""
    ny, nx = 61, 190
    lo, hi = (1, 0), (0, 1)
    dx, dy = 0.7, 0.5
    st = state(np.random.default_rng(34), ny, nx, True)
    dt = 0.4 * O.level_dt(*st[:4], dx, dy)[0]
    set_arith(arith)
    try:
        print(f"edge fluxes FP32: {model} {arith}", flush=True)
        lv = Level(*st, dx, dy, lo, hi, dtype=torch.float32)
        ref = Level(*st, dx, dy, lo, hi, dtype=torch.float32)
        ef = torch.full((2 * ny + 2 * nx, 3), np.nan, dtype=torch.float32, device="cuda")
        sv = SolverVariant(model)
        sv.compute_fluxes(lv, dt, edge_fluxes=ef)
        sv.compute_fluxes(ref, dt)
        again = reeval(model, lv.patch(lv.U_old), edge_faces(nx, ny), torch.float32)
        torch.cuda.synchronize()
        assert torch.equal(lv.U_new, ref.U_new)
        got = ef.cpu().numpy()
        assert np.isfinite(got).all()
        assert np.array_equal(got, again.cpu().numpy())
        W = A.working_state(*st, lo, hi)
        want = np.asarray([A.face_flux(MODELS[model], W, f, dx, dy) for f in edge_faces(nx, ny)])
        err = np.abs(got.astype(np.float64) - want).max() / np.abs(want).max()
        print(f"  rel err vs FP64 oracle {err:.2e}", flush=True)
        assert err <= 2e-6, err
    finally:
        set_arith("exact")


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
def test_edge_fluxes_with_stored_ghosts(gpu, model):
    """A patch whose four sides read caller-filled ghost layers (an AMR fine
    box; arbitrary ghost values): the edge fluxes equal hx_face_fluxes on the
    same patch, and the state equals hx_advance_level's."""
    ny, nx, ng = 45, 140, 2
    rng = np.random.default_rng(32)
    big = state(rng, ny + 2 * ng, nx + 2 * ng)
    h, hu, hv, z, n = big
    fields = [torch.tensor(a, device="cuda").contiguous() for a in (h, hu, hv, h + z, z, n)]
    pitch = nx + 2 * ng
    off = (ng * pitch + ng) * 8
    p = L.HxPatch()
    p.nx, p.ny, p.ng, p.dtype, p.pitch = nx, ny, ng, L.HX_F64, pitch
    p.h, p.hu, p.hv, p.eta, p.z, p.n = (t.data_ptr() + off for t in fields)
    p.dx, p.dy = 0.6, 0.6
    p.bc_lo[0] = p.bc_lo[1] = p.bc_hi[0] = p.bc_hi[1] = L.HX_BC_GHOST
    dt = 0.02
    outs = [torch.zeros((3, ny, nx), dtype=torch.float64, device="cuda") for _ in range(2)]
    ef = torch.full((2 * ny + 2 * nx, 3), np.nan, dtype=torch.float64, device="cuda")
    lib, s = L.load(), MODELS[model]
    print(f"edge fluxes, stored ghosts: {model}", flush=True)
    L.check(lib.hx_advance_level_fluxes(context(), s, C.byref(p), dt, outs[0][0].data_ptr(),
                                        outs[0][1].data_ptr(), outs[0][2].data_ptr(), nx,
                                        ef.data_ptr(), _stream()), "hx_advance_level_fluxes")
    L.check(lib.hx_advance_level(context(), s, C.byref(p), dt, outs[1][0].data_ptr(),
                                 outs[1][1].data_ptr(), outs[1][2].data_ptr(), nx, _stream()),
            "hx_advance_level")
    again = reeval(model, p, edge_faces(nx, ny))
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    got = ef.cpu().numpy()
    assert np.isfinite(got).all()
    assert np.array_equal(got, again.cpu().numpy())


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
def test_box_mass_change_equals_net_edge_flux(gpu, model):
    """Flat, frictionless, wet 700 x 900 level, fast build: with no source
    share the interior faces telescope, so sum(h_new - h_old) dx dy equals
    -dt (sum_j (F_xhigh - F_xlow) dy + sum_i (F_yhigh - F_ylow) dx) for the
    mass component, up to the rounding of the cell updates (bound 1e-16 of
    sum |dh| dA, about 1e-13 absolute)."""
    ny, nx, dx, dy = 700, 900, 0.3, 0.4
    st = state(np.random.default_rng(33), ny, nx, flat=True)
    dt = 0.4 * O.level_dt(*st[:4], dx, dy)[0]
    set_arith("fast")
    try:
        lv = Level(*st, dx, dy, (1, 0), (0, 1))
        ef = torch.empty((2 * ny + 2 * nx, 3), dtype=torch.float64, device="cuda")
        SolverVariant(model).compute_fluxes(lv, dt, edge_fluxes=ef)
        torch.cuda.synchronize()
        dh = (lv.U_new[0] - lv.U_old[0]).cpu().numpy()
        f = ef[:, 0].cpu().numpy()
    finally:
        set_arith("exact")
    lhs = math.fsum(dh.ravel().tolist()) * dx * dy
    net = (math.fsum(f[ny:2 * ny].tolist()) - math.fsum(f[:ny].tolist())) * dy + \
          (math.fsum(f[2 * ny + nx:].tolist()) - math.fsum(f[2 * ny:2 * ny + nx].tolist())) * dx
    rhs = -dt * net
    scale = math.fsum(np.abs(dh).ravel().tolist()) * dx * dy
    print(f"{model}: mass change {lhs:.15e}, net edge flux {rhs:.15e}, sum|dh| dA {scale:.3e}",
          flush=True)
    assert abs(rhs) > 1e-6 * scale            # the edges carry flux
    # measured 7.7e-16 (Roe) and 4.2e-15 (hydrostatic) on sum|dh| dA = 964 and 1170
    assert abs(lhs - rhs) <= 1e-16 * scale
