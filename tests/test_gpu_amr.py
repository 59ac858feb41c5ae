"""GPU parity of the two-level coupling (one fine box, ratio 2) against the
oracle (oracle/amr.py), through the C-ABI entry points hx_fill_fine_ghosts,
hx_face_fluxes, hx_reflux, hx_average_down_state and the TwoLevel driver.

Bars: FP64 within 1e-12 relative L-infinity per field (relative to the
field's max |value|); the fill, reflux and average-down, which are
elementwise arithmetic without transcendentals, bit-exact.  Size-independent
properties at a large size: composite mass conserved to round-off in a
closed box, lake at rest kept.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import amr as A
from oracle import oracle as O
from hydroexa_amd import Level, TwoLevel, context
from hydroexa_amd import _lib as L
from hydroexa_amd.amr import FineBox

pytestmark = pytest.mark.gpu

G = A.G
MODELS = {"Roe": O.ROE, "HydrostaticRoe": O.HYDROSTATIC_ROE}


def rel_linf(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def wet_state(rng, ny, nx, flat_ring=True):
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    z = 0.2 * np.sin(0.37 * ii + 0.2) * np.cos(0.23 * jj) + 0.02 * rng.random((ny, nx))
    if flat_ring:
        z[:2, :] = z[-2:, :] = 0.0
        z[:, :2] = z[:, -2:] = 0.0
    h = 1.2 + 0.3 * np.exp(-((ii - 0.4 * nx) ** 2 + (jj - 0.5 * ny) ** 2) / (0.02 * nx * ny)) - z
    hu = 0.2 * h * rng.standard_normal((ny, nx))
    hv = 0.2 * h * rng.standard_normal((ny, nx))
    n = rng.uniform(0.0, 0.04, (ny, nx))
    return h, hu, hv, z, n


def _np(t):
    return t.double().cpu().numpy()


@pytest.mark.parametrize("bc", [((0, 0), (0, 0)), ((1, 0), (0, 1))])
@pytest.mark.parametrize("alpha", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("whole", [0, 1])
def test_fill_fine_ghosts_bit_exact(gpu, bc, alpha, whole):
    rng = np.random.default_rng(11)
    ny, nx = 20, 24
    box = (1, 2, 7, 5)                  # touches the one-cell margin at x-low
    lo, hi = bc
    so, sn = wet_state(rng, ny, nx), wet_state(rng, ny, nx)
    sn = (sn[0], sn[1], sn[2], so[3], so[4])
    co, cn = Level(*so, 1.0, 1.0, lo, hi), Level(*sn, 1.0, 1.0, lo, hi)
    zf, nf = A.fine_terrain(so[3], so[4], box)
    fb = FineBox(2 * box[2], 2 * box[3], 0.5, 0.5, torch.as_tensor(np.stack([zf, nf])), torch.float64)
    init = rng.standard_normal(fb.U_new.shape)
    fb.U_new.copy_(torch.as_tensor(init))
    print(f"fill: bc {bc} alpha {alpha} whole {whole}", flush=True)
    L.check(L.load().hx_fill_fine_ghosts(context(), C.byref(co.patch()), C.byref(cn.patch()), alpha,
                                         2, box[0], box[1], C.byref(fb.patch()), whole,
                                         fb._ptr(fb.U_new, 0), fb._ptr(fb.U_new, 1),
                                         fb._ptr(fb.U_new, 2), _stream()), "fill")
    torch.cuda.synchronize()
    want = init.copy()
    mk = lambda s: {"h": s[0], "hu": s[1], "hv": s[2], "z": s[3], "n": s[4]}
    A.fill_fine(mk(so), mk(sn), alpha, box, lo, hi, zf, want, whole=bool(whole))
    np.testing.assert_array_equal(_np(fb.U_new), want)


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
@pytest.mark.parametrize("bc", [((0, 0), (0, 0)), ((1, 1), (1, 1))])
def test_face_fluxes_register(gpu, model, bc):
    """CrseInit (pair 1, COPY) then FineAdd-style accumulation (pair 2) over
    the faces around a box of a single level, against the oracle."""
    rng = np.random.default_rng(12)
    ny, nx, dx, dy = 18, 22, 0.7, 0.5
    lo, hi = bc
    st = wet_state(rng, ny, nx)
    lv = Level(*st, dx, dy, lo, hi)
    box = (1, 3, 6, 4)
    cx, cy, _, _ = A.box_faces(box)
    extra = [(0, 0, 5, 0), (0, nx, 7, 1), (1, 4, 0, 0), (1, 9, ny, 1), (0, 11, 9, 1), (1, 3, 12, 0)]
    faces = cx + cy + extra
    dev = torch.as_tensor(np.asarray(faces, np.int32).reshape(-1, 4), device="cuda")
    out = torch.zeros((len(faces), 3), dtype=torch.float64, device="cuda")
    lib = L.load()
    s = MODELS[model]
    print(f"face fluxes: {model} bc {bc}", flush=True)
    L.check(lib.hx_face_fluxes(context(), s, C.byref(lv.patch()), len(faces),
                               C.c_void_p(dev.data_ptr()), 1, -0.3, 0.7, 0,
                               C.c_void_p(out.data_ptr()), _stream()), "face_fluxes copy")
    W = A.working_state(*st, lo, hi)
    sc = lambda f, a, b: a if f[0] == 0 else b
    want = np.asarray([sc(f, -0.3, 0.7) * A.face_flux(s, W, f, dx, dy) for f in faces])
    torch.cuda.synchronize()
    got = _np(out)
    for k in range(3):
        assert rel_linf(got[:, k], want[:, k]) <= 1e-12, (k, rel_linf(got[:, k], want[:, k]))
    # accumulate pairs of faces onto the copied values
    npair = len(faces) // 2
    L.check(lib.hx_face_fluxes(context(), s, C.byref(lv.patch()), 2 * npair,
                               C.c_void_p(dev.data_ptr()), 2, 0.125, -0.25, 1,
                               C.c_void_p(out.data_ptr()), _stream()), "face_fluxes accumulate")
    torch.cuda.synchronize()
    acc = want.copy()
    for q in range(npair):
        for f in faces[2 * q:2 * q + 2]:
            acc[q] = acc[q] + sc(f, 0.125, -0.25) * A.face_flux(s, W, f, dx, dy)
    got = _np(out)
    for k in range(3):
        assert rel_linf(got[:, k], acc[:, k]) <= 1e-12


def test_face_fluxes_outside_level_are_nan(gpu):
    """A face whose stencil would leave the level yields NaN, not a read."""
    lv = Level(*wet_state(np.random.default_rng(2), 8, 9), 1.0, 1.0)
    faces = [(0, 10, 1, 0), (0, 3, 8, 0), (1, 9, 2, 1), (1, 2, 9, 0), (2, 1, 1, 0), (0, -1, 0, 1),
             (0, 4, 4, 0)]
    dev = torch.as_tensor(np.asarray(faces, np.int32), device="cuda")
    out = torch.zeros((len(faces), 3), dtype=torch.float64, device="cuda")
    L.check(L.load().hx_face_fluxes(context(), O.ROE, C.byref(lv.patch()), len(faces),
                                    C.c_void_p(dev.data_ptr()), 1, 1.0, 1.0, 0,
                                    C.c_void_p(out.data_ptr()), _stream()), "face_fluxes")
    got = _np(out)
    assert np.isnan(got[:-1]).all()
    assert np.isfinite(got[-1]).all()


def test_reflux_and_average_down_bit_exact(gpu):
    rng = np.random.default_rng(13)
    ny, nx, dx, dy = 12, 15, 0.7, 0.3
    box = (2, 1, 5, 4)
    st = wet_state(rng, ny, nx)
    lv = Level(*st, dx, dy)
    reg = rng.standard_normal((2 * (box[2] + box[3]), 3))
    regd = torch.as_tensor(reg, device="cuda")
    lib = L.load()
    base = lv.U_new.data_ptr()
    plane = nx * ny * 8
    L.check(lib.hx_reflux(context(), C.byref(lv.patch()), base, base + plane, base + 2 * plane,
                          box[0], box[1], box[2], box[3], C.c_void_p(regd.data_ptr()), _stream()),
            "reflux")
    want = [st[k].copy() for k in range(3)]
    A.reflux(want, reg, box, dx, dy)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(lv.U_new), np.stack(want))
    # average down from a fine box with ghosts
    zf, nf = A.fine_terrain(st[3], st[4], box)
    fb = FineBox(2 * box[2], 2 * box[3], dx / 2, dy / 2, torch.as_tensor(np.stack([zf, nf])),
                 torch.float64)
    fine = rng.standard_normal(fb.U_new.shape)
    fb.U_new.copy_(torch.as_tensor(fine))
    L.check(lib.hx_average_down_state(context(), C.byref(fb.patch()), 2, box[0], box[1],
                                      C.byref(lv.patch()), base, base + plane, base + 2 * plane,
                                      _stream()), "average_down")
    A.average_down([fine[c][G:-G, G:-G] for c in range(3)], want, box)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(lv.U_new), np.stack(want))


def test_entry_points_reject_bad_boxes(gpu):
    lv = Level(*wet_state(np.random.default_rng(1), 10, 10), 1.0, 1.0)
    with pytest.raises(ValueError):
        TwoLevel("Roe", lv, (0, 2, 4, 4))          # no one-cell margin
    with pytest.raises(ValueError):
        TwoLevel("Roe", lv, (2, 2, 8, 4))          # past the high edge
    base = lv.U_new.data_ptr()
    reg = torch.zeros(64, dtype=torch.float64, device="cuda")
    rc = L.load().hx_reflux(context(), C.byref(lv.patch()), base, base, base, 0, 2, 4, 4,
                            C.c_void_p(reg.data_ptr()), _stream())
    assert rc == L.HX_E_BADARG


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
@pytest.mark.parametrize("subcycle", [True, False])
def test_two_level_step_parity(gpu, model, subcycle):
    rng = np.random.default_rng(21)
    ny, nx, dx, dy = 40, 48, 2.0, 2.0
    box = (9, 7, 14, 12)
    lo, hi = (0, 1), (1, 0)
    st = wet_state(rng, ny, nx)
    s = MODELS[model]
    lv = Level(*st, dx, dy, lo, hi)
    T = TwoLevel(model, lv, box, subcycle=subcycle)
    R = A.TwoLevelOracle(s, *st, dx, dy, lo, hi, box, subcycle=subcycle)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(T.fine.U_new), R.F["U"])
    for it in range(5):
        d0, d1 = A.level_dts(s, R.C, R.F, dx, dy)
        dts = A.dt_rule(0.45 * d0, 0.45 * d1, subcycle)
        g0, g1 = T.dt(0.45)
        assert abs(g0 - dts[0]) <= 1e-12 * dts[0] and abs(g1 - dts[1]) <= 1e-12 * dts[1]
        print(f"two-level {model} subcycle {subcycle} step {it} dt {dts}", flush=True)
        T.step(dts=dts)
        R.step(*dts)
        torch.cuda.synchronize()
        gc, gf = _np(lv.U_new), _np(T.fine.valid())
        wc = np.stack([R.C["h"], R.C["hu"], R.C["hv"]])
        wf = R.fine_valid()
        for k in range(3):
            assert rel_linf(gc[k], wc[k]) <= 1e-12, ("coarse", it, k, rel_linf(gc[k], wc[k]))
            assert rel_linf(gf[k], wf[k]) <= 1e-12, ("fine", it, k, rel_linf(gf[k], wf[k]))


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
@pytest.mark.parametrize("subcycle", [True, False])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_two_level_run_equals_stepping(gpu, model, subcycle, dtype):
    """The device-resident run (dt from the fused CFL reductions and the
    multi-level rule on the device) equals host-driven steps bit for bit."""
    rng = np.random.default_rng(41)
    ny, nx = 96, 120
    st = wet_state(rng, ny, nx)
    box = (21, 17, 40, 32)
    la = Level(*st, 2.0, 2.0, (0, 1), (1, 0), dtype=dtype)
    lb = Level(*st, 2.0, 2.0, (0, 1), (1, 0), dtype=dtype)
    A_ = TwoLevel(model, la, box, subcycle=subcycle)
    B_ = TwoLevel(model, lb, box, subcycle=subcycle)
    print(f"run vs step: {model} subcycle {subcycle} {dtype}", flush=True)
    hist = A_.run(7, 0.45)
    dts = [B_.step(0.45) for _ in range(7)]
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hist, np.asarray(dts))
    assert torch.equal(la.U_new, lb.U_new)
    assert torch.equal(A_.fine.valid(), B_.fine.valid())
    assert A_.steps == 7 and abs(A_.time - B_.time) <= 1e-12 * B_.time
    # a second call continues from the swapped buffers
    A_.run(3, 0.45)
    for _ in range(3):
        B_.step(0.45)
    assert torch.equal(la.U_new, lb.U_new)
    assert torch.equal(A_.fine.valid(), B_.fine.valid())


def test_two_level_run_longer_than_the_context_history(gpu):
    """hx_two_level_run does runs longer than the context's 4096-step history
    in chunks: one call of 4096 + 37 steps equals a call of 4096 then one of
    37, bit for bit, with the same dt history."""
    rng = np.random.default_rng(43)
    st = wet_state(rng, 32, 32)
    la = Level(*st, 2.0, 2.0, (0, 0), (0, 0))
    lb = Level(*st, 2.0, 2.0, (0, 0), (0, 0))
    A_ = TwoLevel("HydrostaticRoe", la, (10, 12, 8, 6))
    B_ = TwoLevel("HydrostaticRoe", lb, (10, 12, 8, 6))
    print("two-level chunked run: 4133 steps", flush=True)
    hist = A_.run(4096 + 37, 0.45)
    h1 = B_.run(4096, 0.45)
    h2 = B_.run(37, 0.45)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hist, np.concatenate([h1, h2]))
    assert torch.equal(la.U_new, lb.U_new)
    assert torch.equal(A_.fine.valid(), B_.fine.valid())


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
@pytest.mark.parametrize("subcycle", [True, False])
def test_two_level_graph_replay_equals_eager(gpu, model, subcycle):
    """hx_two_level_enqueue captured in a HIP graph (4 coupled steps) and
    replayed three times, with eager torch work and synchronisations between
    the replays, gives after every replay bit for bit what the same call gives
    eagerly on a twin hierarchy; the eager call equals hx_two_level_run."""
    rng = np.random.default_rng(47)
    st = wet_state(rng, 96, 120)
    box = (21, 17, 40, 32)
    lv = [Level(*st, 2.0, 2.0, (0, 1), (1, 0)) for _ in range(3)]
    G_, E_, R_ = (TwoLevel(model, v, box, subcycle=subcycle) for v in lv)
    dev = lv[0].U_new.device
    hg = torch.zeros(8, dtype=torch.float64, device=dev)
    he = torch.zeros(8, dtype=torch.float64, device=dev)
    sg = torch.full((1,), -1, dtype=torch.int32, device=dev)
    se = torch.full((1,), -1, dtype=torch.int32, device=dev)
    print(f"graph replay: {model} subcycle {subcycle}", flush=True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):             # warm-up outside the capture
        G_.enqueue(4, 0.45, hg, sg)
    torch.cuda.current_stream().wait_stream(side)
    E_.enqueue(4, 0.45, he, se)
    hist = R_.run(4, 0.45)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(he.cpu().numpy().reshape(4, 2), hist)
    assert torch.equal(lv[1].U_new, lv[2].U_new)
    assert torch.equal(E_.fine.valid(), R_.fine.valid())
    assert torch.equal(lv[0].U_new, lv[1].U_new) and int(se.item()) == 0

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        G_.enqueue(4, 0.45, hg, sg)
    torch.cuda.synchronize()
    assert torch.equal(lv[0].U_new, lv[1].U_new)    # capture runs nothing
    for r in range(3):
        sg.fill_(-1)
        g.replay()
        x = torch.randn(256, 256, device=dev, dtype=torch.float64)
        y = (x @ x).sum()                            # eager work between replays
        E_.enqueue(4, 0.45, he, se)
        torch.cuda.synchronize()
        assert torch.isfinite(y)
        assert int(sg.item()) == 0 and int(se.item()) == 0, f"replay {r}"
        assert torch.equal(hg, he), f"replay {r}: dt history"
        assert torch.equal(lv[0].U_new, lv[1].U_new), f"replay {r}: coarse state"
        assert torch.equal(G_.fine.valid(), E_.fine.valid()), f"replay {r}: fine state"
        assert torch.equal(G_.fine.U_new, E_.fine.U_new), f"replay {r}: fine storage"
    del g


def test_two_level_enqueue_rejects_odd_steps_and_reports_collapse(gpu):
    rng = np.random.default_rng(42)
    h, hu, hv, z, n = wet_state(rng, 32, 32)
    hu[3, 3] = 1e13
    lv = Level(h, hu, hv, z, n, 1.0, 1.0)
    T = TwoLevel("Roe", lv, (10, 10, 8, 8))
    dh = torch.zeros(8, dtype=torch.float64, device=lv.U_new.device)
    s = torch.full((1,), -1, dtype=torch.int32, device=lv.U_new.device)
    with pytest.raises(ValueError):
        T.enqueue(3, 0.45, dh, s)
    before = lv.U_new.clone()
    T.enqueue(4, 0.45, dh, s)
    torch.cuda.synchronize()
    assert int(s.item()) == 1                  # the first step's estimate collapsed
    assert torch.equal(lv.U_new, before)


def test_two_level_run_reports_dt_collapse(gpu):
    rng = np.random.default_rng(42)
    h, hu, hv, z, n = wet_state(rng, 32, 32)
    hu[3, 3] = 1e13                       # |u| ~ 1e13: estimate ~ 1e-13 < 1e-12
    lv = Level(h, hu, hv, z, n, 1.0, 1.0)
    T = TwoLevel("Roe", lv, (10, 10, 8, 8))
    before, fbefore = lv.U_new.clone(), T.fine.valid().clone()
    with pytest.raises(L.DtCollapse):
        T.run(5, 0.45)
    assert T.steps == 0
    assert torch.equal(lv.U_new, before)
    assert torch.equal(T.fine.valid(), fbefore)


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
def test_two_level_mass_conservation_large(gpu, model):
    """1024^2 coarse level with a 256^2-cell (coarse) box refined: composite
    mass in a closed box to round-off over 20 coarse steps."""
    rng = np.random.default_rng(31)
    n = 1024
    st = wet_state(rng, n, n)
    lv = Level(*st, 5.0, 5.0, (0, 0), (0, 0))
    T = TwoLevel(model, lv, (384, 400, 256, 256))
    # start from a composite state
    f = T.fine
    L.check(L.load().hx_average_down_state(context(), C.byref(f.patch()), 2, 384, 400,
                                           C.byref(lv.patch()), lv.U_new[0].data_ptr(),
                                           lv.U_new[1].data_ptr(), lv.U_new[2].data_ptr(), _stream()),
            "average_down")
    m0 = T.composite_mass()
    T.run(20, 0.45)
    m1 = T.composite_mass()
    print(f"{model}: composite mass drift {abs(m1 - m0) / m0:.3e} over 20 steps", flush=True)
    assert abs(m1 - m0) / m0 < 1e-13


def test_two_level_lake_at_rest(gpu):
    ny, nx = 64, 80
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    z = 0.5 * np.sin(0.31 * ii) * np.cos(0.19 * jj) + 0.1 * ((7 * ii + 3 * jj) % 5) / 5.0
    eta = 2.0
    zero = np.zeros((ny, nx))
    lv = Level(eta - z, zero, zero, z, np.full((ny, nx), 0.03), 1.0, 1.0, (0, 0), (0, 0))
    T = TwoLevel("HydrostaticRoe", lv, (20, 17, 24, 20))
    for _ in range(10):
        T.step(0.45)
    torch.cuda.synchronize()
    U = _np(lv.U_new)
    np.testing.assert_allclose(U[0] + z, eta, rtol=0, atol=1e-13)
    np.testing.assert_allclose(U[1:], 0.0, rtol=0, atol=1e-13)
    Uf = _np(T.fine.valid())
    zf = _np(T.fine.terrain[0, G:-G, G:-G])
    np.testing.assert_allclose(Uf[0] + zf, eta, rtol=0, atol=1e-13)
    np.testing.assert_allclose(Uf[1:], 0.0, rtol=0, atol=1e-13)
