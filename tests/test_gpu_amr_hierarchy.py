"""GPU tests of the multi-level, multi-box hierarchy (hydroexa_amd.amr.AmrHierarchy).

* Parity against the oracle's restatement (oracle/amr.py AmrOracle) on fixed
  boxes -- boxes touching the domain boundary, adjacent boxes, three levels,
  subcycled and lockstep: FP64 within 1e-12 relative L-infinity per level.
* Composite mass conserved to round-off in a closed box while the grids are
  regridded from gradient tags every two steps.
* The reference's AMR dam break (tests/dambreak: 10 x 1.25 domain, 512 x 64
  base, one level at ratio 2, blocking factor 32, Roe, CFL 0.5, tags at
  |grad| 0.0075): the L1 depth error against the SWASHES solution at t = 15,
  next to the reference's published values.
"""
import numpy as np
import pytest
import torch
from scipy.optimize import brentq

from oracle import amr as A
from oracle import oracle as O
from hydroexa_amd import _lib as L
from hydroexa_amd import set_arith
from hydroexa_amd.amr import AmrHierarchy

pytestmark = pytest.mark.gpu

MODELS = {"Roe": O.ROE, "HydrostaticRoe": O.HYDROSTATIC_ROE}
G = A.G


def rel_linf(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def field(ny, nx, seed=1, flat_ring=True):
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(seed)
    z = 0.05 * np.sin(0.5 * ii) * np.cos(0.4 * jj)
    if flat_ring:
        z[:2, :] = z[-2:, :] = 0.0
        z[:, :2] = z[:, -2:] = 0.0
    h = 1.0 + 0.3 * np.exp(-((ii - 0.4 * nx) ** 2 + (jj - 0.5 * ny) ** 2) / (0.05 * nx * ny)) - z
    hu = 0.1 * h * rng.standard_normal((ny, nx))
    hv = 0.1 * h * rng.standard_normal((ny, nx))
    n = rng.uniform(0.0, 0.03, (ny, nx))
    return h, hu, hv, z, n


CASES = {
    # model, base (ny, nx), bc_lo, bc_hi, fixed boxes per level
    "roe_edges": ("Roe", (24, 40), (1, 0), (0, 1),
                  {1: [(0, 8, 16, 24), (16, 8, 24, 16), (48, 0, 32, 16), (56, 32, 24, 16)]}),
    "hydro_3lev": ("HydrostaticRoe", (20, 24), (0, 0), (0, 0),
                   {1: [(6, 6, 24, 20), (30, 6, 10, 16)], 2: [(18, 18, 32, 18), (50, 18, 16, 12)]}),
    "roe_3lev": ("Roe", (20, 24), (1, 1), (1, 1),
                 {1: [(8, 4, 32, 24)], 2: [(20, 14, 40, 28)]}),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("subcycle", [True, False])
def test_hierarchy_parity_fixed_boxes(gpu, case, subcycle):
    model, (ny, nx), lo, hi, boxes = CASES[case]
    st = field(ny, nx, seed=7)
    H = AmrHierarchy(model, *st, 1.0, 1.0, lo, hi, max_level=max(boxes), boxes=boxes,
                     subcycle=subcycle)
    R = A.AmrOracle(MODELS[model], *st, 1.0, 1.0, lo, hi, boxes, subcycle=subcycle)
    for it in range(4):
        dts = H.compute_dt(0.45)
        want = R.compute_dt(0.45)
        assert all(abs(a - b) <= 1e-12 * b for a, b in zip(dts, want)), (dts, want)
        print(f"hierarchy {case} subcycle {subcycle} step {it} dts {dts}", flush=True)
        H.step(dts=dts)
        R.step(dts=dts)
        torch.cuda.synchronize()
        for l, (lv, ro) in enumerate(zip(H.levels, R.levels)):
            m = ro.interior(ro.mask)
            got = lv.interior(lv.U_new).double().cpu().numpy()[:, m]
            ref = ro.interior(ro.U)[:, m]
            for k in range(3):
                err = rel_linf(got[k], ref[k])
                assert err <= 1e-12, (case, it, l, k, err)


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
def test_hierarchy_regrid_conserves_mass(gpu, model):
    ny, nx = 64, 64
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    z = np.zeros((ny, nx))
    h = 1.0 + 0.5 * (np.hypot(ii - 24.0, jj - 30.0) < 8)
    zero = np.zeros((ny, nx))
    H = AmrHierarchy(model, h, zero, zero, z, zero, 1.0, 1.0, (0, 0), (0, 0), max_level=2,
                     regrid_int=2, blocking_factor=8, max_grid_size=32,
                     tag_thresh=[(0.02, 0.0, 0.0), (0.03, 0.0, 0.0)])
    for l in range(H.finest - 1, -1, -1):
        H._average_down(l)
    m0 = H.composite_mass()
    seen = set()
    for _ in range(16):
        H.step(0.45)
        seen.add(tuple(tuple(lv.boxes) for lv in H.levels[1:]))
    m1 = H.composite_mass()
    print(f"{model}: {len(seen)} distinct grid sets, mass drift {abs(m1 - m0) / m0:.2e}", flush=True)
    assert H.finest == 2
    assert len(seen) > 1
    assert abs(m1 - m0) / m0 < 1e-13


GRAV = 9.81


def swashes(t, x, hl, hr, x0):
    """SWASHES dam break on a wet bed (tests/dambreak/analytical.py:25-65)."""
    cl, cr = np.sqrt(GRAV * hl), np.sqrt(GRAV * hr)
    f = lambda cm: -8 * cr ** 2 * cm ** 2 * (cl - cm) ** 2 + (cm ** 2 - cr ** 2) ** 2 * (cm ** 2 + cr ** 2)
    cm = brentq(f, cr * (1 + 1e-9), cl)
    xa, xb = x0 - t * cl, x0 + t * (2 * cl - 3 * cm)
    xc = x0 + t * (2 * cm ** 2 * (cl - cm)) / (cm ** 2 - cr ** 2)
    return np.where(x <= xa, hl, np.where(x <= xb, (4 / (9 * GRAV)) * (cl - (x - x0) / (2 * t)) ** 2,
                                          np.where(x <= xc, cm ** 2 / GRAV, hr)))


def l1_weighted(x, e):
    """tests/dambreak/analyze.py:160-195: trapezoidal cell widths."""
    w = np.zeros_like(x)
    w[1:-1] = 0.5 * (x[2:] - x[:-2])
    w[0], w[-1] = 0.5 * (x[1] - x[0]), 0.5 * (x[-1] - x[-2])
    return float(np.sum(w * np.abs(e)) / w.sum())


def dambreak_run(max_level, regrid_int, t_end=15.0, nx=512, ny=64):
    L, hl, hr, x0 = 10.0, 0.005, 0.001, 5.0
    dx = L / nx

    def state(nxl, nyl):
        x = (np.arange(nxl) + 0.5) * (L / nxl)
        h = np.where(x < x0, hl, hr)[None, :] * np.ones((nyl, 1))
        z = np.zeros_like(h)
        return h, z.copy(), z.copy(), z.copy(), z.copy()

    H = AmrHierarchy("Roe", *state(nx, ny), dx, dx, (1, 0), (1, 0), max_level=max_level,
                     regrid_int=regrid_int, blocking_factor=32, max_grid_size=128,
                     n_error_buf=1, tag_thresh=[(0.0075, 0.0075, 0.0075)],
                     static=lambda lev: state(nx << lev, ny << lev))
    H.run_native(t_end, 0.5)
    x, h = H.centerline(0.5)
    return H, x, h, l1_weighted(x, h - swashes(H.time, x, hl, hr, x0))


def test_reference_amr_dambreak_accuracy(gpu):
    """The reference publishes L1 (t = 15) = 1.85e-5 on 512 x 64, 1.07e-5 on
    1024 x 128, 1.30e-5 with one AMR level regridded every 50 steps and
    1.06e-5 every 5 steps (tests/dambreak/README.md:272-284, 304-315).  Every
    run must land on the published value to its printed precision (the AMR
    runs use this library's Berger-Rigoutsos grids, not AMReX's, and still
    do: measured 1.851e-5, 1.068e-5, 1.295e-5, 1.058e-5), and meet the
    reference's acceptance rules (tests/dambreak/README.md, metrics section):
    better than the unrefined run, and with regrid 5 within 10 % of the
    uniform fine run."""
    set_arith("exact")
    _, _, _, base = dambreak_run(0, 0)
    _, _, _, fine = dambreak_run(0, 0, nx=1024, ny=128)
    H50, *_, amr50 = dambreak_run(1, 50)
    H5, *_, amr5 = dambreak_run(1, 5)
    print(f"L1 t=15: base {base:.3e} fine {fine:.3e} amr(regrid 50) {amr50:.3e} "
          f"amr(regrid 5) {amr5:.3e}; level-1 boxes {H5.levels[1].boxes if H5.finest else []}",
          flush=True)
    assert float(f"{base:.2e}") == pytest.approx(1.85e-5, abs=0.011e-5)
    assert float(f"{fine:.2e}") == pytest.approx(1.07e-5, abs=0.011e-5)
    assert float(f"{amr50:.2e}") == pytest.approx(1.30e-5, abs=0.011e-5)
    assert float(f"{amr5:.2e}") == pytest.approx(1.06e-5, abs=0.011e-5)
    assert amr50 < base and amr5 < base
    assert amr5 <= 1.1 * fine
    assert amr5 < amr50


def _same_state(H1, H2):
    assert H1.finest == H2.finest
    for a, b in zip(H1.levels, H2.levels):
        assert a.boxes == b.boxes
        m = a.interior(a.mask).bool()
        assert torch.equal(m, b.interior(b.mask).bool())
        assert torch.equal(a.interior(a.U_new)[:, m], b.interior(b.U_new)[:, m])
        assert a.t_new == b.t_new


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("subcycle", [True, False])
def test_native_run_equals_python_steps_fixed(gpu, case, subcycle):
    model, (ny, nx), lo, hi, boxes = CASES[case]
    st = field(ny, nx, seed=9)
    mk = lambda: AmrHierarchy(model, *st, 1.0, 1.0, lo, hi, max_level=max(boxes), boxes=boxes,
                              subcycle=subcycle)
    H1, H2 = mk(), mk()
    print(f"native vs python: {case} subcycle {subcycle}", flush=True)
    hist = H1.run_native(1e30, 0.45, max_steps=5)
    dts = [H2.step(0.45) for _ in range(5)]
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hist, dts)
    assert H1.time == H2.time
    _same_state(H1, H2)


@pytest.mark.parametrize("model", ["Roe", "HydrostaticRoe"])
def test_native_run_equals_python_steps_with_regrid(gpu, model):
    ny, nx = 64, 64
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    h = 1.0 + 0.5 * (np.hypot(ii - 24.0, jj - 30.0) < 8)
    zero = np.zeros((ny, nx))
    mk = lambda: AmrHierarchy(model, h, zero, zero, zero, zero, 1.0, 1.0, (0, 0), (0, 0),
                              max_level=2, regrid_int=2, blocking_factor=8, max_grid_size=32,
                              tag_thresh=[(0.02, 0.0, 0.0), (0.03, 0.0, 0.0)])
    H1, H2 = mk(), mk()
    H1.run_native(1e30, 0.45, max_steps=12)
    for _ in range(12):
        H2.step(0.45)
    torch.cuda.synchronize()
    _same_state(H1, H2)
    assert H1.istep == H2.istep and H1.last_regrid == H2.last_regrid


def test_native_run_reports_dt_collapse(gpu):
    st = list(field(16, 16, seed=2))
    st[1] = st[1].copy()
    st[1][5, 5] = 1e13
    H = AmrHierarchy("Roe", *st, 1.0, 1.0, (0, 0), (0, 0), max_level=1,
                     boxes={1: [(8, 8, 12, 12)]})
    before = H.state(0).clone()
    with pytest.raises(L.DtCollapse):
        H.run_native(1e30, 0.45, max_steps=3)
    assert H.time == 0.0 and torch.equal(H.state(0), before)


@pytest.mark.parametrize("nbuf,b", [(0, 4), (1, 8), (2, 4), (3, 16)])
def test_tag_blocks_match_host_restatement(gpu, nbuf, b):
    from hydroexa_amd import regrid as RG
    from hydroexa_amd.solver import context, _stream
    rng = np.random.default_rng(nbuf * 10 + b)
    ny, nx = 96, 128
    tags = rng.random((ny, nx)) < 0.02
    valid = RG.covers([(0, 0, 64, 96), (64, 32, 48, 48)], ny, nx)
    vpitch = nx + 4
    vpad = np.zeros((ny + 4, vpitch), np.uint8)
    vpad[2:-2, 2:-2] = valid
    td = torch.as_tensor(tags.astype(np.uint8), device="cuda")
    vd = torch.as_tensor(vpad, device="cuda")
    out = torch.empty((ny // b, nx // b), dtype=torch.uint8, device="cuda")
    L.check(L.load().hx_tag_blocks(context(), td.data_ptr(), nx, vd.data_ptr() + 2 * vpitch + 2,
                                   vpitch, nx, ny, nbuf, b, out.data_ptr(), _stream()),
            "hx_tag_blocks")
    want = RG.coarsen_tags(RG.grow_tags(tags & valid, nbuf) & valid, b)
    np.testing.assert_array_equal(out.bool().cpu().numpy(), want)


def test_hierarchy_conservation_and_plotfile(gpu, tmp_path):
    from hydroexa_amd.plotfile import read_native_plotfile
    model, (ny, nx), lo, hi, boxes = CASES["hydro_3lev"]
    st = field(ny, nx, seed=4)
    H = AmrHierarchy(model, *st, 1.0, 1.0, lo, hi, max_level=2, boxes=boxes)
    for l in (1, 0):
        H._average_down(l)
    c0 = H.conservation()
    # per level: the sum over the valid cells (NODATA-free, h > 0 here)
    for l, lv in enumerate(H.levels):
        m = lv.interior(lv.mask).bool()
        hs = lv.interior(lv.U_new)[0][m].double().sum().item() * lv.dx * lv.dy
        assert abs(c0["levels"][l][0] - hs) <= 1e-13 * hs
    assert abs(c0["mass"] - sum(p[0] for p in c0["levels"])) <= 1e-15 * c0["mass"]
    H.run_native(1e30, 0.45, max_steps=3)
    c1 = H.conservation()
    assert abs(c1["composite_mass"] - c0["composite_mass"]) <= 1e-13 * c0["composite_mass"]
    name = H.write_plotfile(str(tmp_path / "plt"), 3)
    pf = read_native_plotfile(name)
    assert pf["finest_level"] == 2 and pf["ref_ratio"] == [2, 2]
    for l, lv in enumerate(H.levels):
        got = pf["levels"][l]
        assert sorted(got["boxes"]) == sorted((i0, j0, i0 + w - 1, j0 + h - 1)
                                             for (i0, j0, w, h) in lv.boxes)
        U = lv.interior(lv.U_new).double().cpu().numpy()
        for fab in got["fabs"]:
            i0, j0, i1, j1 = fab["box"]
            np.testing.assert_array_equal(fab["data"][:3], U[:, j0:j1 + 1, i0:i1 + 1])


# ---------------------------------------------------------------------------
# Config C3: the circular dam break on a 3-level hierarchy with regridding
# ---------------------------------------------------------------------------
from hydroexa_amd import workloads as W  # noqa: E402


def c3_hierarchy(w, **kw):
    args = dict(max_level=2, regrid_int=4, blocking_factor=16, max_grid_size=128,
                tag_thresh=[(0.1, 1e300, 1e300)])
    args.update(kw)
    return AmrHierarchy(w.solver, w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi, **args)


def record_regrids(H):
    """Record every regrid of H after it happened: (level, time, the new box
    lists of the levels above it)."""
    events = []
    make = H._make_grids

    def rec(lbase, time, from_scratch=False):
        make(lbase, time, from_scratch)
        events.append((lbase, time, {l: list(H.levels[l].boxes) for l in range(lbase + 1, H.finest + 1)}))
    H._make_grids = rec
    return events


@pytest.mark.parametrize("n,steps", [(256, 9), (512, 5)])
def test_c3_cylinder_replays_through_the_oracle(gpu, n, steps):
    """Config C3's cylinder dam break (workloads.circular_dam_break: depth 2.0
    inside r = 0.2, 0.5 outside, bed 0.05 U[0,1), Manning 0.03, hydrostatic
    Roe, transmissive sides, tags |grad h| > 0.1, regrid every 4 steps of a
    level, 2 refined levels) against oracle/amr.py's AmrOracle replaying the
    hierarchy's own grids: level 1 regrids level 2 every 2 coarse steps and
    level 0 regrids both every 4.  At a 256^2 base 9 coarse steps cross the
    regrids at steps 2, 4, 6 and 8 (two level-0 intervals); at the config's
    512^2 base 5 steps cross those at 2 and 4.  Every level within 1e-12
    relative L-infinity after every step, the same dts, every event replayed
    where the cadence puts it."""
    import time as _t
    set_arith("exact")
    w = W.circular_dam_break(n)
    H = c3_hierarchy(w)
    boxes0 = {l: list(H.levels[l].boxes) for l in range(1, H.finest + 1)}
    events = record_regrids(H)
    R = A.AmrOracle(MODELS[w.solver], w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi,
                    boxes0, regrid_int=4, max_level=2, events=events)
    print(f"C3 replay n={n}: initial boxes per level {[len(b) for b in boxes0.values()]}", flush=True)
    seen = []
    for it in range(steps):
        t0 = _t.time()
        dts = H.compute_dt(w.cfl)
        want = R.compute_dt(w.cfl)
        assert all(abs(a - b) <= 1e-12 * b for a, b in zip(dts, want)), (it, dts, want)
        nev = len(events)
        H.step(dts=dts)
        torch.cuda.synchronize()
        seen += [(e[0], it) for e in events[nev:]] if len(events) > nev else []
        R.step(dts=dts)
        assert not events, f"step {it}: {len(events)} regrid event(s) not replayed"
        assert R.finest == H.finest
        for l, (lv, ro) in enumerate(zip(H.levels, R.levels)):
            assert lv.boxes == ro.boxes, (it, l)
            m = ro.interior(ro.mask)
            got = lv.interior(lv.U_new).double().cpu().numpy()[:, m]
            ref = ro.interior(ro.U)[:, m]
            for k in range(3):
                err = rel_linf(got[k], ref[k])
                assert err <= 1e-12, (n, it, l, k, err)
        print(f"  step {it}: boxes {[len(lv.boxes) for lv in H.levels]} "
              f"cells {[int(lv.interior(lv.mask).sum().item()) for lv in H.levels]} "
              f"oracle+check {_t.time() - t0:.1f} s", flush=True)
    regrid_levels = [e[0] for e in seen]
    assert 0 in regrid_levels and 1 in regrid_levels, seen


def test_c3_cylinder_full_length_properties(gpu):
    """Config C3 at full size and length: 512^2 base, 2 refined levels,
    1000 coarse steps through hx_amr_run.  In a closed box (reflective sides,
    the bed flat on the two cells along the walls so that the walls hold the
    water: the reference's wall leak on a sloping bed moves mass through them)
    the composite mass stays to round-off; every level stays finite with
    h >= 0; the grids are regridded as the front spreads."""
    set_arith("fast")
    try:
        _c3_full_length()
    finally:
        set_arith("exact")


def _c3_full_length():
    w = W.circular_dam_break(512, bc=(0, 0))
    w.z[:2, :] = w.z[-2:, :] = 0.0
    w.z[:, :2] = w.z[:, -2:] = 0.0
    H = c3_hierarchy(w)
    m0 = H.composite_mass()
    seen = {tuple(len(lv.boxes) for lv in H.levels)}
    prev = H.state(0).clone()
    fine_mass = lambda: H.state(2)[0].double().sum().item() * H.levels[2].dx * H.levels[2].dy
    done = 0
    while done < 1000:
        H.run_native(1e30, w.cfl, max_steps=100)
        done += 100
        seen.add(tuple(len(lv.boxes) for lv in H.levels))
        moved = (H.state(0) - prev).abs().max().item()
        prev = H.state(0).clone()
        print(f"  level 0 max change over the last 100 steps {moved:.3e}; composite mass "
              f"{H.composite_mass()!r} (start {m0!r}); finest-level mass {fine_mass()!r}", flush=True)
        assert moved > 0.0
        for l, lv in enumerate(H.levels):
            U = lv.interior(lv.U_new)
            m = lv.interior(lv.mask).bool()
            assert torch.isfinite(U[:, m]).all(), (done, l)
            assert (U[0][m] >= 0).all(), (done, l)
        drift = abs(H.composite_mass() - m0) / m0
        print(f"C3 full length: {done} steps, t = {H.time:.5f}, boxes {[len(lv.boxes) for lv in H.levels]}, "
              f"mass drift {drift:.2e}", flush=True)
        assert drift < 1e-12
    assert H.finest == 2 and len(seen) > 1
