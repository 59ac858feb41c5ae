"""Pins of the two-level oracle (oracle/amr.py) on the CPU.

Each check is a property the mathematics fixes, chosen so that a plausible
mistake fails it:

* the padded (ghost-fed) step equals the pinned single-level step when its
  ghosts hold the physical halo (wrong ghost offset or stored-vs-rebuilt
  depth fails);
* the side-selected effective fluxes reproduce the level update
  u - dt/dx (D+ right - D- left) - dt/dy (...) (a swapped side fails);
* conservative linear interpolation: the four children average to the
  parent, linear data is reproduced exactly, an extremum gives flat
  children, time blends 0 / 1 give the old / new level (wrong offset sign,
  dropped slope or wrong blend fails);
* average_down is the mean of the children;
* the coupled step conserves composite mass to round-off in a closed
  reflective box with reflux, and visibly does not without it (a wrong
  register sign or scale fails); a uniform flow and a lake at rest over a
  varying bed are kept.
"""
import numpy as np
import pytest

from oracle import amr as A
from oracle import oracle as O

G = A.G


def _wet_state(rng, ny, nx, bed=True):
    z = 0.3 * rng.random((ny, nx)) if bed else np.zeros((ny, nx))
    h = 1.0 + 0.2 * rng.random((ny, nx))
    hu = 0.1 * rng.standard_normal((ny, nx))
    hv = 0.1 * rng.standard_normal((ny, nx))
    n = np.full((ny, nx), 0.02)
    return h, hu, hv, z, n


@pytest.mark.parametrize("solver", [O.ROE, O.HYDROSTATIC_ROE])
@pytest.mark.parametrize("bc", [((0, 0), (0, 0)), ((1, 0), (0, 1))])
def test_padded_step_equals_level_step(solver, bc):
    rng = np.random.default_rng(3)
    h, hu, hv, z, n = _wet_state(rng, 9, 12)
    lo, hi = bc
    ref = O.step(solver, h, hu, hv, z, n, 0.01, 0.5, 0.5, lo, hi)
    # ghosts as the halo: eta carried across with the scalar rule, h = eta - z
    eta = A.pad_physical(h + z, 0, lo, hi)
    zp = A.pad_physical(z, 3, lo, hi)
    hp = eta - zp
    hp[G:-G, G:-G] = h
    got = O.step_padded(solver, hp, A.pad_physical(hu, 1, lo, hi), A.pad_physical(hv, 2, lo, hi),
                        zp, A.pad_physical(n, 3, lo, hi), 0.01, 0.5, 0.5)
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_effective_fluxes_reproduce_roe_update():
    rng = np.random.default_rng(5)
    ny, nx, dx, dy, dt = 6, 7, 0.5, 0.4, 0.01
    h, hu, hv, z, n = _wet_state(rng, ny, nx)
    lo, hi = (0, 1), (1, 0)
    W = A.working_state(h, hu, hv, z, n, lo, hi)
    ref = O.step(O.ROE, h, hu, hv, z, n, dt, dx, dy, lo, hi)
    for j in range(ny):
        for i in range(nx):
            dxr = A.face_flux(O.ROE, W, (0, i + 1, j, 0), dx, dy)
            dxl = A.face_flux(O.ROE, W, (0, i, j, 1), dx, dy)
            dyr = A.face_flux(O.ROE, W, (1, i, j + 1, 0), dx, dy)
            dyl = A.face_flux(O.ROE, W, (1, i, j, 1), dx, dy)
            u0 = np.array([W["he"][j + G, i + G], hu[j, i], hv[j, i]])
            u = u0 - (dt / dx) * (dxr - dxl) - (dt / dy) * (dyr - dyl)
            np.testing.assert_allclose(u, [ref[0][j, i], ref[1][j, i], ref[2][j, i]],
                                       rtol=0, atol=1e-13)


def test_hydrostatic_register_flux_is_mass_conservative_update():
    """Mass component: FL0 = FR0, so 0.5 FL0 + 0.5 FR0 drives the depth
    update (before the dry clamp, which a wet state never reaches)."""
    rng = np.random.default_rng(6)
    ny, nx, dx, dy, dt = 6, 6, 0.5, 0.5, 0.01
    h, hu, hv, z, n = _wet_state(rng, ny, nx)
    lo, hi = (0, 0), (0, 0)
    W = A.working_state(h, hu, hv, z, n, lo, hi)
    ref = O.step(O.HYDROSTATIC_ROE, h, hu, hv, z, n, dt, dx, dy, lo, hi)
    for j in range(ny):
        for i in range(nx):
            fxr = A.face_flux(1, W, (0, i + 1, j, 0), dx, dy)[0]
            fxl = A.face_flux(1, W, (0, i, j, 0), dx, dy)[0]
            fyr = A.face_flux(1, W, (1, i, j + 1, 0), dx, dy)[0]
            fyl = A.face_flux(1, W, (1, i, j, 0), dx, dy)[0]
            hn = h[j, i] - (dt / dx) * (fxr - fxl) - (dt / dy) * (fyr - fyl)
            assert abs(hn - ref[0][j, i]) <= 1e-14


def _coarse(rng, ny=10, nx=12):
    h, hu, hv, z, n = _wet_state(rng, ny, nx)
    return {"h": h, "hu": hu, "hv": hv, "z": z, "n": n}


def test_fill_conserves_parent_and_reproduces_linear_data():
    rng = np.random.default_rng(7)
    box = (2, 3, 4, 3)
    ny, nx = 10, 12
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    C = {"z": np.zeros((ny, nx)), "h": 2.0 + 0.01 * ii - 0.02 * jj,
         "hu": 0.5 - 0.03 * ii + 0.01 * jj, "hv": 0.25 * np.ones((ny, nx)), "n": np.zeros((ny, nx))}
    zf, _ = A.fine_terrain(C["z"], C["n"], box)
    U = np.zeros((3, 2 * box[3] + 4, 2 * box[2] + 4))
    A.fill_fine(C, C, 0.0, box, (1, 1), (1, 1), zf, U, whole=True)
    # linear data: the fine value is the linear function at the fine centre
    fj, fi = np.meshgrid(np.arange(-G, 2 * box[3] + G), np.arange(-G, 2 * box[2] + G), indexing="ij")
    xc = box[0] + (fi + 0.5) / 2 - 0.5
    yc = box[1] + (fj + 0.5) / 2 - 0.5
    np.testing.assert_allclose(U[0], 2.0 + 0.01 * xc - 0.02 * yc, rtol=0, atol=1e-14)
    np.testing.assert_allclose(U[1], 0.5 - 0.03 * xc + 0.01 * yc, rtol=0, atol=1e-14)
    np.testing.assert_allclose(U[2], 0.25, rtol=0, atol=0)
    # random data: children average to the parent (eta, hu, hv)
    C = _coarse(rng, ny, nx)
    zf, _ = A.fine_terrain(C["z"], C["n"], box)
    U = np.zeros((3, 2 * box[3] + 4, 2 * box[2] + 4))
    A.fill_fine(C, C, 0.0, box, (0, 0), (0, 0), zf, U, whole=True)
    eta_f = U[0] + zf        # wet everywhere, so h + z = the interpolated eta
    ci0, cj0, ncx, ncy = box
    for k, (f, c) in enumerate([(eta_f, C["h"] + C["z"]), (U[1], C["hu"]), (U[2], C["hv"])]):
        v = f[G:-G, G:-G]
        mean = 0.25 * (v[0::2, 0::2] + v[0::2, 1::2] + v[1::2, 0::2] + v[1::2, 1::2])
        np.testing.assert_allclose(mean, c[cj0:cj0 + ncy, ci0:ci0 + ncx], rtol=0, atol=1e-14)


def test_fill_extremum_is_flat_and_time_blend():
    box = (2, 2, 2, 2)
    C = {k: np.zeros((6, 6)) for k in ("h", "hu", "hv", "z", "n")}
    C["h"][:] = 1.0
    C["h"][2, 2] = 2.0            # a local maximum: MC slope 0 in both directions
    zf, _ = A.fine_terrain(C["z"], C["n"], box)
    U = np.zeros((3, 8, 8))
    A.fill_fine(C, C, 0.0, box, (0, 0), (0, 0), zf, U, whole=True)
    np.testing.assert_array_equal(U[0][G:G + 2, G:G + 2], 2.0)
    # blend: old everywhere 1, new everywhere 3
    Co = {k: v.copy() for k, v in C.items()}
    Cn = {k: v.copy() for k, v in C.items()}
    Co["h"][:] = 1.0
    Cn["h"][:] = 3.0
    for a, want in ((0.0, 1.0), (1.0, 3.0), (0.5, 2.0)):
        U = np.zeros((3, 8, 8))
        A.fill_fine(Co, Cn, a, box, (0, 0), (0, 0), zf, U)
        ring = np.ones((8, 8), bool)
        ring[G:-G, G:-G] = False
        np.testing.assert_array_equal(U[0][ring], want)
        np.testing.assert_array_equal(U[0][~ring], 0.0)    # valid cells untouched


def test_average_down_is_child_mean():
    rng = np.random.default_rng(9)
    box = (1, 2, 3, 2)
    f = [rng.random((4, 6)) for _ in range(3)]
    U = [np.zeros((6, 6)) for _ in range(3)]
    A.average_down(f, U, box)
    for k in range(3):
        want = f[k].reshape(2, 2, 3, 2).mean(axis=(1, 3))
        np.testing.assert_allclose(U[k][2:4, 1:4], want, rtol=0, atol=1e-15)
        assert U[k][0, 0] == 0.0


def _composite_mass(T: A.TwoLevelOracle):
    return T.C["h"].sum() * T.dx * T.dy


@pytest.mark.parametrize("solver", [O.ROE, O.HYDROSTATIC_ROE])
@pytest.mark.parametrize("subcycle", [True, False])
def test_composite_mass_conserved_with_reflux(solver, subcycle, monkeypatch):
    ny, nx = 12, 14
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    # the bed is flat within two cells of the walls: the hydrostatic scheme
    # mirrors eta but extrapolates the terrain across a reflective wall, so
    # a bed slope at the wall leaks mass through it (a single-level property)
    z = 0.05 * np.sin(0.7 * ii) * np.cos(0.5 * jj)
    z[:2, :] = z[-2:, :] = 0.0
    z[:, :2] = z[:, -2:] = 0.0
    h = 1.0 + 0.3 * np.exp(-((ii - 6.0) ** 2 + (jj - 5.0) ** 2) / 6.0)
    zero = np.zeros((ny, nx))
    box = (4, 3, 5, 5)
    dx = dy = 1.0

    def run(with_reflux):
        T = A.TwoLevelOracle(solver, h, zero, zero, z, zero, dx, dy, (0, 0), (0, 0), box,
                             subcycle=subcycle)
        # start from a composite state (covered cells = their children's mean)
        A.average_down([T.F["U"][c][G:-G, G:-G] for c in range(3)],
                       [T.C["h"], T.C["hu"], T.C["hv"]], box)
        m0 = _composite_mass(T)
        if not with_reflux:
            monkeypatch.setattr(A, "reflux", lambda *a, **k: None)
        for _ in range(6):
            d0, d1 = A.level_dts(solver, T.C, T.F, dx, dy)
            T.step(*A.dt_rule(0.4 * d0, 0.4 * d1, subcycle))
        monkeypatch.undo()
        return abs(_composite_mass(T) - m0) / m0

    with_rf = run(True)
    without = run(False)
    assert with_rf < 1e-14, with_rf
    assert without > 1e3 * max(with_rf, 1e-17), (with_rf, without)


@pytest.mark.parametrize("solver", [O.ROE, O.HYDROSTATIC_ROE])
def test_uniform_flow_is_kept(solver):
    ny, nx = 8, 10
    ones = np.ones((ny, nx))
    T = A.TwoLevelOracle(solver, 1.5 * ones, 0.3 * ones, -0.2 * ones, 0 * ones, 0 * ones,
                         1.0, 1.0, (1, 1), (1, 1), (3, 2, 4, 3))
    for _ in range(3):
        T.step(0.05, 0.025)
    for c, want in zip(range(3), (1.5, 0.3, -0.2)):
        a = [T.C["h"], T.C["hu"], T.C["hv"]][c]
        np.testing.assert_allclose(a, want, rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(T.fine_valid()[c], want, rtol=1e-14, atol=1e-14)


def test_lake_at_rest_hydrostatic():
    ny, nx = 10, 12
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    z = 0.4 * np.sin(0.9 * ii + 0.3) * np.cos(0.6 * jj) + 0.1 * ((ii * 7 + jj * 3) % 5) / 5.0
    eta = 2.0
    zero = np.zeros((ny, nx))
    T = A.TwoLevelOracle(O.HYDROSTATIC_ROE, eta - z, zero, zero, z, np.full((ny, nx), 0.03),
                         0.5, 0.5, (0, 0), (0, 0), (3, 2, 5, 4))
    for _ in range(4):
        T.step(0.02, 0.01)
    np.testing.assert_allclose(T.C["h"] + T.C["z"], eta, rtol=0, atol=1e-13)
    np.testing.assert_allclose(T.C["hu"], 0.0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(T.C["hv"], 0.0, rtol=0, atol=1e-13)
    Uf = T.fine_valid()
    np.testing.assert_allclose(Uf[0] + T.F["z"][G:-G, G:-G], eta, rtol=0, atol=1e-13)
    np.testing.assert_allclose(Uf[1:], 0.0, rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------
# the multi-box hierarchy restatement
# ---------------------------------------------------------------------------
def _field(ny, nx, seed=1, flat_ring=True):
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(seed)
    z = 0.05 * np.sin(0.5 * ii) * np.cos(0.4 * jj)
    if flat_ring:
        z[:2, :] = z[-2:, :] = 0.0
        z[:, :2] = z[:, -2:] = 0.0
    h = 1.0 + 0.3 * np.exp(-((ii - 0.4 * nx) ** 2 + (jj - 0.5 * ny) ** 2) / (0.05 * nx * ny))
    hu = 0.05 * rng.standard_normal((ny, nx))
    hv = 0.05 * rng.standard_normal((ny, nx))
    return h, hu, hv, z, np.zeros((ny, nx))


@pytest.mark.parametrize("solver", [O.ROE, O.HYDROSTATIC_ROE])
def test_hierarchy_single_box_equals_two_level(solver):
    ny, nx = 14, 16
    st = _field(ny, nx)
    box = (4, 3, 6, 5)
    T = A.TwoLevelOracle(solver, *st, 1.0, 1.0, (0, 1), (1, 0), box)
    H = A.AmrOracle(solver, *st, 1.0, 1.0, (0, 1), (1, 0), {1: [(8, 6, 12, 10)]})
    for _ in range(3):
        dts = H.compute_dt(0.4)
        T.step(*dts)
        H.step(dts=dts)
    # the hierarchy blends the second substep's ghosts with AMReX's clock
    # ratio (t - t_old) / (t_new - t_old), which can miss 1/2 by an ulp; the
    # two-level restatement uses 1/2 exactly
    tol = dict(rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(H.levels[0].interior(H.levels[0].U[0]), T.C["h"], **tol)
    np.testing.assert_allclose(H.levels[0].interior(H.levels[0].U[1]), T.C["hu"], **tol)
    fine = H.levels[1].U[:, G + 6:G + 16, G + 8:G + 20]
    np.testing.assert_allclose(fine, T.fine_valid(), **tol)


@pytest.mark.parametrize("solver", [O.ROE, O.HYDROSTATIC_ROE])
def test_hierarchy_split_boxes_equal_merged_box(solver):
    ny, nx = 16, 20
    st = _field(ny, nx, seed=3)
    merged = A.AmrOracle(solver, *st, 1.0, 1.0, (0, 0), (0, 0), {1: [(8, 8, 20, 12)]})
    split = A.AmrOracle(solver, *st, 1.0, 1.0, (0, 0), (0, 0),
                        {1: [(8, 8, 12, 12), (20, 8, 8, 6), (20, 14, 8, 6)]})
    for _ in range(3):
        dts = merged.compute_dt(0.4)
        merged.step(dts=dts)
        split.step(dts=dts)
    for l in range(2):
        a, b = merged.levels[l], split.levels[l]
        np.testing.assert_array_equal(a.interior(a.U)[:, a.interior(a.mask)],
                                      b.interior(b.U)[:, b.interior(b.mask)])


@pytest.mark.parametrize("subcycle", [True, False])
def test_hierarchy_three_levels_conserve_mass(subcycle):
    ny, nx = 16, 16
    st = _field(ny, nx, seed=5)
    boxes = {1: [(6, 6, 14, 12), (20, 6, 6, 10)], 2: [(18, 18, 16, 12)]}
    H = A.AmrOracle(O.HYDROSTATIC_ROE, *st, 1.0, 1.0, (0, 0), (0, 0), boxes, subcycle=subcycle)
    for l in (1, 0):                                   # composite start
        H._average_down(l)
    m0 = H.levels[0].interior(H.levels[0].U[0]).sum()
    for _ in range(4):
        H.step(cfl=0.4)
    m1 = H.levels[0].interior(H.levels[0].U[0]).sum()
    assert abs(m1 - m0) / m0 < 1e-14


def test_hierarchy_lake_at_rest_three_levels():
    ny, nx = 16, 16
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    z = 0.3 * np.sin(0.7 * ii) * np.cos(0.5 * jj)
    zero = np.zeros((ny, nx))
    boxes = {1: [(4, 4, 16, 14), (20, 8, 8, 12)], 2: [(14, 14, 20, 16)]}
    H = A.AmrOracle(O.HYDROSTATIC_ROE, 2.0 - z, zero, zero, z, np.full((ny, nx), 0.02),
                    1.0, 1.0, (0, 0), (0, 0), boxes)
    for _ in range(3):
        H.step(cfl=0.4)
    for lv in H.levels:
        m = lv.interior(lv.mask)
        np.testing.assert_allclose((lv.interior(lv.U[0]) + lv.interior(lv.T[0]))[m], 2.0,
                                   rtol=0, atol=1e-13)
        np.testing.assert_allclose(lv.interior(lv.U[1:])[:, m], 0.0, rtol=0, atol=1e-13)
