"""Pin the oracle's CFL estimate, tagging, masked_average_down and pruning
with closed forms, invariants and brute force on small inputs."""
import numpy as np
import pytest

import oracle as O

G = 9.81


def test_level_dt_is_the_unsplit_cfl_bound():
    """dt*(|u|+c)/dx + dt*(|v|+c)/dy <= 1 on every wet cell with equality at
    the binding cell; max_vel is max(|u|, |v|)."""
    rng = np.random.default_rng(30)
    h = rng.uniform(0.1, 2.0, (17, 23))
    hu, hv = rng.normal(size=h.shape) * h, rng.normal(size=h.shape) * h
    z = rng.uniform(0, 1, h.shape)
    dx, dy = 0.013, 0.021
    dt, vmax = O.level_dt(h, hu, hv, z, dx, dy)
    u, v, c = hu / h, hv / h, np.sqrt(G * h)
    cour = dt * (np.abs(u) + c) / dx + dt * (np.abs(v) + c) / dy
    assert cour.max() == pytest.approx(1.0, rel=1e-15)
    assert vmax == np.max(np.maximum(np.abs(u), np.abs(v)))


def test_level_dt_special_cells():
    h = np.ones((4, 4))
    z = np.zeros((4, 4))
    zero = np.zeros((4, 4))
    z[:] = -9999.0
    assert O.level_dt(h, zero, zero, z, 0.1, 0.1) == (1e30, 0.0)          # all NODATA
    z[:] = 0.0
    assert O.level_dt(zero, zero, zero, z, 0.1, 0.1) == (1e6, 0.0)        # all dry
    hh = np.full((4, 4), 1e-12)                                            # h <= 1e-12 is dry
    assert O.level_dt(hh, zero, zero, z, 0.1, 0.1)[0] == 1e6
    z[0, 0] = -9999.0 + 5e-11                                             # |z + 9999| < 1e-10
    d, _ = O.level_dt(h, zero, zero, z, 0.1, 0.1)
    assert d == pytest.approx(1.0 / (2 * np.sqrt(G) / 0.1), rel=1e-15)


def test_tag_linear_depth_field():
    """Central differences are exact on a linear field: every interior cell
    is tagged iff |grad h| exceeds the threshold; the boundary columns see the
    foextrap halo (a one-sided, halved gradient)."""
    n = 24
    dx = 0.1
    X, Y = np.meshgrid((np.arange(n) + .5) * dx, (np.arange(n) + .5) * dx)
    a, b = 0.03, 0.04                      # |grad h| = 0.05
    h = 1.0 + a * X + b * Y
    zero = np.zeros_like(h)
    for solver in (O.ROE, O.HYDROSTATIC_ROE):
        t = O.tag(solver, h, zero, zero, zero, dx, dx, 1.0, 0.049, 1e9, 1e9)
        assert t[1:-1, 1:-1].all()
        t = O.tag(solver, h, zero, zero, zero, dx, dx, 1.0, 0.051, 1e9, 1e9)
        assert not t[1:-1, 1:-1].any()


def test_tag_velocity_and_bed_criteria():
    n = 20
    dx = 0.05
    X, _ = np.meshgrid((np.arange(n) + .5) * dx, (np.arange(n) + .5) * dx)
    h = np.ones((n, n))
    hu = 0.2 * X                            # |grad |u|| = 0.2
    zero = np.zeros_like(h)
    z = 0.3 * X                             # |grad z| = 0.3
    for solver in (O.ROE, O.HYDROSTATIC_ROE):
        t = O.tag(solver, h, hu, zero, zero, dx, dx, 1.0, 1e9, 0.19, 1e9)
        assert t[1:-1, 1:-1].all()
        # bed criterion only at t <= 1e-8
        assert O.tag(solver, h, zero, zero, z, dx, dx, 0.0, 1e9, 1e9, 0.29)[1:-1, 1:-1].all()
        assert not O.tag(solver, h, zero, zero, z, dx, dx, 1e-7, 1e9, 1e9, 0.29).any()
        # cells with h < 1e-6 count as |u| = 0, whatever their momentum
        hd = np.full((n, n), 9e-7)
        assert not O.tag(solver, hd, hu * 9e-7, zero, zero, dx, dx, 1.0, 1e9, 1e-9, 1e9).any()
        hw = np.full((n, n), 2e-6)
        assert O.tag(solver, hw, hu * 2e-6, zero, zero, dx, dx, 1.0, 1e9, 0.19, 1e9)[1:-1, 1:-1].all()


def test_tag_nodata_neighbour_tests():
    """Roe skips cells with a 4-neighbour bed == -9999; the hydrostatic
    solver uses |z + 9999| < 1e-10 (the two tests differ for z = -9999 +
    1e-11)."""
    n = 10
    h = np.ones((n, n)); h[:, 5:] = 2.0     # a jump tags columns 4 and 5
    zero = np.zeros_like(h)
    z = np.zeros_like(h)
    z[3, 4] = -9999.0 + 1e-11               # neighbour of (2,4), (4,4), (3,3), (3,5)
    tr = O.tag(O.ROE, h, zero, zero, z, 0.1, 0.1, 1.0, 0.5, 1e9, 1e9)
    th = O.tag(O.HYDROSTATIC_ROE, h, zero, zero, z, 0.1, 0.1, 1.0, 0.5, 1e9, 1e9)
    assert tr[2, 4] == 1 and th[2, 4] == 0
    z[3, 4] = -9999.0
    tr = O.tag(O.ROE, h, zero, zero, z, 0.1, 0.1, 1.0, 0.5, 1e9, 1e9)
    assert tr[2, 4] == 0 and tr[3, 5] == 0 and tr[5, 4] == 1


def test_masked_average_down_closed_forms():
    nd = -9999.0
    fine = np.array([[1.0, 3.0, 5.0, nd],
                     [5.0, 7.0, np.nan, nd],
                     [nd, nd, 2.0, 4.0],
                     [nd, 9.0, nd, -9999.0005]])     # within 1e-3 of NODATA
    crse, nv = O.masked_average_down(fine, 2, 2, 0.5, 0.5, nd)
    assert crse[0, 0, 0] == 4.0 and nv[0, 0, 0] == 4            # all valid: mean
    assert crse[0, 0, 1] == nd and nv[0, 0, 1] == 1             # NaN and NODATA invalid
    assert crse[0, 1, 0] == nd and nv[0, 1, 0] == 1             # 1 of 4 < 50 %
    assert crse[0, 1, 1] == 3.0 and nv[0, 1, 1] == 2            # exactly 50 %: mean of valid


def test_masked_average_down_brute_force():
    rng = np.random.default_rng(31)
    nd = -9999.0
    for rx, ry in ((2, 2), (4, 4), (2, 4)):
        f = rng.uniform(-5, 5, (2, 8 * ry, 8 * rx))
        f[rng.uniform(size=f.shape) < 0.35] = nd
        crse, nv = O.masked_average_down(f, rx, ry, 0.1, 0.2, nd)
        blocks = f.reshape(2, 8, ry, 8, rx).transpose(0, 1, 3, 2, 4).reshape(2, 8, 8, rx * ry)
        valid = ~np.isnan(blocks) & (np.abs(blocks - nd) > 1e-3)
        cnt = valid.sum(-1)
        mean = np.where(valid, blocks, 0).sum(-1) / np.maximum(cnt, 1)
        np.testing.assert_array_equal(nv, cnt)
        expect = np.where(2 * cnt >= rx * ry, mean, nd)
        np.testing.assert_allclose(crse, expect, rtol=1e-14)


def test_prune_boxes_brute_force():
    rng = np.random.default_rng(32)
    n0, box, lev = 32, 8, 1
    sc = 2 ** lev
    z = rng.uniform(0, 1, (n0 * sc, n0 * sc))
    z[: 13 * sc, : 19 * sc] = -9999.0
    z[5, 60] = np.nan                               # NaN counts as valid (not <= -9998)
    boxes = [(i, j, i + box - 1, j + box - 1) for j in range(0, n0, box) for i in range(0, n0, box)]
    keep, nod, nk = O.prune_boxes(boxes, z, lev)
    for b, (i0, j0, i1, j1) in enumerate(boxes):
        blk = z[j0 * sc:(j1 + 1) * sc, i0 * sc:(i1 + 1) * sc]
        isnd = blk <= -9998.0
        assert keep[b] == (~isnd).any()
        assert nod[b] == isnd.any()
    assert nk == keep.sum() < len(boxes)


def test_prune_keeps_everything_when_all_nodata():
    z = np.full((16, 16), -9999.0)
    keep, nod, nk = O.prune_boxes([(0, 0, 7, 7), (8, 0, 15, 7)], z, 0)
    assert keep.all() and not nod.any() and nk == 2
