"""Pinning of the tangent-linear reference (oracle/tlm.py), CPU only.

* Its real-valued step is the C restatement's step bit for bit (both
  solvers, every BC mix, NODATA, friction), so the tangent linearises the
  same map the parity tests pin.
* Its complex-step directional derivatives match central finite
  differences of the C restatement, converging at the central-difference
  rate as the step shrinks.
* It is linear in the direction, as a derivative must be.
"""
import numpy as np
import pytest

import oracle as O
from oracle import tlm as T


def state(n=20, seed=0, friction=0.0, nodata=False):
    r = np.random.default_rng(seed)
    X, Y = np.meshgrid((np.arange(n) + 0.5) / n, (np.arange(n) + 0.5) / n)
    z = 0.2 * np.exp(-((X - 0.4) ** 2 + (Y - 0.6) ** 2) / 0.02) + 0.01 * r.random((n, n))
    eta = np.where(X < 0.5, 1.0, 0.4) + 0.01 * r.random((n, n))
    h = np.maximum(eta - z, 0.0)
    hu = 0.1 * r.standard_normal((n, n)) * h
    hv = 0.1 * r.standard_normal((n, n)) * h
    nn = (friction + 0.01 * r.random((n, n))) if friction else np.zeros((n, n))
    if nodata:
        z[:4, :4] = -9999.0
        nn[:4, :4] = -9999.0
        h[:4, :4] = hu[:4, :4] = hv[:4, :4] = 0.0
    return h, hu, hv, z, nn


BCS = [((1, 1), (1, 1)), ((0, 0), (0, 0)), ((0, 1), (1, 0))]


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("friction,nodata", [(0.0, False), (0.03, False), (0.03, True)])
def test_real_step_equals_c_restatement(solver, bc, friction, nodata):
    if solver == 0 and nodata:
        pytest.skip("NODATA is a hydrostatic-solver feature")
    h, hu, hv, z, n = state(friction=friction, nodata=nodata, seed=solver)
    a = O.step(solver, h, hu, hv, z, n, 2e-4, 1 / 20, 1 / 20, bc[0], bc[1])
    b = T.step(solver, h, hu, hv, z, n, 2e-4, 1 / 20, 1 / 20, bc[0], bc[1])
    for k in range(3):
        np.testing.assert_array_equal(a[k], b[k])


def _c_rollout(solver, h, hu, hv, z, n, dts, dx, bc, gauges):
    series = []
    for dt in dts:
        h, hu, hv = O.step(solver, h, hu, hv, z, n, dt, dx, dx, bc[0], bc[1])
        series.append(h.reshape(-1)[gauges])
    return np.array(series), (h, hu, hv)


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("bc", BCS[:2])
def test_complex_step_matches_central_differences(solver, bc):
    n = 20
    h, hu, hv, z, nn = state(n, seed=solver, friction=0.03)
    dts = [1.5e-3] * 8
    g = np.array([5 * n + 7, 12 * n + 12, 16 * n + 3])
    r = np.random.default_rng(7)
    dh = r.standard_normal((n, n)) * (h > 0.01)
    dz = r.standard_normal((n, n))
    dn = r.standard_normal((n, n))
    ds, df, _ = T.directional_derivative(solver, (h, hu, hv), z, nn, dts, 1 / n, 1 / n, bc[0],
                                         bc[1], g, d_state=(dh, None, None), d_z=dz, d_n=dn)
    errs = []
    for eps in (1e-5, 1e-6):
        sp, fp = _c_rollout(solver, h + eps * dh, hu, hv, z + eps * dz, nn + eps * dn, dts, 1 / n, bc, g)
        sm, fm = _c_rollout(solver, h - eps * dh, hu, hv, z - eps * dz, nn - eps * dn, dts, 1 / n, bc, g)
        fd = (fp[1] - fm[1]) / (2 * eps)
        errs.append(np.abs(fd - df[1]).max() / np.abs(df[1]).max())
    # central differences converge to the tangent at O(eps^2) (10x smaller
    # eps -> ~100x smaller error) until rounding takes over
    assert errs[1] < 2e-5, errs
    assert errs[1] < errs[0] / 20 or errs[1] < 1e-8, errs


@pytest.mark.parametrize("solver", [0, 1])
def test_tangent_is_linear_in_the_direction(solver):
    n = 16
    h, hu, hv, z, nn = state(n, seed=4, friction=0.02)
    dts = [1e-3] * 5
    g = np.array([3 * n + 3, 9 * n + 10])
    r = np.random.default_rng(9)
    d1 = r.standard_normal((n, n))
    d2 = r.standard_normal((n, n))
    f = lambda d: T.directional_derivative(solver, (h, hu, hv), z, nn, dts, 1 / n, 1 / n, (1, 1),
                                           (1, 1), g, d_z=d)[1][0]
    a, b, c = f(d1), f(d2), f(2.0 * d1 - 3.0 * d2)
    np.testing.assert_allclose(c, 2.0 * a - 3.0 * b, rtol=1e-9, atol=1e-12 * np.abs(c).max())
