"""Pinning of the oracle's conservation diagnostics
(src/state/AmrMeshState_Utils.cpp:213-343) against exactly rounded sums
(math.fsum) of the per-cell terms, and the masking rules: roughness NODATA
and non-positive depth are skipped, min h is DBL_MAX on an empty level."""
import math

import numpy as np

import oracle as O

DBL_MAX = 1.7976931348623157e308
EPS = np.finfo(np.float64).eps


def fields(ny=37, nx=53, seed=0):
    r = np.random.default_rng(seed)
    h = r.uniform(-0.2, 2.0, (ny, nx))          # some dry (h <= 0) cells
    hu = r.standard_normal((ny, nx))
    hv = r.standard_normal((ny, nx))
    n = np.full((ny, nx), 0.03)
    n[:5, :7] = -9999.0                          # NODATA by roughness
    return h, hu, hv, n


def test_sums_match_exact_sums_within_serial_bound():
    h, hu, hv, n = fields()
    dx, dy = 0.37, 0.21
    dA = dx * dy
    out = O.conservation(h, hu, hv, n, dx, dy)
    m = (n != -9999.0) & (h > 0.0)
    terms = [h[m] * dA, hu[m] * dA, hv[m] * dA, 0.5 * (hu[m] * hu[m] + hv[m] * hv[m]) / h[m] * dA]
    for k, t in enumerate(terms):
        exact = math.fsum(t)
        bound = len(t) * EPS * math.fsum(np.abs(t))
        assert abs(out[k] - exact) <= bound, (k, out[k], exact)
    assert out[4] == h[m].min()


def test_masks_and_empty_level():
    h, hu, hv, n = fields(seed=1)
    # a cell with NODATA roughness but positive depth is skipped
    n2 = n.copy()
    n2[10, 10] = -9999.0
    a = O.conservation(h, hu, hv, n, 1.0, 1.0)
    b = O.conservation(h, hu, hv, n2, 1.0, 1.0)
    if h[10, 10] > 0:
        assert abs((a[0] - b[0]) - h[10, 10]) < 1e-9 * abs(a[0])
    # the bed is not the mask: z = -9999 with valid roughness still counts
    out = O.conservation(np.full((3, 4), -1.0), np.ones((3, 4)), np.ones((3, 4)),
                         np.zeros((3, 4)), 1.0, 1.0)
    assert list(out[:4]) == [0.0, 0.0, 0.0, 0.0] and out[4] == DBL_MAX
