"""Level 0 from an input file on the GPU against a restatement of the
reference's initialisation: terrain_ref_lev 0 (copy, h = max(0, eta - z)
with eta = z at NODATA) and 1, 2 (terrain by the oracle's
masked_average_down, momenta and eta by plain r x r means summed jref-outer,
iref-inner, then scaled by 1/r^2, as AMReX's average_down)."""
import os

import numpy as np
import pytest
import torch

import oracle as O
from hydroexa_amd import h5input as H

pytestmark = pytest.mark.gpu
if not os.path.exists(H._IO_PATH):
    pytest.skip("libhydroexa_io.so not built (no HDF5)", allow_module_level=True)


def avg(fine, r):
    c = np.zeros(fine[::r, ::r].shape)
    for jref in range(r):
        for iref in range(r):
            c = c + fine[jref::r, iref::r]
    return (1.0 / (r * r)) * c


@pytest.mark.parametrize("ref", [0, 1, 2])
def test_level0_from_file(gpu, tmp_path, ref):
    ny, nx = 32, 64
    r = np.random.default_rng(ref)
    z = r.uniform(0, 1, (ny, nx))
    n = np.full((ny, nx), 0.03)
    z[:6, :10] = -9999.0
    n[:6, :10] = -9999.0
    h = np.where(z > -9000, np.maximum(1.2 - z, 0.0), 0.0)
    hu = r.standard_normal((ny, nx)) * h
    hv = r.standard_normal((ny, nx)) * h
    p = str(tmp_path / "in.h5")
    H.write_input(p, np.stack([z, n]), np.stack([h, hu, hv]), 5.0, 5.0, 10.0, 20.0)
    lv, geo = H.level0_from_file(p, terrain_ref_lev=ref, blocking_factor=4)
    s = 2 ** ref
    assert (lv.nx, lv.ny) == (nx // s, ny // s) and lv.dx == 5.0 * s
    eta_s = np.where(n == -9999.0, z, h + z)
    if ref == 0:
        zc, nc, eta, qx, qy = z, n, eta_s, hu, hv
    else:
        t, _ = O.masked_average_down(np.stack([z, n]), s, s, 5.0, 5.0)
        zc, nc = t[0], t[1]
        eta, qx, qy = avg(eta_s, s), avg(hu, s), avg(hv, s)
    x = eta - zc
    hc = np.where(x > 0, x, 0.0)
    np.testing.assert_array_equal(lv.terrain[0].cpu().numpy(), zc)
    np.testing.assert_array_equal(lv.terrain[1].cpu().numpy(), nc)
    np.testing.assert_array_equal(lv.U_new[0].cpu().numpy(), hc)
    np.testing.assert_array_equal(lv.U_new[1].cpu().numpy(), qx)
    np.testing.assert_array_equal(lv.U_new[2].cpu().numpy(), qy)
