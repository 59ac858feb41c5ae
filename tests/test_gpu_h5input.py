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


def test_levels_from_static_layer_and_hierarchy(gpu, tmp_path):
    """Every level of a run from the file's static layer (terrain_ref_lev 1):
    level 0 as level0_from_file, level 1 a copy with h rebuilt, level 2 the
    conservative interpolation of eta_s and momenta on a piecewise-constant
    bed (the oracle's fill restatement over the whole level); then a
    hierarchy built from the file steps and keeps its mass."""
    from oracle import amr as A
    ny, nx = 32, 64
    r = np.random.default_rng(11)
    z = r.uniform(0, 0.5, (ny, nx))
    n = np.full((ny, nx), 0.03)
    z[:4, :8] = -9999.0
    n[:4, :8] = -9999.0
    h = np.where(z > -9000, np.maximum(1.2 - z, 0.0), 0.0)
    hu = 0.1 * r.standard_normal((ny, nx)) * h
    hv = 0.1 * r.standard_normal((ny, nx)) * h
    p = str(tmp_path / "in.h5")
    H.write_input(p, np.stack([z, n]), np.stack([h, hu, hv]), 5.0, 5.0, 0.0, 0.0)
    meta, terr, fluid = H._static_layer(p)
    lv0, _ = H.level0_from_file(p, terrain_ref_lev=1, blocking_factor=4)
    a0 = H.level_from_static(meta, terr, fluid, 1, 0)
    np.testing.assert_array_equal(a0[0].cpu().numpy(), lv0.U_new[0].cpu().numpy())
    np.testing.assert_array_equal(a0[3].cpu().numpy(), lv0.terrain[0].cpu().numpy())
    a1 = H.level_from_static(meta, terr, fluid, 1, 1)
    eta_s = np.where(n == -9999.0, z, h + z)
    np.testing.assert_array_equal(a1[0].cpu().numpy(), np.maximum(eta_s - z, 0.0))
    np.testing.assert_array_equal(a1[1].cpu().numpy(), hu)
    a2 = H.level_from_static(meta, terr, fluid, 1, 2, (1, 0), (0, 1))
    zf = np.pad(np.repeat(np.repeat(z, 2, 0), 2, 1), 2, mode="edge")
    U = np.zeros((3, 2 * ny + 4, 2 * nx + 4))
    Cd = {"h": np.where(n == -9999.0, 0.0, h), "hu": hu, "hv": hv, "z": z, "n": n}
    A.fill_fine(Cd, Cd, 0.0, (0, 0, nx, ny), (1, 0), (0, 1), zf, U, whole=True)
    for k in range(3):
        np.testing.assert_array_equal(a2[k].cpu().numpy(), U[k, 2:-2, 2:-2])
    np.testing.assert_array_equal(a2[3].cpu().numpy(), zf[2:-2, 2:-2])
    # a hierarchy from the file: levels 0..2, NODATA corner, closed box
    Hy, geo = H.hierarchy_from_file(p, "HydrostaticRoe", terrain_ref_lev=1, max_level=2,
                                    blocking_factor=4, bc_lo=(0, 0), bc_hi=(0, 0),
                                    regrid_int=2, tag_thresh=[(1e-4, 1e300, 1e300)])
    assert geo["base_nx"] == nx // 2 and Hy.finest >= 1
    for l in range(Hy.finest - 1, -1, -1):
        Hy._average_down(l)
    m0 = Hy.conservation()["composite_mass"]
    Hy.run_native(1e30, 0.45, max_steps=6)
    m1 = Hy.conservation()["composite_mass"]
    assert np.isfinite(m1) and abs(m1 - m0) <= 1e-12 * m0
