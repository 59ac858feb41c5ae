"""HDF5 input schema (CPU): write/read round trip in the PyHydroEXA schema,
the reader's NODATA fill outside the file box, metadata, and every branch of
the geometry contract (src/HydroEXA.cpp:95-205) with the reference's
assertion text."""
import os

import numpy as np
import pytest

from hydroexa_amd import h5input as H

pytest.importorskip("ctypes")
if not os.path.exists(H._IO_PATH):
    pytest.skip("libhydroexa_io.so not built (no HDF5)", allow_module_level=True)


def sample(ny=24, nx=40, seed=0):
    r = np.random.default_rng(seed)
    terrain = np.stack([r.uniform(0, 1, (ny, nx)), np.full((ny, nx), 0.03)])
    terrain[:, :3, :5] = -9999.0
    fluid = np.stack([r.uniform(0, 2, (ny, nx)), r.standard_normal((ny, nx)),
                      r.standard_normal((ny, nx))])
    return terrain, fluid


def test_round_trip_and_metadata(tmp_path):
    t, f = sample()
    p = str(tmp_path / "in.h5")
    H.write_input(p, t, f, 2.5, 2.5, 100.0, -50.0)
    m = H.read_metadata(p, "terrain")
    assert (m["nx"], m["ny"], m["ndims"], m["ncomp"]) == (40, 24, 3, 2)
    assert (m["dx"], m["dy"], m["x_ll"], m["y_ll"]) == (2.5, 2.5, 100.0, -50.0)
    assert m["x_hi"] == 100.0 + 40 * 2.5 and m["y_hi"] == -50.0 + 24 * 2.5
    assert H.read_metadata(p, "fluid")["ncomp"] == 3
    np.testing.assert_array_equal(H.read_components(p, "terrain", 0, 2, 0, 0, 40, 24, -9999.0), t)
    np.testing.assert_array_equal(H.read_components(p, "fluid", 0, 3, 0, 0, 40, 24, 0.0), f)
    np.testing.assert_array_equal(H.read_components(p, "fluid", 1, 2, 0, 0, 40, 24, 0.0), f[1:])


def test_box_past_the_file_is_nodata_filled(tmp_path):
    t, f = sample()
    p = str(tmp_path / "in.h5")
    H.write_input(p, t, f, 1.0, 1.0)
    out = H.read_components(p, "terrain", 0, 2, -3, 20, 50, 8, -9999.0)
    ref = np.full((2, 8, 50), -9999.0)
    ref[:, :4, 3:43] = t[:, 20:24, 0:40]
    np.testing.assert_array_equal(out, ref)
    # a box entirely outside the file
    out = H.read_components(p, "fluid", 0, 3, 100, 100, 4, 4, 0.0)
    assert (out == 0.0).all()


def test_errors(tmp_path):
    t, f = sample()
    p = str(tmp_path / "in.h5")
    with pytest.raises(H.InputError):
        H.read_metadata(str(tmp_path / "missing.h5"))
    H.write_input(p, t, f, 1.0, 1.0)
    with pytest.raises(H.InputError):
        H.read_metadata(p, "nothere")
    with pytest.raises(H.InputError):
        H.read_components(p, "terrain", 1, 2, 0, 0, 4, 4, 0.0)     # only 2 components
    with pytest.raises(H.InputError):
        H.write_input(str(tmp_path / "b.h5"), t, f, 1.0, 1.1)       # non-square cells


def meta(nx=64, ny=32, dx=2.0, dy=2.0, x_ll=0.0, y_ll=0.0, **kw):
    m = {"nx": nx, "ny": ny, "ndims": 3, "ncomp": 2, "dx": dx, "dy": dy, "x_ll": x_ll,
         "y_ll": y_ll, "x_hi": x_ll + nx * dx, "y_hi": y_ll + ny * dy}
    m.update(kw)
    return m


def test_geometry_contract_accepts_and_derives():
    g = H.check_geometry(meta(), terrain_ref_lev=1, blocking_factor=8)
    assert (g["base_nx"], g["base_ny"], g["refinement_scale"]) == (32, 16, 2)
    assert g["level0_dx"] == 4.0 and g["prob_hi"] == (128.0, 64.0)


@pytest.mark.parametrize("m,ref,bf,text", [
    (meta(), 0, 0, "amr.blocking_factor must be greater than zero."),
    (meta(), -1, 8, "amr.terrain_ref_lev must be non-negative."),
    (meta(nx=0), 0, 8, "HDF5 terrain dimensions must be positive."),
    (meta(dx=0.0), 0, 8, "HDF5 terrain cell spacing must be positive."),
    (meta(dy=2.0001), 0, 8, "HydroEXA requires square HDF5 cells: dx must equal dy."),
    (meta(nx=66), 2, 1, "HDF5 global_nx must be divisible by 2^amr.terrain_ref_lev."),
    (meta(ny=34), 2, 1, "HDF5 global_ny must be divisible by 2^amr.terrain_ref_lev."),
    (meta(nx=72), 0, 16, "Level-0 nx is incompatible with amr.blocking_factor. HydroEXA will not "
                         "pad the physical domain."),
    (meta(ny=40), 0, 16, "Level-0 ny is incompatible with amr.blocking_factor. HydroEXA will not "
                         "pad the physical domain."),
    (meta(x_hi=0.0), 0, 8, "HDF5 physical domain must have positive extents."),
    (meta(x_hi=130.0), 0, 8, "HDF5 x extent is inconsistent with global_nx, terrain_ref_lev, and dx."),
    (meta(y_hi=70.0), 0, 8, "HDF5 y extent is inconsistent with global_ny, terrain_ref_lev, and dy."),
])
def test_geometry_contract_rejects(m, ref, bf, text):
    with pytest.raises(H.GeometryError) as e:
        H.check_geometry(m, ref, bf)
    assert str(e.value) == text
