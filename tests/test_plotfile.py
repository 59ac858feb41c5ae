"""Native AMReX plotfile writer (CPU): the reference's naming and variable
list (src/io/NativePlotFile.cpp:19-97), the plotfile / VisMF / FAB layout of
amrex::WriteMultiLevelPlotfile, and a bit-for-bit data round trip over
single- and multi-box, single- and two-level files.

These tests read the files back with the writer's own parser; an
independent reader written from the AMReX format checks the same files in
test_plotfile_format.py."""
import os
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from hydroexa_amd import h5input as H
from hydroexa_amd.plotfile import plotfile_name, read_native_plotfile, write_native_plotfile

if not os.path.exists(H._IO_PATH):
    pytest.skip("libhydroexa_io.so not built", allow_module_level=True)


def level(nx, ny, dx, seed):
    r = np.random.default_rng(seed)
    U = torch.as_tensor(r.standard_normal((3, ny, nx)))
    T = torch.as_tensor(r.standard_normal((2, ny, nx)))
    return SimpleNamespace(U_new=U, terrain=T, nx=nx, ny=ny, dx=dx, dy=dx)


def test_name_and_single_level_round_trip(tmp_path):
    assert plotfile_name("plt", 42) == "plt00042"
    lv = level(40, 24, 0.25, 0)
    d = write_native_plotfile(str(tmp_path / "plt"), 7, 1.75, [lv], prob_lo=(10.0, -2.0))
    assert d.endswith("plt00007")
    pf = read_native_plotfile(d)
    assert pf["version"] == "HyperCLaw-V1.1"
    assert pf["varnames"] == ["h_fluid", "hu_momentum", "hv_momentum", "z_bathymetry",
                              "extra_comp_4"]
    assert pf["dim"] == 2 and pf["time"] == 1.75 and pf["finest_level"] == 0
    assert pf["prob_lo"] == [10.0, -2.0] and pf["prob_hi"] == [20.0, 4.0]
    assert pf["domains"] == [(0, 0, 39, 23)] and pf["steps"] == [7]
    assert pf["cell_size"] == [[0.25, 0.25]] and pf["coord"] == 0 and pf["bwidth"] == 0
    L = pf["levels"][0]
    assert (L["vismf_version"], L["how"], L["ncomp"], L["ngrow"]) == (1, 1, 5, 0)
    assert L["extents"] == [(10.0, 20.0, -2.0, 4.0)]
    fab = L["fabs"][0]
    assert fab["fab_header"] == ("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
                                 "((0,0) (39,23) (0,0)) 5")
    full = np.concatenate([lv.U_new.numpy(), lv.terrain.numpy()])
    np.testing.assert_array_equal(fab["data"], full)
    np.testing.assert_allclose(L["min"][0], full.reshape(5, -1).min(1), rtol=1e-14)
    np.testing.assert_allclose(L["max"][0], full.reshape(5, -1).max(1), rtol=1e-14)
    with open(os.path.join(d, "Header")) as f:
        head = f.read().split("\n")
    assert head[12] == "" and head[13] == "((0,0) (39,23) (0,0)) "
    assert head[-2] == "Level_0/Cell"


def test_multi_box_two_level_round_trip(tmp_path):
    c = level(32, 16, 1.0, 1)
    f = level(24, 16, 0.5, 2)
    boxes = [[(0, 0, 15, 15), (16, 0, 31, 15)], [(0, 0, 11, 7), (12, 0, 23, 7), (0, 8, 23, 15)]]
    d = write_native_plotfile(str(tmp_path / "run"), 12, 3.0, [c, f], ref_ratio=[2], boxes=boxes)
    pf = read_native_plotfile(d)
    assert pf["finest_level"] == 1 and pf["ref_ratio"] == [2]
    assert pf["domains"] == [(0, 0, 31, 15), (0, 0, 23, 15)]
    assert pf["cell_size"] == [[1.0, 1.0], [0.5, 0.5]]
    for lv, L, bl in zip((c, f), pf["levels"], boxes):
        full = np.concatenate([lv.U_new.numpy(), lv.terrain.numpy()])
        assert L["boxes"] == [tuple(b) for b in bl]
        for fab, (ilo, jlo, ihi, jhi) in zip(L["fabs"], bl):
            np.testing.assert_array_equal(fab["data"], full[:, jlo:jhi + 1, ilo:ihi + 1])
    # physical extents of the fine boxes on the fine cell size
    assert pf["levels"][1]["extents"][1] == (6.0, 12.0, 0.0, 4.0)
