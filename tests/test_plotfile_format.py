"""Native plotfiles read back by an independent reader written from the AMReX
plotfile format (HyperCLaw-V1.1 Header, VisMF version-1 ``Cell_H``, FAB data
files with their RealDescriptor), not by the writer's own parser
(hydroexa_amd.plotfile.read_native_plotfile).

The reader exposes what the reference's analysis scripts take from yt's BoxLib
frontend: ``field_list`` of ("boxlib", name), ``current_time``, the domain
edges and dimensions, and per grid its level, left and right edges, cell size
and data indexed [i, j] (x first).  The checks then do what those scripts do:
resolve h, hu, hv and z through their alias tables
(tests/dambreak/plot_output.py:28-52, PyHydroEXA/src/pyhydroexa/io/fields.py:26-97),
extract a centreline profile from the finest level covering each stretch
(tests/dambreak/plot_output.py:145-190) and a level-0 covering grid
(fields.py:100-166), and locate run directories by their ``plt#####`` names
(tests/dambreak/analyze.py:61-75).  Every written value is a dyadic function
of its cell centre, so a misplaced box, a transposed FAB, a swapped
component or a wrong byte order changes what the reader returns; one test
plants each such mistake in a copy of the files and checks that the reader
sees it.  Writer: src/io/NativePlotFile.cpp:19-97."""
import glob
import os
import re
import shutil
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from hydroexa_amd import h5input as H
from hydroexa_amd.plotfile import write_native_plotfile

if not os.path.exists(H._IO_PATH):
    pytest.skip("libhydroexa_io.so not built", allow_module_level=True)


# --------------------------------------------------------------------------
# reader, from the format
# --------------------------------------------------------------------------

_BOX = re.compile(r"\(\((-?\d+),(-?\d+)\) \((-?\d+),(-?\d+)\) \((\d+),(\d+)\)\)")
_FAB = re.compile(r"^FAB \(\((\d+), \(([0-9 ]+)\)\),\((\d+), \(([0-9 ]+)\)\)\)"
                  r"(\(\(-?\d+,-?\d+\) \(-?\d+,-?\d+\) \(\d+,\d+\)\)) (\d+)$")
IEEE_BINARY64 = [64, 11, 52, 0, 1, 12, 0, 1023]   # bits, exp bits, mantissa bits,
                                                  # sign/exp/mantissa start bits, bias


def parse_box(text):
    m = _BOX.fullmatch(text.strip())
    assert m, text
    lo = (int(m.group(1)), int(m.group(2)))
    hi = (int(m.group(3)), int(m.group(4)))
    assert (m.group(5), m.group(6)) == ("0", "0"), "plotfile boxes are cell centred"
    return lo, hi


def fab_dtype(fmt, order):
    """numpy dtype of a FAB from its RealDescriptor: the format must describe
    IEEE binary64 and the byte order lists, per byte of a value as stored,
    its significance (1 = most significant)."""
    fmt = [int(v) for v in fmt.split()]
    order = [int(v) for v in order.split()]
    assert fmt == IEEE_BINARY64, fmt
    if order == list(range(1, 9)):
        return np.dtype(">f8")
    if order == list(range(8, 0, -1)):
        return np.dtype("<f8")
    raise AssertionError(f"unsupported byte order {order}")


class Grid:
    def __init__(self, ds, level, left, right, lo, hi, fields):
        self.Level, self.LeftEdge, self.RightEdge = level, left, right
        self.lo, self.hi = lo, hi
        self.ActiveDimensions = (hi[0] - lo[0] + 1, hi[1] - lo[1] + 1)
        self.dds = ds.level_dds[level]
        self._fields = fields

    def __getitem__(self, field):
        return self._fields[field[1]]


class BoxlibDataset:
    """A plotfile directory, parsed token by token from the Header and per
    level from Cell_H and the FAB data files."""

    def __init__(self, path):
        self.path = path
        with open(os.path.join(path, "Header")) as f:
            lines = f.read().split("\n")
        pos = 0

        def line():
            nonlocal pos
            pos += 1
            return lines[pos - 1]

        assert line() == "HyperCLaw-V1.1"
        nvar = int(line())
        names = [line() for _ in range(nvar)]
        self.field_list = [("boxlib", n) for n in names]
        self.dimensionality = int(line())
        assert self.dimensionality == 2
        self.current_time = float(line())
        self.max_level = int(line())
        self.domain_left_edge = np.array(line().split(), dtype=float)
        self.domain_right_edge = np.array(line().split(), dtype=float)
        self.refine_by = [int(v) for v in line().split()]
        assert len(self.refine_by) == self.max_level
        self.level_domains = [parse_box(m.group(0)) for m in _BOX.finditer(line())]
        assert len(self.level_domains) == self.max_level + 1
        self.level_steps = [int(v) for v in line().split()]
        self.level_dds = [np.array(line().split(), dtype=float) for _ in range(self.max_level + 1)]
        self.geometry = int(line())
        self.boundary_width = int(line())
        (lo, hi) = self.level_domains[0]
        self.domain_dimensions = np.array([hi[0] - lo[0] + 1, hi[1] - lo[1] + 1])
        self.grids = []
        for lev in range(self.max_level + 1):
            tok = line().split()
            assert int(tok[0]) == lev
            ngrids, t = int(tok[1]), float(tok[2])
            assert t == self.current_time
            assert int(line()) == self.level_steps[lev]
            edges = []
            for _g in range(ngrids):
                x0, x1 = (float(v) for v in line().split())
                y0, y1 = (float(v) for v in line().split())
                edges.append((np.array([x0, y0]), np.array([x1, y1])))
            rel = line()
            assert rel == f"Level_{lev}/Cell"
            self.grids += self._read_level(lev, os.path.join(path, rel), names, edges)

    def _read_level(self, lev, prefix, names, edges):
        with open(prefix + "_H") as f:
            toks = f.read().split("\n")
        version, how, ncomp, ngrow = (int(toks[k]) for k in range(4))
        assert (version, how, ngrow) == (1, 1, 0) and ncomp == len(names)
        head = toks[4].split()
        nbox = int(head[0].lstrip("("))
        assert head[1] == "0"
        boxes = [parse_box(toks[5 + b]) for b in range(nbox)]
        k = 5 + nbox
        assert toks[k] == ")"
        assert int(toks[k + 1]) == nbox
        fods = []
        for b in range(nbox):
            tag, fname, off = toks[k + 2 + b].split()
            assert tag == "FabOnDisk:"
            fods.append((fname, int(off)))
        k += 2 + nbox
        assert toks[k] == ""
        minmax = []
        for _pass in range(2):
            assert toks[k + 1] == f"{nbox},{ncomp}"
            rows = [toks[k + 2 + b] for b in range(nbox)]
            assert all(r.endswith(",") for r in rows)
            minmax.append(np.array([[float(v) for v in r[:-1].split(",")] for r in rows]))
            k += 2 + nbox
            assert toks[k] == ""
        grids = []
        dx = self.level_dds[lev]
        dlo = np.array(self.level_domains[lev][0])
        for b, ((fname, off), (lo, hi), (left, right)) in enumerate(zip(fods, boxes, edges)):
            with open(os.path.join(os.path.dirname(prefix), fname), "rb") as f:
                f.seek(off)
                m = _FAB.match(f.readline().decode().rstrip("\n"))
                assert m, "FAB header"
                assert int(m.group(1)) == 8 and int(m.group(3)) == 8
                dt = fab_dtype(m.group(2), m.group(4))
                assert parse_box(m.group(5)) == (lo, hi)
                assert int(m.group(6)) == ncomp
                nx, ny = hi[0] - lo[0] + 1, hi[1] - lo[1] + 1
                raw = np.frombuffer(f.read(dt.itemsize * ncomp * nx * ny), dtype=dt)
            assert raw.size == ncomp * nx * ny
            # Fortran order within a FAB: i fastest, then j, then the component
            arr = raw.astype(np.float64).reshape((nx, ny, ncomp), order="F")
            # the Header's physical extents place the box on its level
            np.testing.assert_array_equal(left, self.domain_left_edge + (np.array(lo) - dlo) * dx)
            np.testing.assert_array_equal(right, self.domain_left_edge + (np.array(hi) + 1 - dlo) * dx)
            np.testing.assert_array_equal(minmax[0][b], arr.reshape(-1, ncomp).min(0))
            np.testing.assert_array_equal(minmax[1][b], arr.reshape(-1, ncomp).max(0))
            grids.append(Grid(self, lev, left, right, lo, hi,
                              {n: arr[:, :, c] for c, n in enumerate(names)}))
        return grids

    def covering_grid(self, level):
        """Level ``level`` over its whole domain, [i, j]; cells no box covers are NaN."""
        (lo, hi) = self.level_domains[level]
        out = np.full((hi[0] - lo[0] + 1, hi[1] - lo[1] + 1), np.nan)
        res = {}
        for g in self.grids:
            if g.Level != level:
                continue
            for name in (f[1] for f in self.field_list):
                a = res.setdefault(name, out.copy())
                a[g.lo[0] - lo[0]:g.hi[0] - lo[0] + 1, g.lo[1] - lo[1]:g.hi[1] - lo[1] + 1] = g[("boxlib", name)]
        return res


# --------------------------------------------------------------------------
# what the reference's analysis does with a dataset
# --------------------------------------------------------------------------

# logical field -> names tried in turn (tests/dambreak/plot_output.py:28-33)
ANALYSIS_ALIASES = {
    "h": ["h_fluid", "h", "water_depth", "Depth"],
    "hu": ["hu", "hu_momentum", "MomentumX"],
    "hv": ["hv", "hv_momentum", "MomentumY"],
    "z": ["z_bathymetry", "z", "bathymetry", "BedElevation", "elev"],
}


def find_variable(ds, key):
    """Exact names first, then the first field whose name contains one
    (tests/dambreak/plot_output.py:42-52)."""
    for alias in ANALYSIS_ALIASES[key]:
        if ("boxlib", alias) in ds.field_list:
            return ("boxlib", alias)
    for f in ds.field_list:
        if any(alias in f[1] for alias in ANALYSIS_ALIASES[key]):
            return f
    return None


def centreline(ds, field, y):
    """Values along y = const, each x stretch from the finest grid that spans
    it (the rule of tests/dambreak/plot_output.py:145-190)."""
    segs = []
    for g in ds.grids:
        if not g.LeftEdge[1] <= y <= g.RightEdge[1]:
            continue
        jy = min(int(np.floor((y - g.LeftEdge[1]) / g.dds[1])), g.ActiveDimensions[1] - 1)
        xc = g.LeftEdge[0] + (np.arange(g.ActiveDimensions[0]) + 0.5) * g.dds[0]
        segs.append((g.Level, g.LeftEdge[0], g.RightEdge[0], xc, g[field][:, jy]))
    cuts = np.unique(np.concatenate([[s[1], s[2]] for s in segs]))
    xs, vs, levs = [], [], []
    for a, b in zip(cuts[:-1], cuts[1:]):
        best = max((s for s in segs if s[1] <= a and s[2] >= b), key=lambda s: s[0])
        keep = (best[3] > a) & (best[3] < b)
        xs.append(best[3][keep])
        vs.append(best[4][keep])
        levs.append(np.full(int(keep.sum()), best[0]))
    return np.concatenate(xs), np.concatenate(vs), np.concatenate(levs)


# --------------------------------------------------------------------------
# dyadic fields of the cell centre
# --------------------------------------------------------------------------

PROB_LO = (-8.0, 2.0)


def field_values(nx, ny, dx, dy, lev):
    """(5, ny, nx): component c at cell (i, j) = a distinct dyadic function of
    the cell centre (x, y), exact in binary64."""
    x = PROB_LO[0] + (np.arange(nx) + 0.5) * dx
    y = PROB_LO[1] + (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    return np.stack([1000.0 * c + 64.0 * lev + X + 256.0 * Y for c in range(5)])


def make_level(nx, ny, dx, lev):
    v = torch.as_tensor(field_values(nx, ny, dx, dx, lev))
    return SimpleNamespace(U_new=v[:3], terrain=v[3:], nx=nx, ny=ny, dx=dx, dy=dx)


def centre_value(c, lev, x, y):
    return 1000.0 * c + 64.0 * lev + x + 256.0 * y


def two_level_file(tmp_path):
    crse = make_level(32, 16, 0.5, 0)
    fine = make_level(64, 32, 0.25, 1)
    boxes = [[(0, 0, 15, 15), (16, 0, 31, 15)],
             [(20, 8, 35, 23), (36, 8, 43, 15), (36, 16, 43, 23)]]
    return write_native_plotfile(str(tmp_path / "plt"), 120, 2.5, [crse, fine],
                                 prob_lo=PROB_LO, ref_ratio=[2], boxes=boxes), boxes


def test_reader_sees_the_reference_layout_and_names(tmp_path):
    d, boxes = two_level_file(tmp_path)
    assert glob.glob(os.path.join(str(tmp_path), "plt[0-9]*")) == [d]
    ds = BoxlibDataset(d)
    assert ds.current_time == 2.5 and ds.max_level == 1 and ds.refine_by == [2]
    assert ds.level_steps == [120, 120]
    assert ds.level_domains == [((0, 0), (31, 15)), ((0, 0), (63, 31))]
    np.testing.assert_array_equal(ds.domain_left_edge, PROB_LO)
    np.testing.assert_array_equal(ds.domain_right_edge, [8.0, 10.0])
    np.testing.assert_array_equal(ds.domain_dimensions, [32, 16])
    np.testing.assert_array_equal(ds.level_dds[1], [0.25, 0.25])
    assert (ds.geometry, ds.boundary_width) == (0, 0)
    # the names the analysis resolves, and the component each one carries
    for key, name, comp in (("h", "h_fluid", 0), ("hu", "hu_momentum", 1),
                            ("hv", "hv_momentum", 2), ("z", "z_bathymetry", 3)):
        f = find_variable(ds, key)
        assert f == ("boxlib", name), (key, f)
        for g in ds.grids:
            xc = g.LeftEdge[0] + (np.arange(g.ActiveDimensions[0]) + 0.5) * g.dds[0]
            yc = g.LeftEdge[1] + (np.arange(g.ActiveDimensions[1]) + 0.5) * g.dds[1]
            np.testing.assert_array_equal(g[f], centre_value(comp, g.Level, xc[:, None], yc[None, :]))
    assert ds.field_list[4] == ("boxlib", "extra_comp_4")
    assert [(g.lo, g.hi) for g in ds.grids if g.Level == 1] == [((a, b), (c, e)) for a, b, c, e in boxes[1]]


def test_covering_grid_and_centreline_as_the_analysis_reads_them(tmp_path):
    d, _ = two_level_file(tmp_path)
    ds = BoxlibDataset(d)
    full = ds.covering_grid(0)
    h = full[find_variable(ds, "h")[1]]
    assert h.shape == (32, 16) and not np.isnan(h).any()
    np.testing.assert_array_equal(h.T, field_values(32, 16, 0.5, 0.5, 0)[0])   # (ny, nx) view
    fine = ds.covering_grid(1)["h_fluid"]
    covered = ~np.isnan(fine)
    assert covered.sum() == 16 * 16 + 8 * 8 + 8 * 8
    # a line through the refined patch: coarse, then fine, then coarse
    y = 5.1
    x, v, lev = centreline(ds, find_variable(ds, "h"), y)
    assert np.all(np.diff(x) > 0)
    assert x[0] == PROB_LO[0] + 0.25 and x[-1] == 8.0 - 0.25
    assert list(np.unique(lev)) == [0, 1]
    fine_x = x[lev == 1]
    np.testing.assert_array_equal(fine_x, PROB_LO[0] + (np.arange(20, 44) + 0.5) * 0.25)
    for L in (0, 1):
        dy = 0.5 / (1 + L)
        yc = PROB_LO[1] + (np.floor((y - PROB_LO[1]) / dy) + 0.5) * dy
        np.testing.assert_array_equal(v[lev == L], centre_value(0, L, x[lev == L], yc))


def _plant(src, dst, edit):
    shutil.copytree(src, dst)
    edit(dst)
    return dst


@pytest.mark.parametrize("mistake", ["byte_order", "transposed_fab", "shifted_extent",
                                     "swapped_names", "stale_max"])
def test_reader_catches_planted_layout_mistakes(tmp_path, mistake):
    """Each plausible writer mistake, planted in a copy of a good file, is
    either rejected by the reader or changes what it returns."""
    d, _ = two_level_file(tmp_path)
    good = BoxlibDataset(d)
    cd = os.path.join(d, "Level_0", "Cell_D_00000")

    def byte_order(p):
        path = os.path.join(p, "Level_0", "Cell_D_00000")
        b = open(path, "rb").read().replace(b"(8 7 6 5 4 3 2 1)", b"(1 2 3 4 5 6 7 8)")
        open(path, "wb").write(b)

    def transposed_fab(p):
        # write box 0 with j fastest instead of i (equal cell count)
        path = os.path.join(p, "Level_0", "Cell_D_00000")
        raw = open(cd, "rb").read()
        nl = raw.index(b"\n") + 1
        a = np.frombuffer(raw[nl:nl + 8 * 5 * 256], dtype="<f8").reshape(5, 16, 16)
        open(path, "wb").write(raw[:nl] + np.ascontiguousarray(a.transpose(0, 2, 1)).tobytes()
                               + raw[nl + 8 * 5 * 256:])

    def shifted_extent(p):
        path = os.path.join(p, "Header")
        t = open(path).read().split("\n")
        k = t.index("Level_1/Cell") - 6                # the first fine box's x extent
        x0, x1 = (float(v) for v in t[k].split())
        t[k] = f"{x0 + 0.25:.17g} {x1 + 0.25:.17g}"
        open(path, "w").write("\n".join(t))

    def swapped_names(p):
        path = os.path.join(p, "Header")
        t = open(path).read().replace("hu_momentum", "@@").replace("hv_momentum", "hu_momentum")
        open(path, "w").write(t.replace("@@", "hv_momentum"))

    def stale_max(p):
        path = os.path.join(p, "Level_0", "Cell_H")
        t = open(path).read().split("\n")
        k = len(t) - 3                                 # last max row
        vals = t[k].rstrip(",").split(",")
        vals[0] = f"{float(vals[0]) + 1.0:.15e}"
        t[k] = ",".join(vals) + ","
        open(path, "w").write("\n".join(t))

    bad = _plant(d, str(tmp_path / "bad"), locals()[mistake])
    try:
        ds = BoxlibDataset(bad)
    except AssertionError:
        return                                          # rejected by the reader
    if mistake == "swapped_names":
        hu = find_variable(ds, "hu")
        g = ds.grids[0]
        assert not np.array_equal(g[hu], good.grids[0][("boxlib", "hu_momentum")])
        return
    changed = any(not np.array_equal(a[("boxlib", "h_fluid")], b[("boxlib", "h_fluid")])
                  for a, b in zip(ds.grids, good.grids))
    assert changed, mistake
