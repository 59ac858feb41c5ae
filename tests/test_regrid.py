"""Host box layer of the AMR hierarchy (hydroexa_amd/regrid.py) on the CPU.

AMReX's grid generation is absent from the reference tree, so these checks
are the properties its algorithm guarantees: every buffered tag covered,
boxes disjoint, aligned to the blocking factor, inside the domain and the
proper-nesting region, no side above max_grid_size, Berger-Rigoutsos
clusters at or above the efficiency, and the cut rules of FindCut.
"""
import numpy as np
import pytest

from hydroexa_amd import regrid as R


def test_find_cut_rules():
    # a hole: the centre-most zero
    off, st = R.find_cut(np.array([3, 2, 0, 4, 0, 0, 1, 2]))
    assert st == R.HOLE and off == 4
    # no hole, a steep Laplacian sign change
    off, st = R.find_cut(np.array([5, 5, 5, 5, 1, 1, 1, 1]))
    assert st == R.STEEP and off in (3, 4)
    # flat: bisect
    off, st = R.find_cut(np.array([2, 2, 2, 2, 2, 2]))
    assert st == R.BISECT and off == 3


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("eff", [0.5, 0.7, 0.9])
def test_cluster_covers_tags_disjoint_and_efficient(seed, eff):
    rng = np.random.default_rng(seed)
    tags = np.zeros((40, 56), bool)
    for _ in range(5):
        j, i = rng.integers(0, 36), rng.integers(0, 50)
        tags[j:j + rng.integers(1, 6), i:i + rng.integers(1, 8)] = True
    tags |= rng.random(tags.shape) < 0.01
    boxes = R.cluster(tags, eff)
    cov = np.zeros(tags.shape, int)
    for (i0, j0, w, h) in boxes:
        cov[j0:j0 + h, i0:i0 + w] += 1
        e = tags[j0:j0 + h, i0:i0 + w].sum() / float(w * h)
        assert e >= eff or w * h == 1
    assert (cov[tags] == 1).all()
    assert cov.max() <= 1


def test_max_size_chops_by_powers_of_two():
    out = R.max_size([(0, 0, 192, 64), (64, 64, 96, 32)], 64)
    assert all(w <= 64 and h <= 64 for _, _, w, h in out)
    assert all(i % 32 == 0 and j % 32 == 0 and w % 32 == 0 and h % 32 == 0 for i, j, w, h in out)
    cov = R.covers(out, 128, 256)
    assert cov.sum() == 192 * 64 + 96 * 32 == sum(w * h for *_, w, h in out)


def test_erode_keeps_domain_boundary():
    v = np.zeros((10, 10), bool)
    v[0:6, 0:6] = True
    e = R.erode(v, 2)
    assert e[0, 0] and e[3, 3] and not e[4, 3] and not e[3, 4]


@pytest.mark.parametrize("seed", range(4))
def test_new_grids_invariants(seed):
    rng = np.random.default_rng(seed)
    ny, nx = 64, 96
    # level k covers two boxes; tags scattered over the whole level
    valid = R.covers([(0, 0, 48, 64), (48, 16, 32, 32)], ny, nx)
    tags = rng.random((ny, nx)) < 0.02
    bf, ratio, nest, nbuf = 16, 2, 2, 1
    boxes = R.new_grids(tags, valid, ratio, bf, 32, nbuf, 0.7, nest)
    fine_cov = R.covers(boxes, 2 * ny, 2 * nx)
    assert fine_cov.sum() == sum(w * h for *_, w, h in boxes)           # disjoint
    for (i0, j0, w, h) in boxes:
        assert i0 % bf == 0 and j0 % bf == 0 and w % bf == 0 and h % bf == 0
        assert w <= 32 and h <= 32
        assert i0 >= 0 and j0 >= 0 and i0 + w <= 2 * nx and j0 + h <= 2 * ny
    # nested: the coarsened boxes lie in the valid region shrunk by nest
    allowed = R.erode(valid, nest)
    crse_cov = fine_cov.reshape(ny, 2, nx, 2).any(axis=(1, 3))
    assert not (crse_cov & ~allowed).any()
    # every buffered tag in an allowed block is covered
    b = bf // ratio
    grown = R.grow_tags(tags & valid, nbuf) & valid
    blk_ok = allowed.reshape(ny // b, b, nx // b, b).all(axis=(1, 3))
    ok_cells = np.kron(blk_ok, np.ones((b, b), bool)).astype(bool)
    assert not (grown & ok_cells & ~crse_cov).any()


def test_new_grids_empty_without_tags():
    v = np.ones((32, 32), bool)
    assert R.new_grids(np.zeros((32, 32), bool), v, 2, 8, 64, 1, 0.7, 2) == []


@pytest.mark.parametrize("seed", range(4))
def test_native_clustering_equals_the_python_restatement(seed):
    """hx_cluster_blocks (the library's grid generation) gives the boxes of
    the Python restatement, in the same order, on random tag and nesting
    patterns: clustering at three efficiencies, nesting cut-outs, two block
    sizes and three max_grid_size chops."""
    rng = np.random.default_rng(100 + seed)
    for t in range(150):
        ny, nx = rng.integers(1, 70, 2)
        tb = np.zeros((ny, nx), bool)
        for _ in range(rng.integers(0, 6)):
            j, i = rng.integers(0, ny), rng.integers(0, nx)
            tb[j:j + rng.integers(1, 12), i:i + rng.integers(1, 12)] = True
        tb |= rng.random((ny, nx)) < rng.random() * 0.05
        allowed = R.erode(rng.random((ny, nx)) < 0.9, 1) if t % 2 else np.ones((ny, nx), bool)
        b = int(rng.choice([4, 8]))
        eff = float(rng.choice([0.5, 0.7, 0.9]))
        mgs = int(rng.choice([32, 64, 128]))
        assert R.grids_from_blocks(tb, allowed, b, 2, eff, mgs) == \
            R.grids_from_blocks_reference(tb, allowed, b, 2, eff, mgs), (seed, t)
