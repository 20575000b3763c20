"""The vectorised node2 checker (node2_fast.FastNode2, used on full-size GPU grids) agrees with the
per-voxel reader (node2_reader.Node2Grid) on oracle grids, and rejects grids with a planted fault
in a prefix lane, a value index, a child mask or the tree header."""
import numpy as np
import pytest

import oracle
from helpers import preprocessed, small_uniform
from nanovdb_editor_amd import synthetic as syn
from node2_fast import FastNode2
from node2_reader import LEAF, LOWER, OFF_CM, OFF_VALUE_IDX, OFF_VMP, Node2Grid


def _grid(kind):
    if kind == "uniform":
        c = small_uniform(400, 0.01, 81)
        kw = {}
    elif kind == "clustered":
        c = syn.clustered_cloud(600, 4, 0.3, 0.004, -5.0, 0.7, seed=82)
        kw = {}
    else:
        c = small_uniform(500, 0.01, 83)
        kw = dict(prim_batch_size=61, ijk_batch_size=30000)
    m, op, q, s, col = preprocessed(c)
    return oracle.raster_node2(m, q, s, col, op, c.voxel_size, **kw)[0]


@pytest.fixture(scope="module", params=["uniform", "clustered", "windowed"])
def grid(request):
    return _grid(request.param)


def test_fast_reader_agrees_with_walk(grid):
    slow = Node2Grid(grid)
    nodes, vox = slow.check_invariants()
    fast = FastNode2(grid)
    summ = fast.check_invariants()
    assert (summ["upper"], summ["lower"], summ["leaf"]) == (slow.count_upper, slow.count_lower, slow.count_leaf)
    assert summ["voxels"] == vox.shape[0]
    assert np.array_equal(summ["bbox"][0], vox[:, :3].min(0)) and np.array_equal(summ["bbox"][1], vox[:, :3].max(0))
    assert np.array_equal(summ["bbox"], slow.bbox())
    # every voxel with its value index, over the whole box and over a sub-box
    lo, hi = vox[:, :3].min(0), vox[:, :3].max(0)
    got = fast.voxel_ijk_in_box(lo, hi)
    key = lambda a: a[np.lexsort((a[:, 2], a[:, 1], a[:, 0]))]
    assert np.array_equal(key(got), key(vox))
    mid = (lo + hi) // 2
    sub = fast.voxel_ijk_in_box(mid - 5, mid + 5)
    inside = vox[np.all(vox[:, :3] >= mid - 5, 1) & np.all(vox[:, :3] <= mid + 5, 1)]
    assert np.array_equal(key(sub), key(inside))
    # node addresses per level equal the walk's
    lv = fast.levels()
    for t in (LOWER, LEAF):
        assert sorted(lv[t]["addr"].tolist()) == sorted(d["addr"] for d in nodes if d["type"] == t)


def _leaf_addr(g):
    return int(FastNode2(g).levels()[LEAF]["addr"][0])


def test_fast_reader_rejects_planted_faults():
    g = _grid("clustered")
    FastNode2(g).check_invariants()
    a = _leaf_addr(g)
    w = lambda b: b.view("<u8")
    # a prefix lane of a leaf value mask
    bad = g.copy()
    w(bad)[a + OFF_VMP[LEAF]] ^= np.uint64(1 << 16)
    with pytest.raises(AssertionError):
        FastNode2(bad).check_invariants()
    # a value index
    bad = g.copy()
    w(bad)[a + OFF_VALUE_IDX] += np.uint64(1)
    with pytest.raises(AssertionError):
        FastNode2(bad).check_invariants()
    # an extra child mask bit in a lower node (its prefix lanes no longer match)
    lo = int(FastNode2(g).levels()[LOWER]["addr"][0])
    bad = g.copy()
    cm = w(bad)[lo + OFF_CM[LOWER]:lo + OFF_CM[LOWER] + 64]
    free = np.nonzero(cm != np.uint64(0xFFFFFFFFFFFFFFFF))[0][0]
    x = int(cm[free])
    inv = ~x & 0xFFFFFFFFFFFFFFFF
    cm[free] = np.uint64(x | (inv & -inv))                  # set the lowest clear bit
    with pytest.raises(AssertionError):
        FastNode2(bad).check_invariants()
    # the leaf count of the tree header
    bad = g.copy()
    bad[672 + 32:672 + 36] = np.frombuffer(np.uint32(int(bad[672 + 32:672 + 36].view("<u4")[0]) + 1).tobytes(),
                                           np.uint8)
    with pytest.raises(AssertionError):
        FastNode2(bad).check_invariants()
