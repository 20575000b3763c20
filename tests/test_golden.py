"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle
reproduces every stored grid and pair list bit for bit, on CPU; test_gpu_golden.py holds the
HIP side of the same fixtures."""
import glob
import hashlib
import os

import numpy as np
import pytest

import oracle
from node2_reader import Node2Grid

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def pair_digest(gid, ijk, rgba):
    rec = np.zeros((gid.size, 5), np.uint32)
    rec[:, 0] = gid
    rec[:, 1:4] = ijk.view(np.uint32)
    rec[:, 4] = rgba if rgba is not None else 0
    return np.frombuffer(hashlib.sha256(rec.tobytes()).digest(), np.uint8)


def test_fixtures_present():
    assert len(GOLDEN) >= 3


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_reproduces_golden(path):
    g = load(path)
    m, op, q, s, col = oracle.preprocess(g["means"], g["opacity_logits"], g["quats"], g["log_scales"], g["sh0"])
    vs = float(g["voxel_size"])
    grid, st = oracle.raster_node2(m, q, s, col, op, vs, prim_batch_size=int(g["prim_batch_size"]),
                                   ijk_batch_size=int(g["ijk_batch_size"]), value_mode=int(g["value_mode"]))
    assert grid.size == g["grid"].size and np.array_equal(grid, g["grid"])
    assert np.array_equal(np.frombuffer(hashlib.sha256(grid.tobytes()).digest(), np.uint8), g["grid_sha256"])
    assert st["window_count"] == int(g["window_count"])
    gid, ijk, rgba = oracle.pairs(m, q, s, col, op, vs)
    assert gid.size == int(g["pair_count"])
    assert np.array_equal(pair_digest(gid, ijk, rgba), g["pair_sha256"])
    if "pair_gid" in g:
        assert np.array_equal(gid, g["pair_gid"]) and np.array_equal(ijk, g["pair_ijk"])
    Node2Grid(g["grid"]).check_invariants()
