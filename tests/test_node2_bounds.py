"""The bounded node2 descent (vdb_node2.h: rank < fanout, child node inside the grid, key table
inside the root), compiled for the host with AddressSanitizer + UndefinedBehaviorSanitizer, on
oracle grids: every active voxel is found at its value index on an intact grid, and grids with a
planted fault (truncated, a child offset past the end, a corrupt child count or prefix lane) are
reported as corrupt without any read outside the grid."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import preprocessed, small_clustered
from node2_fast import FastNode2
from node2_reader import LOWER, OFF_CHILDREN, OFF_CMP, ROOT, UPPER

ROOT_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT_DIR, "nanovdb-editor_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = tmp_path_factory.mktemp("n2b") / "node2_bounds_check"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", CSRC,
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    os.path.join(ROOT_DIR, "tests", "native", "node2_bounds_check.cpp"), "-o", str(exe)], check=True)
    return str(exe)


@pytest.fixture(scope="module")
def grid_and_queries(tmp_path_factory):
    c = small_clustered(3000, 0.005, 95)
    m, op, q, s, col = preprocessed(c)
    grid, _ = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    ijk, vidx = oracle.node2_active_voxels(grid)
    rng = np.random.default_rng(1)
    lo, hi = ijk.min(0) - 40, ijk.max(0) + 40
    extra = rng.integers(lo, hi + 1, (20000, 3)).astype(np.int32)
    queries = np.concatenate([ijk.astype(np.int32), extra])
    d = tmp_path_factory.mktemp("n2q")
    qpath = str(d / "q.bin")
    queries.tofile(qpath)
    return grid, ijk, vidx, qpath, d


def _run(harness, grid_bytes, qpath, d, name):
    path = str(d / f"{name}.node2")
    np.asarray(grid_bytes, np.uint8).tofile(path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    out = subprocess.run([harness, path, qpath], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "AddressSanitizer" not in out.stderr and "runtime error" not in out.stderr, out.stderr
    f = out.stdout.split()
    return dict(found=int(f[1]), bad=int(f[3]), sum=int(f[5]))


def test_intact_grid(harness, grid_and_queries):
    grid, ijk, vidx, qpath, d = grid_and_queries
    r = _run(harness, grid, qpath, d, "intact")
    assert r["bad"] == 0
    assert r["found"] >= ijk.shape[0] and r["sum"] >= int(vidx.sum())


def test_faults_are_rejected_in_bounds(harness, grid_and_queries):
    grid, ijk, vidx, qpath, d = grid_and_queries
    lv = FastNode2(grid).levels()
    w = lambda b: b.view("<u8")
    cases = {}
    # truncated in the middle of the lower nodes: the root survives, half the lowers and every leaf
    # are gone
    cut = int(np.sort(lv[LOWER]["addr"])[lv[LOWER]["addr"].size // 2]) * 8 + 1000
    assert cut < grid.size
    cases["truncated"] = grid[:cut // 8 * 8].copy()
    # a lower node's first child offset far past the end
    b = grid.copy()
    lo = int(lv[LOWER]["addr"][0])
    w(b)[lo + OFF_CHILDREN[LOWER]] = np.uint64(0xFFFFFFFFFFFFFFF8)
    cases["child_past_end"] = b
    # the root's child count lane at its 16-bit maximum: the key table would leave the root
    b = grid.copy()
    root = int(lv[ROOT]["addr"][0])
    w(b)[root + OFF_CMP[ROOT]] |= np.uint64(0xFFFF)
    cases["child_count"] = b
    # an upper node's prefix lanes at their maximum: child ranks beyond the fanout
    b = grid.copy()
    up = int(lv[UPPER]["addr"][0])
    w(b)[up + OFF_CMP[UPPER]:up + OFF_CMP[UPPER] + 128] = np.uint64(0xFFFFFFFFFFFFFFFF)
    cases["prefix_lanes"] = b
    for name, g in cases.items():
        r = _run(harness, g, qpath, d, name)
        assert r["bad"] > 0, name
