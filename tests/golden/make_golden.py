"""Writes the golden fixtures of the raster3d path (inputs and expected outputs).

Each fixture is a compressed .npz with the raw Gaussian arrays in the layout the entry point
takes (means, opacity logits, quaternions (w, x, y, z), log-scales, SH0), the build
parameters, and the expected node2 grid bytes and surviving pairs computed by the oracle.
The cases follow the reference's own tests in size and shape (a 32-Gaussian cloud as in
pymodule tests, small clustered clouds, windowed builds with small prim batches).

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "nanovdb-editor_amd")]

import oracle  # noqa: E402
from nanovdb_editor_amd import synthetic as syn  # noqa: E402

CASES = {
    "uniform_32_rgba8": dict(cloud=lambda: syn.uniform_cloud(32, 0.02, -3.5, 0.4, 101), value_mode=0,
                             prim_batch_size=0, ijk_batch_size=0),
    "clustered_600_rgba8_windowed": dict(cloud=lambda: syn.clustered_cloud(600, 6, 0.15, 0.005, -5.0, 0.7, 102),
                                         value_mode=0, prim_batch_size=128, ijk_batch_size=60_000),
    "uniform_150_density": dict(cloud=lambda: syn.uniform_cloud(150, 0.01, -4.5, 0.5, 103), value_mode=1,
                                prim_batch_size=0, ijk_batch_size=0),
}


def pair_digest(gid, ijk, rgba):
    """sha256 over the pairs in fragment order: gid (u32), ijk (3 x i32), rgba (u32) per pair."""
    rec = np.zeros((gid.size, 5), np.uint32)
    rec[:, 0] = gid
    rec[:, 1:4] = ijk.view(np.uint32)
    rec[:, 4] = rgba if rgba is not None else 0
    return np.frombuffer(hashlib.sha256(rec.tobytes()).digest(), np.uint8)


def make(name, spec):
    c = spec["cloud"]()
    m, op, q, s, col = oracle.preprocess(c.means, c.opacity_logits, c.quats, c.log_scales, c.sh0)
    grid, st = oracle.raster_node2(m, q, s, col, op, c.voxel_size, prim_batch_size=spec["prim_batch_size"],
                                   ijk_batch_size=spec["ijk_batch_size"], value_mode=spec["value_mode"])
    gid, ijk, rgba = oracle.pairs(m, q, s, col, op, c.voxel_size)
    path = os.path.join(HERE, name + ".npz")
    extra = {}
    if gid.size <= 50_000:  # the full pair list only for the small case; a digest otherwise
        extra = dict(pair_gid=gid, pair_ijk=ijk, pair_rgba=rgba)
    np.savez_compressed(path, means=c.means, opacity_logits=c.opacity_logits, quats=c.quats,
                        log_scales=c.log_scales, sh0=c.sh0, voxel_size=np.float32(c.voxel_size),
                        value_mode=np.int32(spec["value_mode"]), prim_batch_size=np.uint32(spec["prim_batch_size"]),
                        ijk_batch_size=np.uint32(spec["ijk_batch_size"]), grid=grid,
                        pair_count=np.uint64(gid.size), pair_sha256=pair_digest(gid, ijk, rgba),
                        window_count=np.uint64(st["window_count"]),
                        grid_sha256=np.frombuffer(hashlib.sha256(grid.tobytes()).digest(), np.uint8), **extra)
    print(f"{name}: {c.count} Gaussians, {gid.size} pairs, grid {grid.size} B, windows {st['window_count']}")


if __name__ == "__main__":
    for name, spec in CASES.items():
        make(name, spec)
