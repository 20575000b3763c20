"""Pins of the oracle's Raster2D tile binning (Raster2D.cpp:347-535) by independent formulations:
brute-force per-tile interval tests, a stable sort by (tile, depth bits), left binary search."""
import numpy as np
import pytest

import oracle


def _cloud(n, width, height, seed, radius_max=40):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-60, width + 60, n), rng.uniform(-60, height + 60, n)], 1).astype(np.float32)
    radii = rng.integers(-3, radius_max, n).astype(np.int32)
    depths = rng.uniform(0.1, 50.0, n).astype(np.float32)
    if n:
        depths[::7] = depths[0]  # equal depths exercise stability
    return means, radii, depths


def _brute(means, radii, depths, width, height, ts):
    tw, th = (width + ts - 1) // ts, (height + ts - 1) // ts
    rows = []
    t32 = np.float32(ts)
    for g in range(radii.size):
        if radii[g] <= 0:
            continue
        r = np.float32(radii[g]) / t32
        u = means[g, 0] / t32
        v = means[g, 1] / t32
        lo_u, hi_u, lo_v, hi_v = np.float32(u - r), np.float32(u + r), np.float32(v - r), np.float32(v + r)
        for ty in range(th):
            # tile row ty spans [ty, ty + 1); the Gaussian's box spans [lo, hi]; ceil() excludes a
            # tile that the box only touches at its left edge
            if not (ty + 1 > lo_v and ty < hi_v):
                continue
            for tx in range(tw):
                if tx + 1 > lo_u and tx < hi_u:
                    rows.append((ty * tw + tx, int(depths[g:g + 1].view(np.uint32)[0]), g))
    return rows, tw * th


@pytest.mark.parametrize("n,w,h,ts,seed", [(0, 64, 64, 16, 1), (1, 640, 480, 16, 2), (300, 640, 480, 16, 3),
                                           (500, 1000, 37, 8, 4), (200, 31, 900, 32, 5), (50, 16, 16, 16, 6)])
def test_tile_bin_matches_brute_force(n, w, h, ts, seed):
    means, radii, depths = _cloud(n, w, h, seed)
    lo, hi, va, offsets = oracle.tile_bin(means, radii, depths, w, h, ts)
    rows, tiles = _brute(means, radii, depths, w, h, ts)
    rows_sorted = sorted(rows, key=lambda r: (r[0], r[1]))   # Timsort is stable: Gaussian order kept
    assert lo.size == len(rows)
    if rows:
        exp = np.array(rows_sorted, np.int64)
        assert np.array_equal(hi.astype(np.int64), exp[:, 0])
        assert np.array_equal(lo.astype(np.int64), exp[:, 1])
        assert np.array_equal(va.astype(np.int64), exp[:, 2])
        assert np.array_equal(offsets, np.searchsorted(exp[:, 0], np.arange(tiles), "left").astype(np.uint32))
    else:
        assert not offsets.any()


def test_tile_bin_clamps_offscreen_and_nan():
    means = np.array([[-1e9, -1e9], [1e9, 5.0], [np.nan, 3.0], [8.0, 8.0]], np.float32)
    radii = np.array([5, 5, 5, 1], np.int32)
    depths = np.ones(4, np.float32)
    lo, hi, va, offsets = oracle.tile_bin(means, radii, depths, 32, 32, 16)
    # only the last Gaussian (pixel (8, 8), radius 1 -> tile box [0, 1) x [0, 1)) intersects; the NaN
    # mean converts to tile 0 on both ends of x and therefore covers no tile
    assert list(va) == [3] and list(hi) == [0]
    assert list(offsets) == [0, 1, 1, 1]
