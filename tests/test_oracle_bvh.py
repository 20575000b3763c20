"""Pins of the oracle's VoxelBVH Gaussian keys (VoxelBVH.cpp:1225-1435) and of its 4x4 inverse
(Camera.h:316-375) by independent formulations: numpy's inverse, closed-form boxes of
axis-aligned Gaussians, a float64 ellipsoid/box distance test per cell, and np.unique runs."""
import numpy as np
import pytest

import oracle

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


def test_mat_inverse_matches_numpy():
    rng = np.random.default_rng(1)
    for _ in range(20):
        m = rng.standard_normal((4, 4)).astype(np.float32) + 3 * np.eye(4, dtype=np.float32)
        got = oracle.mat_inverse(m)
        ref = np.linalg.inv(m.astype(np.float64))
        assert np.allclose(got, ref, rtol=1e-4, atol=1e-5)
        assert np.allclose(got.astype(np.float64) @ m.astype(np.float64), np.eye(4), atol=1e-5)


def _axis_aligned(n, seed, spread=1.0, opacity_min=0.0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    scales = np.exp(rng.normal(-3.0, 0.6, (n, 3))).astype(np.float32)
    opac = rng.uniform(opacity_min, 1.0, n).astype(np.float32)
    quats = np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1))
    return means, opac, quats, scales


def _radius(opac):
    return np.sqrt(2.0 * np.log(np.maximum(0.01, opac.astype(np.float64)) / 0.01))


def test_world_box_closed_form():
    means, opac, quats, scales = _axis_aligned(300, 2)
    _, _, _, bbox = oracle.voxelbvh_ijkl(means, opac, quats, scales, 256)
    ext = scales.astype(np.float64) * _radius(opac)[:, None]
    assert np.allclose(bbox[:3], (means - ext).min(0), rtol=1e-6, atol=1e-6)
    assert np.allclose(bbox[3:], (means + ext).max(0), rtol=1e-6, atol=1e-6)
    # translation + uniform scale transform (rows; translation in the last row)
    xf = np.diag([2.0, 2.0, 2.0, 1.0]).astype(np.float32)
    xf[3, :3] = (1.0, -3.0, 0.5)
    _, _, _, b2 = oracle.voxelbvh_ijkl(means, opac, quats, scales, 256, xf)
    assert np.allclose(b2[:3], 2 * bbox[:3].astype(np.float64) + xf[3, :3], rtol=1e-6, atol=1e-6)
    assert np.allclose(b2[3:], 2 * bbox[3:].astype(np.float64) + xf[3, :3], rtol=1e-6, atol=1e-6)


def _cells_f64(mean, scale, r, bbox, res, level, lo, hi):
    """Cells of `level` in the aligned range [lo, hi] whose box meets the axis-aligned ellipsoid
    |(x - mean) / (r scale)| < 1, with the distance evaluated in float64; also the smallest
    |1 - distance^2| seen (to skip cases decided by rounding)."""
    d = bbox[3:].astype(np.float64) - bbox[:3]
    sc = (res - 1) / d.max()
    off = -bbox[:3].astype(np.float64) * sc + 0.5
    step = 1 << level
    msk = ~(step - 1)
    cells, margin = [], np.inf
    for i in range(lo[0] & msk, (hi[0] & msk) + 1, step):
        for j in range(lo[1] & msk, (hi[1] & msk) + 1, step):
            for k in range(lo[2] & msk, (hi[2] & msk) + 1, step):
                cmin = (np.array([i, j, k], np.float64) - off) / sc
                cmax = (np.array([i, j, k], np.float64) + step - off) / sc
                near = np.clip(mean, cmin, cmax)
                dist2 = (((near - mean) / (r * scale)) ** 2).sum()
                margin = min(margin, abs(1.0 - dist2))
                if dist2 < 1.0:
                    cells.append((i, j, k))
    return cells, margin


@pytest.mark.parametrize("res,seed", [(64, 3), (512, 4), (4096, 5)])
def test_keys_match_float64_ellipsoid_box_test(res, seed):
    # opacities above 0.01 (at or below it the cutoff radius is 0 and the local frame degenerates)
    means, opac, quats, scales = _axis_aligned(60, seed, spread=0.3, opacity_min=0.02)
    keys, ids, ranges, bbox = oracle.voxelbvh_ijkl(means, opac, quats, scales, res)
    d = bbox[3:].astype(np.float64) - bbox[:3]
    sc = (res - 1) / d.max()
    off = -bbox[:3].astype(np.float64) * sc + 0.5
    checked = 0
    for g in range(opac.size):
        r = _radius(opac[g:g + 1])[0]
        m = means[g].astype(np.float64)
        s = scales[g].astype(np.float64)
        lo = np.floor(sc * (m - r * s) + off).astype(int)
        hi = np.floor(sc * (m + r * s) + off).astype(int)
        # level rule: from 11 down, stop before the first level with more than 8 overlapping cells
        level, last, ok = 11, 11, True
        while True:
            cells, margin = _cells_f64(m, s, r, bbox, res, level, lo, hi)
            ok = ok and margin > 1e-4
            if len(cells) > 8:
                level = last
                break
            if level == 0:
                break
            last, level = level, level - 1
        cells, margin = _cells_f64(m, s, r, bbox, res, level, lo, hi)
        if not ok or margin <= 1e-4:
            continue
        exp = sorted((i << 48) | (j << 32) | (k << 16) | level for i, j, k in cells[:8])
        got = sorted(int(k) for k in keys[ids == g])
        assert got == exp, g
        checked += 1
    assert checked >= 40


def test_sorted_order_and_ranges():
    rng = np.random.default_rng(7)
    n = 400
    means = rng.normal(0, 0.2, (n, 3)).astype(np.float32)
    q = rng.standard_normal((n, 4)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scales = np.exp(rng.normal(-3.5, 0.5, (n, 3))).astype(np.float32)
    opac = rng.uniform(0.0, 1.0, n).astype(np.float32)
    keys, ids, ranges, _ = oracle.voxelbvh_ijkl(means, opac, q, scales, 1024)
    assert np.all(keys[1:] >= keys[:-1])
    same = keys[1:] == keys[:-1]
    assert np.all(ids[1:][same] >= ids[:-1][same])               # stable: Gaussian order kept
    assert np.all((keys == NONE) == (ids == np.uint32(0xFFFFFFFF)))
    assert np.all(np.bincount(ids[ids != 0xFFFFFFFF], minlength=n) >= 1)   # every Gaussian keeps a cell
    u, first, counts = np.unique(keys, return_index=True, return_counts=True)
    assert np.array_equal(ranges[:u.size, 0], first)
    assert np.array_equal(ranges[:u.size, 1], first + counts)
    assert not ranges[u.size:].any()


def test_resolution_bounds():
    means, opac, quats, scales = _axis_aligned(4, 1)
    with pytest.raises(ValueError):
        oracle.voxelbvh_ijkl(means, opac, quats, scales, 4097)
    with pytest.raises(ValueError):
        oracle.voxelbvh_ijkl(means, opac, quats, scales, 0)
