"""GPU parity of the raster3d pipeline against the CPU oracle (node2 bytes, pairs, values)."""
import numpy as np
import pytest

import oracle
from helpers import preprocessed, small_clustered, small_uniform, sort_pairs

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32).ravel()).cuda()


def _run_gpu(vox, cloud, raw=True, **kw):
    means = _dev(cloud.means)
    op = _dev(cloud.opacity_logits)
    q = _dev(cloud.quats)
    s = _dev(cloud.log_scales)
    sh0 = _dev(cloud.sh0)
    colors = torch.empty_like(sh0)
    grid, st = vox.raster_node2(means, q, s, colors, op, cloud.voxel_size, raw=raw, sh0=sh0, **kw)
    return grid.cpu().numpy(), st, (op.cpu().numpy(), q.cpu().numpy(), s.cpu().numpy(), colors.cpu().numpy())


@pytest.mark.parametrize("n,vs,seed,W", [(800, 0.01, 3, 0), (1500, 0.01, 5, 0), (1500, 0.01, 5, 200_000),
                                         (3000, 0.005, 7, 0), (3000, 0.005, 7, 1_000_003)])
def test_node2_bytes_match_oracle_uniform(vox, n, vs, seed, W):
    c = small_uniform(n, vs, seed)
    m, op, q, s, col = preprocessed(c)
    ref, rst = oracle.raster_node2(m, q, s, col, op, vs, ijk_batch_size=W)
    got, st, _ = _run_gpu(vox, c, ijk_batch_size=W)
    print(f"case n={n} vs={vs} seed={seed} W={W}", rst, st, flush=True)
    assert st["pair_count"] == rst["pair_count"]
    assert st["fragment_count"] == rst["fragment_count"]
    assert (st["upper_count"], st["lower_count"], st["leaf_count"]) == (rst["upper_count"], rst["lower_count"],
                                                                       rst["leaf_count"])
    assert got.size == ref.size
    diff = np.nonzero(got != ref)[0]
    assert diff.size == 0, f"{diff.size} differing bytes, first at {diff[:8]}"


def test_node2_bytes_match_oracle_clustered(vox):
    c = small_clustered(4000, 0.005, 11)
    m, op, q, s, col = preprocessed(c)
    ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    got, st, _ = _run_gpu(vox, c)
    assert got.size == ref.size
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("value_mode", [0, 1])
def test_segment_record_forms_agree(vox, value_mode):
    # the emission's 8-byte candidate-segment records (the default where the voxel key, the
    # Gaussian id and the count fit 64 bits) and the 16-byte form give the same grid bytes
    # (RGBA8 maxima and exact density sums do not depend on the pair order); RGBA8 equals the oracle
    c = small_clustered(4000, 0.005, 13)
    compact, st, _ = _run_gpu(vox, c, value_mode=value_mode)
    wide, wst, _ = _run_gpu(vox, c, value_mode=value_mode, wide_segments=True)
    print(f"segment records mode={value_mode}", st, wst, flush=True)
    assert st["pair_count"] == wst["pair_count"] and st["voxel_count"] == wst["voxel_count"]
    assert np.array_equal(compact, wide)
    if value_mode == 0:
        m, op, q, s, col = preprocessed(c)
        ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
        assert st["pair_count"] == rst["pair_count"]
        assert np.array_equal(compact, ref)


def test_small_prim_batches_and_windows(vox):
    # many prim batches and windows: exercises the windowed node order across batches
    c = small_uniform(1200, 0.01, 21)
    m, op, q, s, col = preprocessed(c)
    ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size, prim_batch_size=300, ijk_batch_size=500_000)
    got, st, _ = _run_gpu(vox, c, prim_batch_size=300, ijk_batch_size=500_000)
    assert st["window_count"] == rst["window_count"]
    assert np.array_equal(got, ref)


def test_preprocessing_in_place_matches_oracle(vox):
    c = small_uniform(5000, 0.01, 9)
    m, op, q, s, col = preprocessed(c)
    _, _, (gop, gq, gs, gcol) = _run_gpu(vox, c)
    for got, ref in ((gop, op), (gq, q), (gs, s), (gcol, col)):
        assert np.array_equal(got.ravel().view(np.uint32), ref.ravel().view(np.uint32))


@pytest.mark.parametrize("value_mode", [0, 1])
def test_pairs_match_oracle(vox, value_mode):
    c = small_uniform(1000, 0.01, 13)
    m, op, q, s, col = preprocessed(c)
    _run_gpu(vox, c, value_mode=value_mode)
    ijk, gid = vox.last_pairs()
    ijk = ijk.cpu().numpy()
    gid = gid.cpu().numpy().astype(np.uint32)
    rg, rijk, _ = oracle.pairs(m, q, s, col, op, c.voxel_size)
    a_ijk, a_gid = sort_pairs(ijk, gid)
    b_ijk, b_gid = sort_pairs(rijk, rg)
    assert np.array_equal(a_ijk, b_ijk) and np.array_equal(a_gid, b_gid)
    # pairs are grouped by leaf (8^3 voxels)
    unit = np.floor_divide(ijk, 8)
    same = np.all(unit[1:] == unit[:-1], axis=1)
    assert np.unique(unit, axis=0).shape[0] == unit.shape[0] - same.sum()


def test_density_values_within_tolerance(vox):
    c = small_uniform(1500, 0.01, 17)
    m, op, q, s, col = preprocessed(c)
    ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size, value_mode=1)
    got, st, _ = _run_gpu(vox, c, value_mode=1)
    assert got.size == ref.size
    meta1 = int(np.frombuffer(ref[640:648].tobytes(), np.uint64)[0]) + 288
    values = meta1 + int(np.frombuffer(ref[meta1:meta1 + 8].tobytes(), np.uint64)[0])
    # topology bytes identical up to the value array
    assert np.array_equal(got[:values], ref[:values])
    gv = got[values:].view(np.float32)
    rv = ref[values:].view(np.float32)
    rel = np.abs(gv[2:].astype(np.float64) - rv[2:]) / np.maximum(np.abs(rv[2:].astype(np.float64)), 1e-30)
    print("density rel err max", rel.max(), "mean", rel.mean())
    assert rel.max() < 1e-5


@pytest.mark.parametrize("n", [0, 1, 2])
def test_tiny_inputs(vox, n):
    c = small_uniform(max(n, 1), 0.01, 31)
    c = c.subset(np.arange(n))
    m, op, q, s, col = preprocessed(c)
    ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    got, st, _ = _run_gpu(vox, c)
    assert np.array_equal(got, ref)


def test_dropped_low_opacity(vox):
    c = small_uniform(600, 0.01, 41)
    c.opacity_logits[::3] = -6.0  # sigmoid < 0.01: radius undefined, Gaussian dropped
    m, op, q, s, col = preprocessed(c)
    ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    got, st, _ = _run_gpu(vox, c)
    assert np.array_equal(got, ref)


def test_spatial_order_same_topology(vox):
    c = small_uniform(1500, 0.01, 5)
    m, op, q, s, col = preprocessed(c)
    g1, _, _ = _run_gpu(vox, c, ijk_batch_size=200_000)
    g2, _, _ = _run_gpu(vox, c, ijk_batch_size=200_000, spatial_order=True)
    a1, v1 = oracle.node2_active_voxels(g1)
    a2, v2 = oracle.node2_active_voxels(g2)
    assert np.array_equal(a1, a2)


@pytest.mark.parametrize("value_mode", [0, 1])
def test_wide_domain_alpha_outside_key(vox, value_mode):
    # two clusters 1000 units apart: 51 voxel-key bits + 11 Gaussian-id bits leave no room for the
    # RGBA8 alpha in the 64-bit pair key, so the pair value is evaluated again in the reduce
    c = small_uniform(1200, 0.01, 41)
    c.means[:600] += np.float32(500.0)
    c.means[600:] -= np.float32(500.0)
    m, op, q, s, col = preprocessed(c)
    got, st, _ = _run_gpu(vox, c, value_mode=value_mode)
    print(f"wide domain mode={value_mode}", st, flush=True)
    assert st["key_bits"] + 8 > 64
    if value_mode == 0:
        ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
        assert st["pair_count"] == rst["pair_count"]
        assert (st["upper_count"], st["lower_count"], st["leaf_count"]) == (rst["upper_count"], rst["lower_count"],
                                                                           rst["leaf_count"])
        assert np.array_equal(got, ref)
    else:
        # density: same topology as the RGBA8 build of the cloud
        rgba, rst, _ = _run_gpu(vox, c, value_mode=0)
        assert st["pair_count"] == rst["pair_count"] and st["voxel_count"] == rst["voxel_count"]
        assert got.size == rgba.size


@pytest.mark.parametrize("seed", [51, 52])
def test_extreme_shapes_pairs(vox, seed):
    # strongly anisotropic, tiny and large Gaussians, opacities near the 0.01 cut: the row
    # candidate interval must never drop a pair the exact predicate keeps
    c = small_uniform(700, 0.01, seed, log_scale_mean=-4.5, log_scale_std=1.4)
    rng = np.random.default_rng(seed)
    c.opacity_logits[: 200] = rng.uniform(-4.59, -4.0, 200).astype(np.float32)
    c.opacity_logits[200: 300] = rng.uniform(4.0, 9.0, 100).astype(np.float32)
    m, op, q, s, col = preprocessed(c)
    ref, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    got, st, _ = _run_gpu(vox, c)
    print("extreme", rst["pair_count"], rst["fragment_count"], st)
    assert st["pair_count"] == rst["pair_count"]
    assert np.array_equal(got, ref)
