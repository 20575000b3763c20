"""Batched raster3d (one captured HIP graph for the clouds it takes): every cloud of a batch
equals a single call on that cloud (and the oracle)."""
import numpy as np
import pytest

import oracle
from helpers import preprocessed, small_clustered, small_uniform
from nanovdb_editor_amd import synthetic as syn

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _tensors(cloud, value_mode=0):
    d = {k: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32).ravel()).cuda()
         for k, a in (("means", cloud.means), ("opacities", cloud.opacity_logits), ("quats", cloud.quats),
                      ("scales", cloud.log_scales), ("sh0", cloud.sh0))}
    d["colors"] = torch.empty_like(d["sh0"])
    d.update(voxel_size=cloud.voxel_size, raw=True, value_mode=value_mode)
    return d


def _single(vox, d):
    grid, st = vox.raster_node2(d["means"], d["quats"], d["scales"], d["colors"], d["opacities"], d["voxel_size"],
                                raw=True, sh0=d["sh0"], value_mode=d["value_mode"])
    return grid.cpu().numpy(), st


def _mixed_clouds():
    rng = np.random.default_rng(21)
    clouds = []
    for i, n in enumerate((1, 40_000, 700, 0, 5_000, 90_000, 2_000, 12_000, 300, 25_000)):
        vs = float(rng.uniform(0.004, 0.02))
        seed = 100 + i
        if i % 2:
            clouds.append((syn.clustered_cloud(n, 64, 0.1, vs, -5.0, 0.7, seed), i % 3 == 2))
        else:
            clouds.append((syn.uniform_cloud(n, vs, -4.0, 0.5, seed), i % 3 == 2))
    return clouds


@pytest.mark.parametrize("workers", [1, 3, 8])
def test_batch_equals_single_calls(vox, workers):
    clouds = _mixed_clouds()
    expect = []
    for c, density in clouds:
        expect.append(_single(vox, _tensors(c, 1 if density else 0)))
    batch = [_tensors(c, 1 if density else 0) for c, density in clouds]
    print(f"batch of {len(batch)} clouds, workers={workers}", flush=True)
    grids, stats = vox.raster_node2_batch(batch, workers=workers)
    # every cloud with fragments runs inside the captured graph (the empty cloud does not)
    with_pairs = sum(1 for _, rst in expect if rst["fragment_count"] > 0)
    assert vox.last_batch_graph_clouds == with_pairs
    for i, ((ref, rst), got, st) in enumerate(zip(expect, grids, stats)):
        assert st["pair_count"] == rst["pair_count"], i
        assert st["voxel_count"] == rst["voxel_count"], i
        g = got.cpu().numpy()
        assert g.size == ref.size, i
        assert np.array_equal(g, ref), f"cloud {i}: {(g != ref).sum()} differing bytes"


def test_batch_matches_oracle(vox):
    clouds = [small_uniform(900, 0.01, 31), small_clustered(2500, 0.005, 32), small_uniform(400, 0.02, 33)]
    grids, _ = vox.raster_node2_batch([_tensors(c) for c in clouds], workers=2)
    assert vox.last_batch_graph_clouds == len(clouds)
    for c, g in zip(clouds, grids):
        m, op, q, s, col = preprocessed(c)
        ref, _ = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
        assert np.array_equal(g.cpu().numpy(), ref)


def test_batch_reuses_arena_and_reports_failing_cloud(vox):
    clouds = [small_uniform(300, 0.01, 41), small_uniform(500, 0.01, 42), small_uniform(200, 0.01, 43)]
    first, _ = vox.raster_node2_batch([_tensors(c) for c in clouds], workers=2)
    first = [g.cpu().numpy() for g in first]
    again, _ = vox.raster_node2_batch([_tensors(c) for c in clouds], workers=2)
    for a, b in zip(first, again):
        assert np.array_equal(a, b.cpu().numpy())
    assert vox.last_batch_graph_clouds == len(clouds)
    bad = [_tensors(c) for c in clouds]
    bad[1]["voxel_size"] = -1.0
    with pytest.raises(RuntimeError, match="cloud 1"):
        vox.raster_node2_batch(bad, workers=2)
    assert vox.raster_node2_batch([], workers=2) == ([], [])


def test_batch_relaunches_captured_graph(vox):
    """A second call with the same clouds, plans and workspaces launches the graph captured by the
    first call again (status words and output cursor cleared per launch): same grids."""
    clouds = [small_uniform(800, 0.01, 51), small_clustered(3000, 0.005, 52), small_uniform(300, 0.02, 53)]
    items = [_tensors(c) for c in clouds]
    first, _ = vox.raster_node2_batch(items, workers=2)  # preprocesses the raw arrays in place
    for d in items:
        d["raw"] = False
    second, _ = vox.raster_node2_batch(items, workers=2)
    assert vox.last_batch_timing["captured"]
    third, st3 = vox.raster_node2_batch(items, workers=2)
    assert not vox.last_batch_timing["captured"], vox.last_batch_timing
    assert vox.last_batch_graph_clouds == len(clouds)
    for c, a, b, g in zip(clouds, first, second, third):
        ref = a.cpu().numpy()
        assert np.array_equal(b.cpu().numpy(), ref)
        assert np.array_equal(g.cpu().numpy(), ref)
        m, op, q, s, col = preprocessed(c)
        exp, _ = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
        assert np.array_equal(ref, exp)
