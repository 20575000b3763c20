"""Slab windowing of the pair stages: grids built in many voxel-x slabs (forced by a small pair
budget, including slabs that overflow and split) are byte-identical to the single-slab build and
to the oracle, in both value modes; the density adjoint is unchanged."""
import numpy as np
import pytest

import oracle
from helpers import preprocessed, small_clustered, small_uniform

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _run(vox, cloud, mode, slab_pairs):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()
    m, op, q, s, sh0 = d(cloud.means), d(cloud.opacity_logits), d(cloud.quats), d(cloud.log_scales), d(cloud.sh0)
    grid, st = vox.raster_node2(m, q, s, torch.empty_like(sh0), op, cloud.voxel_size, raw=True, sh0=sh0,
                                value_mode=mode, slab_pairs=slab_pairs)
    return grid.cpu().numpy(), st


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("kind,slab_pairs", [("uniform", 40_000), ("uniform", 3_000), ("clustered", 25_000),
                                             ("clustered", 500)])
def test_slabbed_grid_equals_single_slab(vox, mode, kind, slab_pairs):
    cloud = small_uniform(2000, 0.01, 71) if kind == "uniform" else small_clustered(4000, 0.005, 72)
    ref, rst = _run(vox, cloud, mode, 0)
    got, st = _run(vox, cloud, mode, slab_pairs)
    print(f"slab case {kind} mode={mode} budget={slab_pairs} pairs={st['pair_count']} slabs={st['slab_count']}",
          flush=True)
    assert rst["slab_count"] == 1 and st["slab_count"] > 1
    assert st["pair_count"] == rst["pair_count"] and st["voxel_count"] == rst["voxel_count"]
    assert np.array_equal(got, ref)
    if mode == 0:
        m, op, q, s, col = preprocessed(cloud)
        exp, _ = oracle.raster_node2(m, q, s, col, op, cloud.voxel_size)
        assert np.array_equal(got, exp)
    # the pair list is kept only for single-slab builds
    assert vox.last_pairs()[1].numel() == 0


def test_slabbed_density_backward_matches(vox):
    from nanovdb_editor_amd.device import VALUE_DENSITY
    cloud = small_uniform(1200, 0.01, 73)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()
    raw = [d(a) for a in (cloud.means, cloud.log_scales, cloud.quats, cloud.opacity_logits)]

    def grads(slab_pairs):
        q, s, o = raw[2].clone(), raw[1].clone(), raw[3].clone()
        sh0 = torch.zeros_like(raw[0])
        grid, st = vox.raster_node2(raw[0], q, s, torch.empty_like(sh0), o, cloud.voxel_size, raw=True, sh0=sh0,
                                    value_mode=VALUE_DENSITY, slab_pairs=slab_pairs)
        gv = torch.linspace(-1, 1, st["voxel_count"] + 2, device="cuda")
        return st, vox.density_backward(raw[0], raw[1], raw[2], raw[3], cloud.voxel_size, grid, gv)

    st1, g1 = grads(0)
    st2, g2 = grads(5_000)
    assert st2["slab_count"] > 1
    for a, b in zip(g1, g2):
        assert torch.equal(a, b)


def test_slabs_with_alpha_outside_the_key(vox):
    """Wide domain (clusters at +-500): voxel and Gaussian-id bits leave no room for the alpha in the
    pair key, so each slab's leaf reduce evaluates the pair values again from the voxel boxes; many
    slabs must still give the oracle's grid."""
    c = small_uniform(1200, 0.01, 41)
    c.means[:600] += np.float32(500.0)
    c.means[600:] -= np.float32(500.0)
    got, st = _run(vox, c, 0, 20_000)
    print(f"wide-domain slabs: key bits {st['key_bits']}, slabs {st['slab_count']}", flush=True)
    assert st["key_bits"] + 8 > 64 and st["slab_count"] > 1
    m, op, q, s, col = preprocessed(c)
    exp, rst = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    assert st["pair_count"] == rst["pair_count"]
    assert np.array_equal(got, exp)
