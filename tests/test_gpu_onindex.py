"""GPU node2 -> NanoVDB OnIndex conversion (putil/Convert.h) against the oracle's conversion,
and the raster table entry points end to end (RasterUtils.cpp:187-204 returns OnIndex)."""
import glob
import os

import numpy as np
import pytest

import oracle
from helpers import preprocessed, small_uniform
from nanovdb_editor_amd import synthetic as syn
from node2_reader import OnIndexGrid

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "*.npz")))


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _check_padded(got, ref, size):
    assert size == ref.size
    assert got.size >= ref.size and got.size % 8 == 0
    assert np.array_equal(got[:size], ref)
    assert not np.any(got[size:])


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_conversion_of_golden_grids(vox, path):
    with np.load(path, allow_pickle=False) as z:
        node2 = z["grid"]
    print("onindex golden", os.path.basename(path), flush=True)
    ref = oracle.node2_to_onindex(node2)
    got, size = vox.node2_to_onindex(torch.from_numpy(node2).cuda())
    _check_padded(got.cpu().numpy(), ref, size)


@pytest.mark.parametrize("n,vs,seed", [(2000, 0.01, 41), (3000, 0.004, 42)])
def test_gpu_raster_then_conversion(vox, n, vs, seed):
    cloud = small_uniform(n, vs, seed)  # means in [-1, 1]^3: eight root tiles around the origin
    m, op, q, s, col = preprocessed(cloud)
    ref_node2, _ = oracle.raster_node2(m, q, s, col, op, vs)
    ref = oracle.node2_to_onindex(ref_node2)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()

    colors = dev(col)
    grid, st = vox.raster_node2(dev(m), dev(q), dev(s), colors, dev(op), vs)
    print(f"onindex raster n={n} vs={vs} seed={seed} uppers={st['upper_count']}", flush=True)
    got, size = vox.node2_to_onindex(grid)
    _check_padded(got.cpu().numpy(), ref, size)
    o = OnIndexGrid(ref)
    assert int(ref[o.root + 24:o.root + 28].view("<u4")[0]) == st["upper_count"] >= 2


def test_raster_table_from_arrays_returns_onindex():
    from nanovdb_editor_amd.compute import Compute
    from nanovdb_editor_amd.raster import Raster
    cloud = syn.clustered_cloud(1500, 5, 0.3, 0.005, -5.0, 0.7, seed=43)
    compute = Compute()
    raster = Raster(compute)
    arrays = [compute.create_array(np.ascontiguousarray(a, np.float32).ravel())
              for a in (cloud.means, cloud.opacity_logits, cloud.quats, cloud.log_scales, cloud.sh0)]
    arrays.append(compute.create_array(np.zeros(0, np.float32)))
    out = raster.raster_to_nanovdb_from_arrays(cloud.voxel_size, arrays)
    got = compute.array_bytes(out)
    m, op, q, s, col = preprocessed(cloud)
    ref = oracle.node2_to_onindex(oracle.raster_node2(m, q, s, col, op, cloud.voxel_size)[0])
    assert int(got[636:640].view("<u4")[0]) == 20
    assert got.size >= ref.size and np.array_equal(got[:ref.size], ref)
    # the raw arrays received their preprocessed values in place (RasterUtils.cpp:59-117)
    assert np.array_equal(compute.map_array(arrays[2], np.float32), q.ravel())
    assert np.array_equal(compute.map_array(arrays[1], np.float32), op.ravel())
    for a in arrays + [out]:
        compute.destroy_array(a)


def test_raster_table_to_nanovdb_returns_node2():
    from nanovdb_editor_amd.compute import Compute
    from nanovdb_editor_amd.raster import Raster
    cloud = small_uniform(800, 0.01, 44)
    m, op, q, s, col = preprocessed(cloud)
    compute = Compute()
    raster = Raster(compute)
    arrs = {k: compute.create_array(np.ascontiguousarray(v, np.float32).ravel())
            for k, v in (("m", m), ("q", q), ("s", s), ("c", col), ("o", op), ("sh0", cloud.sh0))}
    shn = compute.create_array(np.zeros(0, np.float32))
    out = raster.raster_to_nanovdb(cloud.voxel_size, arrs["m"], arrs["q"], arrs["s"], arrs["c"], arrs["sh0"], shn,
                                   arrs["o"])
    got = compute.array_bytes(out)
    ref, _ = oracle.raster_node2(m, q, s, col, op, cloud.voxel_size)
    assert got.size >= ref.size and np.array_equal(got[:ref.size], ref) and not np.any(got[ref.size:])
    for a in list(arrs.values()) + [shn, out]:
        compute.destroy_array(a)
