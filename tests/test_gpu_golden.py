"""HIP path on the golden fixtures: node2 bytes equal the stored oracle grid (RGBA8 bit for bit;
density grids bit for bit outside the value array, values within 1e-5 relative)."""
import glob
import os

import numpy as np
import pytest

from node2_reader import Node2Grid

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "*.npz")))


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_gpu_reproduces_golden(vox, path):
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    print("golden case", os.path.basename(path), flush=True)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()

    means, op, q, s, sh0 = (dev(g[k]) for k in ("means", "opacity_logits", "quats", "log_scales", "sh0"))
    colors = torch.empty_like(sh0)
    grid, st = vox.raster_node2(means, q, s, colors, op, float(g["voxel_size"]), raw=True, sh0=sh0,
                                value_mode=int(g["value_mode"]), prim_batch_size=int(g["prim_batch_size"]),
                                ijk_batch_size=int(g["ijk_batch_size"]))
    got = grid.cpu().numpy()
    ref = g["grid"]
    assert got.size == ref.size
    assert st["pair_count"] == int(g["pair_count"])
    if int(g["value_mode"]) == 0:
        assert np.array_equal(got, ref)
        return
    off, count, _, _ = Node2Grid(ref).blind(1)
    mask = np.ones(ref.size, bool)
    mask[off:off + 4 * count] = False
    assert np.array_equal(got[mask], ref[mask])
    np.testing.assert_allclose(got[off:off + 4 * count].view(np.float32), ref[off:off + 4 * count].view(np.float32),
                               rtol=1e-5, atol=0)
