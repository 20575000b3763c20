"""raster_file (RasterUtils.cpp:295-395): a Gaussian PLY / NPZ file rasterizes to the same OnIndex
grid as the oracle on the arrays written to the file."""
import numpy as np
import pytest

import oracle
from gaussian_files import write_ply
from helpers import preprocessed
from nanovdb_editor_amd import synthetic as syn

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _expected(cloud):
    m, op, q, s, col = preprocessed(cloud)
    return oracle.node2_to_onindex(oracle.raster_node2(m, q, s, col, op, cloud.voxel_size)[0])


@pytest.mark.parametrize("kind", ["ply", "ply_be", "npz", "npz_sh_alias"])
def test_raster_file_matches_oracle(tmp_path, kind):
    from nanovdb_editor_amd.compute import Compute
    from nanovdb_editor_amd.raster import Raster
    cloud = syn.clustered_cloud(2500, 6, 0.3, 0.005, -5.0, 0.7, seed=51)
    n = cloud.count
    rest = np.random.default_rng(1).standard_normal((n, 45)).astype(np.float32)
    if kind.startswith("ply"):
        path = str(tmp_path / "g.ply")
        write_ply(path, cloud.means, cloud.opacity_logits, cloud.quats, cloud.log_scales, cloud.sh0, rest,
                  big_endian=kind == "ply_be", extra_props=1)
    elif kind == "npz":
        path = str(tmp_path / "g.npz")
        np.savez_compressed(path, means=cloud.means, opacities=cloud.opacity_logits, quaternions=cloud.quats,
                            scales=cloud.log_scales, sh_0=cloud.sh0, sh_n=rest)
    else:
        path = str(tmp_path / "g.npz")
        sh = np.concatenate([cloud.sh0[:, None, :], rest.reshape(n, 15, 3)], 1)
        np.savez(path, means=cloud.means, opacities=cloud.opacity_logits, quaternions=cloud.quats,
                 scales=cloud.log_scales, sh=sh)
    print(f"raster_file case {kind} n={n}", flush=True)
    compute = Compute()
    raster = Raster(compute)
    out = raster.raster_file(path, cloud.voxel_size)
    got = compute.array_bytes(out)
    compute.destroy_array(out)
    ref = _expected(cloud)
    assert int(got[636:640].view("<u4")[0]) == 20       # OnIndex grid type
    assert got.size >= ref.size and np.array_equal(got[:ref.size], ref)


def test_raster_file_rejects_bad_files(tmp_path):
    from nanovdb_editor_amd.raster import Raster
    raster = Raster()
    with pytest.raises(RuntimeError):
        raster.raster_file(str(tmp_path / "missing.ply"), 0.01)
    partial = str(tmp_path / "partial.npz")
    np.savez(partial, means=np.zeros((3, 3), np.float32))
    with pytest.raises(RuntimeError):
        raster.raster_file(partial, 0.01)
