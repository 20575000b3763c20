"""Density values of the GPU grid against the FP64 host oracle (vdbo_density_f64 on the same pair
set): maximum and mean relative error, stated and bounded by 1e-5."""
import numpy as np
import pytest

import oracle
from helpers import preprocessed
from nanovdb_editor_amd import synthetic as syn

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["uniform", "clustered"])
def test_density_error_vs_fp64(kind):
    from nanovdb_editor_amd.device import VALUE_DENSITY, Voxelizer
    if kind == "uniform":
        cloud = syn.uniform_cloud(6000, 0.01, -4.0, 0.5, 81)
    else:
        cloud = syn.clustered_cloud(8000, 16, 0.1, 0.005, -5.0, 0.7, 82)
    vox = Voxelizer(0)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()
    sh0 = d(cloud.sh0)
    grid, st = vox.raster_node2(d(cloud.means), d(cloud.quats), d(cloud.log_scales), torch.empty_like(sh0),
                                d(cloud.opacity_logits), cloud.voxel_size, raw=True, sh0=sh0,
                                value_mode=VALUE_DENSITY, voxel_coords=True)
    count = st["voxel_count"] + 2
    values = grid[grid.numel() - 4 * count:].view(torch.float32).cpu().numpy()[2:].astype(np.float64)
    ijk = vox.last_voxels().cpu().numpy()[2:]
    vox.close()
    m, op, q, s, col = preprocessed(cloud)
    gid, pijk, _ = oracle.pairs(m, q, s, col, op, cloud.voxel_size, want_rgba=False)
    vijk, dens = oracle.density_f64(gid, pijk, m, q, s, op, cloud.voxel_size)
    ref = {tuple(v): x for v, x in zip(map(tuple, vijk.tolist()), dens)}
    assert len(ref) == ijk.shape[0]
    exp = np.array([ref[tuple(v)] for v in ijk.tolist()])
    rel = np.abs(values - exp) / np.maximum(np.abs(exp), 1e-300)
    print(f"density vs FP64 ({kind}, {ijk.shape[0]} voxels): max rel err {rel.max():.3e}, "
          f"mean rel err {rel.mean():.3e}", flush=True)
    assert rel.max() < 1e-5
