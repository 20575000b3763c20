"""Voxel coordinates of the value slots (VDBR_FLAG_VOXEL_COORDS) and node2 sampling at voxel
coordinates, against the oracle's tree walk of the same grids."""
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


def _raster(vox, cloud, mode):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()
    m, op, q, s, sh0 = d(cloud.means), d(cloud.opacity_logits), d(cloud.quats), d(cloud.log_scales), d(cloud.sh0)
    colors = torch.empty_like(sh0)
    grid, st = vox.raster_node2(m, q, s, colors, op, cloud.voxel_size, raw=True, sh0=sh0, value_mode=mode,
                                voxel_coords=True)
    return grid, st, vox.last_voxels()


@pytest.mark.parametrize("mode", [0, 1])
def test_voxel_coords_follow_value_order(vox, mode):
    cloud = small_clustered(3000, 0.005, 61)
    grid, st, ijk = _raster(vox, cloud, mode)
    g = grid.cpu().numpy()
    exp_ijk, exp_idx = oracle.node2_active_voxels(g)
    got = ijk.cpu().numpy()
    assert got.shape == (st["voxel_count"] + 2, 3)
    assert np.array_equal(got[exp_idx.astype(np.int64)], exp_ijk)
    assert np.all(got[:2] == np.int32(-2139062144))      # background slots: sentinels


def test_node2_sample_matches_tree_walk(vox):
    cloud = small_uniform(1500, 0.01, 62)
    grid, st, ijk = _raster(vox, cloud, 1)
    count = st["voxel_count"] + 2
    values = grid[grid.numel() - 4 * count:].view(torch.float32)
    # at its own voxels the sample returns the value array
    own = vox.node2_sample(grid, values, ijk[2:].contiguous())
    assert torch.equal(own, values[2:])
    # at random coordinates: the value where the oracle finds an active voxel, else 0
    g = grid.cpu().numpy()
    exp_ijk, exp_idx = oracle.node2_active_voxels(g)
    active = {tuple(map(int, v)): int(i) for v, i in zip(exp_ijk, exp_idx)}
    rng = np.random.default_rng(3)
    lo, hi = exp_ijk.min(0) - 20, exp_ijk.max(0) + 20
    q = rng.integers(lo, hi + 1, (20000, 3)).astype(np.int32)
    q[:500] = exp_ijk[rng.integers(0, exp_ijk.shape[0], 500)]
    q[500] = (70000, 0, 0)                                 # outside the node2 root range
    got = vox.node2_sample(grid, values, torch.from_numpy(q).cuda()).cpu().numpy()
    vals = values.cpu().numpy()
    exp = np.array([vals[active[tuple(map(int, p))]] if tuple(map(int, p)) in active else 0.0 for p in q],
                   np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_fit_loss_gradient_through_autograd(vox):
    """L2 fit of one cloud's density to another grid sampled at the fit voxels: the gradient equals
    density_backward with 2 (values - target) on the value slots, over the same pair list (bitwise),
    and the re-walk adjoint to float32 rounding."""
    from nanovdb_editor_amd.autograd import voxelize_density
    fit, tgt = small_uniform(800, 0.01, 63), small_uniform(800, 0.01, 64)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()
    tgrid, tst, _ = _raster(vox, tgt, 1)
    tcount = tst["voxel_count"] + 2
    tvals = tgrid[tgrid.numel() - 4 * tcount:].view(torch.float32)
    params = [d(fit.means).requires_grad_(True), d(fit.log_scales).requires_grad_(True),
              d(fit.quats).requires_grad_(True), d(fit.opacity_logits).requires_grad_(True)]
    values, grid, ijk = voxelize_density(*params, fit.voxel_size, vox=vox, with_coords=True)
    target = vox.node2_sample(tgrid, tvals, ijk[2:].contiguous())
    loss = ((values[2:] - target) ** 2).sum()
    loss.backward()
    gv = torch.zeros_like(values)
    gv[2:] = 2 * (values[2:].detach() - target)
    ref = vox.density_backward(*[p.detach() for p in params], fit.voxel_size, grid, gv,
                               pair_list_id=vox.last_pair_list_id)
    assert vox.last_backward_used_pairs
    for p, r in zip(params, ref):
        assert torch.equal(p.grad, r)
    rewalk = vox.density_backward(*[p.detach() for p in params], fit.voxel_size, grid, gv)
    for p, r in zip(params, rewalk):
        scale = float(r.abs().max())
        assert float((p.grad - r).abs().max()) <= 1e-6 * scale
