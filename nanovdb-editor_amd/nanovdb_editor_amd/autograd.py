"""Differentiable density voxelization (torch.autograd) on the HIP path.

``voxelize_density(means, log_scales, quats, opacity_logits, voxel_size)`` returns the value
array of the node2 density grid (float32, one value per value index; slots 0 and 1 are the
background) and the grid itself. The values are differentiable with respect to the four raw
parameter tensors; the backward pass is ``vdbr_density_backward_pairs`` over the pair list the
forward kept on the voxelizer's context (double precision terms summed in exact fixed point,
deterministic), or ``vdbr_density_backward`` when another call has replaced that list. The
voxel set of the grid is a piecewise-constant function of the parameters and carries no gradient.
"""
import torch

from .device import VALUE_DENSITY, Voxelizer

_DEFAULT = {}


def _voxelizer(device):
    idx = device.index if device.index is not None else torch.cuda.current_device()
    v = _DEFAULT.get(idx)
    if v is None:
        v = _DEFAULT[idx] = Voxelizer(idx)
    return v


def density_value_offset(grid, value_count):
    """Byte offset of the value array: the builder places it last in the grid."""
    return grid.numel() - 4 * value_count


class DensityVoxelize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, voxel_size, vox, with_coords=False):
        m = means.detach().contiguous()
        work_q = quats.detach().contiguous().clone()
        work_s = log_scales.detach().contiguous().clone()
        work_o = opacity_logits.detach().contiguous().clone()
        sh0 = torch.zeros_like(m)
        colors = torch.empty_like(m)
        grid, st = vox.raster_node2(m, work_q, work_s, colors, work_o, voxel_size, value_mode=VALUE_DENSITY, raw=True,
                                    sh0=sh0, voxel_coords=with_coords)
        count = st["voxel_count"] + 2
        off = density_value_offset(grid, count)
        values = grid[off:].view(torch.float32).clone()
        ctx.save_for_backward(means, log_scales, quats, opacity_logits, grid)
        ctx.voxel_size = float(voxel_size)
        ctx.vox = vox
        ctx.pair_list_id = st.get("pair_list_id", 0)
        ctx.mark_non_differentiable(grid)
        if with_coords:
            ijk = vox.last_voxels()
            ctx.mark_non_differentiable(ijk)
            return values, grid, ijk
        return values, grid

    @staticmethod
    def backward(ctx, grad_values, grad_grid, grad_ijk=None):
        means, log_scales, quats, opacity_logits, grid = ctx.saved_tensors
        if grad_values is None:
            return None, None, None, None, None, None, None
        gm, gs, gq, go = ctx.vox.density_backward(
            means.detach().contiguous(), log_scales.detach().contiguous(), quats.detach().contiguous(),
            opacity_logits.detach().contiguous(), ctx.voxel_size, grid, grad_values.contiguous().float(),
            pair_list_id=ctx.pair_list_id)
        return gm, gs, gq, go, None, None, None


def voxelize_density(means, log_scales, quats, opacity_logits, voxel_size, vox=None, with_coords=False):
    """(values float32 [V + 2], node2 grid uint8) with values differentiable w.r.t. the inputs;
    with ``with_coords=True`` also the int32 [V + 2, 3] voxel coordinates of the value slots.

    Inputs are flat or shaped float32 CUDA tensors: means [N, 3], log_scales [N, 3], quats [N, 4]
    (w, x, y, z, any norm), opacity_logits [N]."""
    vox = vox or _voxelizer(means.device)
    return DensityVoxelize.apply(means, log_scales, quats, opacity_logits, voxel_size, vox, with_coords)
