"""Density adjoint on the GPU (vdbr_density_backward through torch.autograd) against two
references: the oracle's float64 adjoint (itself pinned by finite differences on CPU) and torch
float64 autograd over the same pair list. Tolerance: 1e-5 of each gradient group's largest
magnitude (float32 outputs of double accumulations)."""
import numpy as np
import pytest

import oracle
from helpers import preprocessed, small_uniform
from nanovdb_editor_amd import synthetic as syn
from node2_reader import Node2Grid

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _params(cloud, requires_grad=True):
    def t(a):
        x = torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
        return x.requires_grad_(requires_grad)
    return t(cloud.means), t(cloud.log_scales), t(cloud.quats), t(cloud.opacity_logits)


def _run(vox, cloud, seed=0):
    from nanovdb_editor_amd.autograd import voxelize_density
    m, s, q, o = _params(cloud)
    values, grid = voxelize_density(m, s, q, o, cloud.voxel_size, vox=vox)
    rng = np.random.default_rng(seed)
    w = torch.from_numpy(rng.standard_normal(values.numel()).astype(np.float32)).cuda()
    w[:2] = 0
    (values * w).sum().backward()
    grads = [x.grad.detach().cpu().numpy().astype(np.float64) for x in (m, s, q, o)]
    return values.detach().cpu().numpy(), grid.cpu().numpy(), w.cpu().numpy(), grads


def _close(got, ref, name):
    scale = max(np.abs(ref).max(), 1e-30)
    err = np.abs(got - ref).max() / scale
    assert err <= TOL, f"{name}: max error {err:.3e} of max |grad| {scale:.3e}"


def _oracle_grads(cloud, grid_np, w):
    m, op, q, s, col = preprocessed(cloud)
    vs = float(np.float32(cloud.voxel_size))
    gid, ijk, _ = oracle.pairs(m, q, s, col, op, vs, want_rgba=False)
    g = Node2Grid(grid_np)
    _, vox = g.check_invariants()
    lut = {tuple(v[:3]): v[3] for v in vox}
    vidx = np.array([lut[tuple(x)] for x in ijk], np.int64)
    res = w[vidx].astype(np.float64)
    return oracle.density_grad_f64(gid, ijk, cloud.means, cloud.log_scales, cloud.quats, cloud.opacity_logits, vs, res), (gid, ijk, vidx)


@pytest.mark.parametrize("case", ["uniform", "mixed_sizes"])
def test_backward_matches_oracle_adjoint(vox, case):
    if case == "uniform":
        cloud = small_uniform(400, 0.01, 61)
    else:  # large Gaussians whose rows cross many 64-row chunks, next to small ones
        cloud = syn.clustered_cloud(300, 3, 0.2, 0.005, -4.0, 1.0, seed=62)
    print("backward case", case, flush=True)
    values, grid, w, grads = _run(vox, cloud)
    (gm, gs, gq, go), _ = _oracle_grads(cloud, grid, w)
    for got, ref, name in zip(grads, (gm, gs, gq, go), ("means", "log_scales", "quats", "opacity_logits")):
        _close(got.reshape(ref.shape), ref, name)


def test_backward_matches_torch_autograd(vox):
    cloud = small_uniform(250, 0.01, 63)
    values, grid, w, grads = _run(vox, cloud, seed=1)
    _, (gid, ijk, vidx) = _oracle_grads(cloud, grid, w)
    dev = torch.device("cpu")
    m = torch.tensor(cloud.means, dtype=torch.float64, device=dev, requires_grad=True)
    ls = torch.tensor(cloud.log_scales, dtype=torch.float64, device=dev, requires_grad=True)
    q = torch.tensor(cloud.quats, dtype=torch.float64, device=dev, requires_grad=True)
    ol = torch.tensor(cloud.opacity_logits, dtype=torch.float64, device=dev, requires_grad=True)
    vs = float(np.float32(cloud.voxel_size))
    gi = torch.from_numpy(gid.astype(np.int64))
    qn = q / q.norm(dim=1, keepdim=True)
    W, X, Y, Z = qn[:, 0], qn[:, 1], qn[:, 2], qn[:, 3]
    R = torch.stack([torch.stack([1 - 2 * (Y * Y + Z * Z), 2 * (X * Y - Z * W), 2 * (X * Z + Y * W)], -1),
                     torch.stack([2 * (X * Y + Z * W), 1 - 2 * (X * X + Z * Z), 2 * (Y * Z - X * W)], -1),
                     torch.stack([2 * (X * Z - Y * W), 2 * (Y * Z + X * W), 1 - 2 * (X * X + Y * Y)], -1)], 1)
    s = torch.clamp(torch.exp(ls), min=float(np.float32(0.125) * np.float32(vs)))
    o = torch.sigmoid(ol)
    d = (torch.from_numpy(ijk.astype(np.float64)) + 0.5) * vs - m[gi]
    u = torch.einsum("pba,pb->pa", R[gi], d) / s[gi]
    rho = o[gi] * torch.exp(-0.5 * (u * u).sum(1))
    loss = (torch.from_numpy(w[vidx].astype(np.float64)) * rho).sum()
    loss.backward()
    for got, ref, name in zip(grads, (m.grad, ls.grad, q.grad, ol.grad), ("means", "log_scales", "quats", "logits")):
        _close(got.reshape(ref.shape), ref.numpy(), name)


def test_forward_values_and_determinism(vox):
    cloud = small_uniform(300, 0.01, 64)
    v1, grid1, w, g1 = _run(vox, cloud, seed=2)
    v2, grid2, _, g2 = _run(vox, cloud, seed=2)
    assert np.array_equal(grid1, grid2) and np.array_equal(v1, v2)
    for a, b in zip(g1, g2):
        assert np.array_equal(a, b)
    m, op, q, s, col = preprocessed(cloud)
    vs = float(np.float32(cloud.voxel_size))
    ref, _ = oracle.raster_node2(m, q, s, col, op, vs, value_mode=1)
    gref = Node2Grid(ref)
    off, count, _, _ = gref.blind(1)
    np.testing.assert_allclose(v1, ref[off:off + 4 * count].view(np.float32), rtol=1e-5, atol=0)
