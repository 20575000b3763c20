"""Every benchmark configuration at its stated size on the GPU, checked against the oracle where
the oracle finishes in seconds to a minute, and through size-independent properties elsewhere:

  cfg 1  100k uniform Gaussians, voxel 0.01: node2 bytes equal the oracle's (full run).
  cfg 2  2M clustered Gaussians, voxel 0.005: a seeded 6,000-Gaussian subsample byte-exact; the
         full grid's structural invariants (prefix lanes = popcounts, no empty child, value
         indices = 2 + exclusive scan in node order, node counts, bounding box); every voxel of
         three 16^3 sub-boxes equal to the oracle run on all Gaussians whose AABB meets the box;
         two runs byte-identical.
  cfg 3  1M fit cloud, voxel 0.005: GPU gradients of the L2 fit loss on a seeded 512-Gaussian
         subset against float64 central finite differences (pair support fixed); the full 1M
         forward + backward finite and bitwise reproducible.
  cfg 4  the 256-cloud batch: every grid equal to its single call, the 20 smallest equal to the
         oracle.
  cfg 5  16M clustered Gaussians, voxel 0.002: the slab-windowed build at the default pair budget
         and at half of it byte-identical, full-grid invariants, and three sub-boxes against the
         oracle.
"""
import time

import numpy as np
import pytest

import oracle
from helpers import oracle_threads, preprocessed
from nanovdb_editor_amd import synthetic as syn
from fd_density import FD_GROUPS, fd_l2_gradients
from node2_fast import FastNode2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()


def _gpu_grid(vox, cloud, **kw):
    """Raw cloud -> node2 grid tensor on the GPU (inputs preprocessed in place on device copies)."""
    means, op, q, s, sh0 = (_dev(a) for a in (cloud.means, cloud.opacity_logits, cloud.quats, cloud.log_scales,
                                               cloud.sh0))
    grid, st = vox.raster_node2(means, q, s, torch.empty_like(sh0), op, cloud.voxel_size, raw=True, sh0=sh0, **kw)
    return grid, st


def _say(msg):
    print(msg, flush=True)


def _same_counts(st, rst):
    assert st["fragment_count"] == rst["fragment_count"]
    assert st["pair_count"] == rst["pair_count"]
    assert (st["upper_count"], st["lower_count"], st["leaf_count"]) == (rst["upper_count"], rst["lower_count"],
                                                                       rst["leaf_count"])


def _values(grid_np):
    off, count, size = FastNode2(grid_np).blind(1)
    assert size == 4
    return grid_np[off:off + 4 * count].view("<u4")


def _check_boxes(cloud, grid_np, st, seed, nboxes=3, side=16):
    """Voxels and values of leaf-aligned side^3 boxes around random Gaussians against the oracle run
    on every Gaussian whose voxel AABB meets the box (all of a voxel's pairs come from those)."""
    vs = cloud.voxel_size
    m, op, q, s, col = preprocessed(cloud)
    bbox, _ = oracle.prim_aabb(m, q, s, op, vs)
    fast = FastNode2(grid_np)
    gvals = _values(grid_np)
    rng = np.random.default_rng(seed)
    checked = 0
    # random Gaussians (dense regions) and the extreme ones along x (the sparse edge of the cloud,
    # where the box holds inactive voxels as well)
    picks = list(rng.integers(0, cloud.count, nboxes)) + [int(np.argmin(cloud.means[:, 0])),
                                                         int(np.argmax(cloud.means[:, 0]))]
    partial = 0
    for k in picks:
        c = np.floor(cloud.means[k] / np.float32(vs)).astype(np.int64)
        lo = (c >> 3) * 8 - 8
        hi = lo + side - 1
        sel = np.nonzero(np.all(bbox[:, 3:] >= lo, 1) & np.all(bbox[:, :3] <= hi, 1))[0]
        sub = cloud.subset(sel)
        sm, sop, sq, ss, scol = preprocessed(sub)
        ref, _ = oracle.raster_node2(sm, sq, ss, scol, sop, vs, nthreads=oracle_threads())
        rijk, ridx = oracle.node2_active_voxels(ref)
        inside = np.all(rijk >= lo, 1) & np.all(rijk <= hi, 1)
        rvals = _values(ref)
        rv = {tuple(map(int, v)): int(rvals[i]) for v, i in zip(rijk[inside], ridx[inside])}
        got = fast.voxel_ijk_in_box(lo, hi)
        gv = {tuple(map(int, v[:3])): int(gvals[v[3]]) for v in got}
        assert gv == rv, f"box {lo}..{hi}: {len(gv)} GPU voxels vs {len(rv)} oracle voxels"
        checked += len(rv)
        partial += 0 < len(rv) < side ** 3
        _say(f"  box {lo.tolist()}: {sel.size} Gaussians meet it, {len(rv)} voxels equal")
    assert checked > 0 and partial > 0
    return checked


# ----------------------------------------------------------------------------------------- cfg 1
def test_cfg1_uniform_100k_bytes_match_oracle(vox):
    cloud = syn.workload_uniform_100k()
    _say(f"cfg1: {cloud.count} Gaussians, voxel {cloud.voxel_size}, oracle threads {oracle_threads()}")
    t0 = time.perf_counter()
    m, op, q, s, col = preprocessed(cloud)
    ref, rst = oracle.raster_node2(m, q, s, col, op, cloud.voxel_size, nthreads=oracle_threads())
    _say(f"  oracle {time.perf_counter() - t0:.1f} s: {rst['pair_count']} pairs, {ref.size} bytes")
    grid, st = _gpu_grid(vox, cloud)
    _same_counts(st, rst)
    assert st["window_count"] == rst["window_count"]
    got = grid.cpu().numpy()
    assert got.size == ref.size
    diff = np.nonzero(got != ref)[0]
    assert diff.size == 0, f"{diff.size} differing bytes, first at {diff[:8]}"


# ----------------------------------------------------------------------------------------- cfg 2
@pytest.fixture(scope="module")
def cloud_2m():
    return syn.workload_clustered_2m()


def test_cfg2_clustered_2m_subsample_matches_oracle(vox, cloud_2m):
    rng = np.random.default_rng(12345)
    idx = np.sort(rng.choice(cloud_2m.count, size=6000, replace=False))
    sub = cloud_2m.subset(idx)
    m, op, q, s, col = preprocessed(sub)
    ref, rst = oracle.raster_node2(m, q, s, col, op, sub.voxel_size, nthreads=oracle_threads())
    grid, st = _gpu_grid(vox, sub)
    _say(f"cfg2 subsample: {rst['pair_count']} pairs, {rst['voxel_count'] if 'voxel_count' in rst else ''}")
    _same_counts(st, rst)
    assert np.array_equal(grid.cpu().numpy(), ref)


def test_cfg2_clustered_2m_full_size(vox, cloud_2m):
    t0 = time.perf_counter()
    grid, st = _gpu_grid(vox, cloud_2m)
    grid2, st2 = _gpu_grid(vox, cloud_2m)
    assert torch.equal(grid, grid2), "two runs differ"
    del grid2
    _say(f"cfg2: {st['pair_count']} pairs, {st['voxel_count']} voxels, {st['leaf_count']} leaves, "
         f"grid {st['grid_size']} bytes ({time.perf_counter() - t0:.1f} s)")
    assert st["pair_count"] > 10**9
    g = grid.cpu().numpy()
    t0 = time.perf_counter()
    summ = FastNode2(g).check_invariants()
    _say(f"  invariants {time.perf_counter() - t0:.1f} s")
    assert summ["voxels"] == st["voxel_count"] and summ["leaf"] == st["leaf_count"]
    assert np.array_equal(summ["bbox"], FastNode2(g).bbox())
    # every mean's voxel is active when its Gaussian is not dropped (opacity > 0.01): sampling an
    # array of value indices returns the index (>= 2) at active voxels and 0 elsewhere
    keep = 1.0 / (1.0 + np.exp(-cloud_2m.opacity_logits.astype(np.float64))) > 0.0101
    mv = np.floor(cloud_2m.means[keep] / np.float32(cloud_2m.voxel_size)).astype(np.int32)
    index = torch.arange(st["voxel_count"] + 2, dtype=torch.int32, device="cuda")
    got = vox.node2_sample(grid, index, torch.from_numpy(mv).cuda())
    assert int((got < 2).sum()) == 0
    _check_boxes(cloud_2m, g, st, seed=21)


# ----------------------------------------------------------------------------------------- cfg 3
@pytest.fixture(scope="module")
def fit_target(vox):
    """The 1M target cloud (seed 32) voxelised as a density grid on the GPU."""
    from nanovdb_editor_amd.device import VALUE_DENSITY
    tgt = syn.workload_fit_1m(32)
    grid, st = _gpu_grid(vox, tgt, value_mode=VALUE_DENSITY)
    count = st["voxel_count"] + 2
    return grid, grid[grid.numel() - 4 * count:].view(torch.float32)


def test_cfg3_fit_gradients_vs_finite_differences_512(vox, fit_target):
    """L = sum_v (rho(v) - target(v))^2 over the subset's voxels, target = the 1M seed-32 grid sampled
    there. GPU adjoint vs float64 central differences of L (Richardson-extrapolated, each Gaussian's
    pair support held fixed, residuals = the forward's values - target). Tolerance: 1e-5 of each
    parameter group's largest |gradient|; the relative error of the larger gradients is reported."""
    from nanovdb_editor_amd.autograd import voxelize_density
    fit = syn.workload_fit_1m(31)
    idx = np.sort(np.random.default_rng(512).choice(fit.count, 512, replace=False))
    sub = fit.subset(idx)
    vs = float(np.float32(sub.voxel_size))
    tgrid, tvals = fit_target
    params = [_dev(a).requires_grad_(True) for a in (sub.means, sub.log_scales, sub.quats, sub.opacity_logits)]
    values, grid, ijk = voxelize_density(*params, sub.voxel_size, vox=vox, with_coords=True)
    target = vox.node2_sample(tgrid, tvals, ijk[2:].contiguous())
    loss = ((values[2:] - target) ** 2).sum()
    loss.backward()
    g_gpu = [p.grad.detach().cpu().numpy().astype(np.float64).reshape(512, -1) for p in params]
    # float64 central differences over the oracle's pair support, residuals at the GPU's voxels
    residual = (values[2:].detach() - target).cpu().numpy().astype(np.float64)
    fd = fd_l2_gradients(sub, ijk[2:].cpu().numpy(), residual=residual, nthreads=oracle_threads())
    for name, a, b, k in FD_GROUPS:
        ref = fd[:, a:b]
        got = g_gpu[k]
        scale = np.abs(ref).max()
        err = np.abs(got - ref).max() / scale
        big = np.abs(ref) > 1e-2 * scale
        rel = np.abs(got - ref)[big] / np.abs(ref)[big]
        _say(f"cfg3 FD {name}: max error {err:.2e} of max |g| {scale:.3e}; relative error over |g| > 1% of max: "
             f"max {rel.max():.2e}, mean {rel.mean():.2e}")
        assert err <= 1e-5, name


def test_cfg3_fit_1m_full_forward_backward(vox, fit_target):
    from nanovdb_editor_amd.autograd import voxelize_density
    fit = syn.workload_fit_1m(31)
    tgrid, tvals = fit_target
    grads = []
    for _ in range(2):
        params = [_dev(a).requires_grad_(True) for a in (fit.means, fit.log_scales, fit.quats, fit.opacity_logits)]
        values, grid, ijk = voxelize_density(*params, fit.voxel_size, vox=vox, with_coords=True)
        target = vox.node2_sample(tgrid, tvals, ijk[2:].contiguous())
        loss = ((values[2:] - target) ** 2).sum()
        loss.backward()
        loss = loss.detach()
        grads.append([p.grad.detach().clone() for p in params])
        _say(f"cfg3 full: {values.numel() - 2} voxels, loss {loss.item():.6e}")
        del values, grid, ijk, target, params
    for a, b in zip(*grads):
        assert torch.isfinite(a).all()
        assert torch.equal(a, b)
    assert float(grads[0][3].abs().max()) > 0


# ----------------------------------------------------------------------------------------- cfg 4
def test_cfg4_batch_256(vox):
    clouds = syn.workload_batch_256()
    items = []
    for c in clouds:
        d = {k: _dev(a) for k, a in (("means", c.means), ("opacities", c.opacity_logits), ("quats", c.quats),
                                     ("scales", c.log_scales), ("sh0", c.sh0))}
        d["colors"] = torch.empty_like(d["sh0"])
        d.update(voxel_size=c.voxel_size, raw=True)
        items.append(d)
    grids, stats = vox.raster_node2_batch(items, workers=4)
    _say(f"cfg4: {len(clouds)} clouds, {sum(c.count for c in clouds)} Gaussians, "
         f"{sum(s['pair_count'] for s in stats)} pairs, {vox.last_batch_graph_clouds} in the captured graph")
    assert vox.last_batch_graph_clouds == len(clouds)
    for i, c in enumerate(clouds):
        single, st = _gpu_grid(vox, c)
        assert st["pair_count"] == stats[i]["pair_count"]
        assert torch.equal(single, grids[i]), f"cloud {i} differs from its single call"
    smallest = np.argsort([c.count for c in clouds], kind="stable")[:20]
    for i in smallest:
        c = clouds[i]
        m, op, q, s, col = preprocessed(c)
        ref, _ = oracle.raster_node2(m, q, s, col, op, c.voxel_size, nthreads=oracle_threads())
        assert np.array_equal(grids[i].cpu().numpy(), ref), f"cloud {i} ({c.count} Gaussians) differs from the oracle"


# ----------------------------------------------------------------------------------------- cfg 5
def test_cfg5_clustered_16m(vox):
    vox.trim()  # the workspace of the earlier configurations (batch workers and arena included)
    cloud = syn.workload_clustered_16m()
    t0 = time.perf_counter()
    grid, st = _gpu_grid(vox, cloud, max_node_count=0xFFFFFFFF)
    _say(f"cfg5: {st['pair_count']} pairs in {st['slab_count']} slabs, {st['voxel_count']} voxels, "
         f"grid {st['grid_size']} bytes ({time.perf_counter() - t0:.1f} s)")
    grid_half, st_half = _gpu_grid(vox, cloud, max_node_count=0xFFFFFFFF, slab_pairs=1 << 30)
    _say(f"  half budget: {st_half['slab_count']} slabs")
    assert st_half["slab_count"] > st["slab_count"] > 1
    assert torch.equal(grid, grid_half), "grids differ between slab budgets"
    del grid_half
    vox.trim()
    g = grid.cpu().numpy()
    del grid
    t0 = time.perf_counter()
    summ = FastNode2(g).check_invariants()
    _say(f"  invariants {time.perf_counter() - t0:.1f} s")
    assert summ["voxels"] == st["voxel_count"] and summ["leaf"] == st["leaf_count"]
    _check_boxes(cloud, g, st, seed=51)
