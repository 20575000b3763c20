"""Density adjoint over the forward's sorted pair list (vdbr_density_backward_pairs) against the
re-walk adjoint (vdbr_density_backward, itself pinned against the oracle's float64 adjoint and
finite differences): the same gradients to float32 rounding (both evaluate every pair term in
double precision; the pair-list path rounds each term to 2^-48 of max |dL/drho| and sums
integers), bitwise reproducible across forwards whose in-leaf pair order differs, and the
fallbacks: a pair list replaced by a later raster call, a build of several slabs, a sort autotune,
and a non-finite gradient all run the re-walk."""
import numpy as np
import pytest

from helpers import small_uniform
from nanovdb_editor_amd import synthetic as syn

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()


def _forward(vox, cloud, slab_pairs=0):
    """Density forward on copies of the raw parameters; returns (params, grid, stats)."""
    from nanovdb_editor_amd.device import VALUE_DENSITY
    m, s, q, o = _t(cloud.means), _t(cloud.log_scales), _t(cloud.quats), _t(cloud.opacity_logits)
    wq, ws, wo = q.clone(), s.clone(), o.clone()
    sh0 = torch.zeros_like(m)
    grid, st = vox.raster_node2(m, wq, ws, torch.empty_like(m), wo, cloud.voxel_size, raw=True, sh0=sh0,
                                value_mode=VALUE_DENSITY, slab_pairs=slab_pairs)
    return (m, s, q, o), grid, st


def _weights(st, seed):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(st["voxel_count"] + 2).astype(np.float32)
    w[:2] = 0
    return _t(w)


def _backward(vox, params, cloud, grid, w, pair_list_id):
    m, s, q, o = params
    g = vox.density_backward(m, s, q, o, cloud.voxel_size, grid, w, pair_list_id=pair_list_id)
    return [x.cpu().numpy().astype(np.float64) for x in g], vox.last_backward_used_pairs


def _agree(a, b, what):
    for x, y, name in zip(a, b, ("means", "log_scales", "quats", "opacity_logits")):
        scale = max(np.abs(y).max(), 1e-30)
        err = np.abs(x - y).max() / scale
        assert err <= 1e-6, f"{what} {name}: {err:.3e} of max {scale:.3e}"


CASES = {
    "uniform": lambda: small_uniform(400, 0.01, 71),
    "mixed_sizes": lambda: syn.clustered_cloud(300, 3, 0.2, 0.005, -4.0, 1.0, seed=72),
    "clustered_20k": lambda: syn.clustered_cloud(20000, 16, 0.05, 0.002, -5.0, 0.4, seed=73),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_pairs_path_matches_rewalk(vox, case):
    cloud = CASES[case]()
    print("pairs adjoint case", case, flush=True)
    params, grid, st = _forward(vox, cloud)
    assert st["slab_count"] == 1 and st["pair_list_id"] != 0
    w = _weights(st, 5)
    got, used = _backward(vox, params, cloud, grid, w, st["pair_list_id"])
    assert used, "the pair list of the forward just run was not used"
    ref, used_ref = _backward(vox, params, cloud, grid, w, 0)
    assert not used_ref
    _agree(got, ref, case)
    assert all(np.isfinite(g).all() for g in got)


def test_pairs_path_independent_of_pair_order(vox):
    """Two forwards of one cloud emit the pairs of a leaf in different orders; the integer sums
    make the gradients bitwise equal, as are two backward calls on one list."""
    cloud = CASES["clustered_20k"]()
    params, grid, st = _forward(vox, cloud)
    w = _weights(st, 6)
    g1, u1 = _backward(vox, params, cloud, grid, w, st["pair_list_id"])
    g2, u2 = _backward(vox, params, cloud, grid, w, st["pair_list_id"])
    params_b, grid_b, st_b = _forward(vox, cloud)
    assert torch.equal(grid, grid_b) and st_b["pair_list_id"] != st["pair_list_id"]
    g3, u3 = _backward(vox, params_b, cloud, grid_b, w, st_b["pair_list_id"])
    assert u1 and u2 and u3
    for a, b, c in zip(g1, g2, g3):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_replaced_pair_list_falls_back(vox):
    a = CASES["uniform"]()
    b = small_uniform(350, 0.01, 74)
    params_a, grid_a, st_a = _forward(vox, a)
    w = _weights(st_a, 7)
    _forward(vox, b)                               # replaces the context's pair list
    got, used = _backward(vox, params_a, a, grid_a, w, st_a["pair_list_id"])
    assert not used
    ref, _ = _backward(vox, params_a, a, grid_a, w, 0)
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)


def test_autotune_drops_the_pair_list(vox):
    cloud = CASES["uniform"]()
    params, grid, st = _forward(vox, cloud)
    w = _weights(st, 8)
    vox.autotune_sort(1 << 16, 40)
    _, used = _backward(vox, params, cloud, grid, w, st["pair_list_id"])
    assert not used


def test_multi_slab_forward_keeps_no_pair_list(vox):
    cloud = small_uniform(2000, 0.01, 75)
    params, grid, st = _forward(vox, cloud, slab_pairs=40_000)
    assert st["slab_count"] > 1 and st["pair_list_id"] == 0
    w = _weights(st, 9)
    got, used = _backward(vox, params, cloud, grid, w, st["pair_list_id"])
    assert not used and all(np.isfinite(g).all() for g in got)


def test_nonfinite_gradient_takes_the_rewalk(vox):
    cloud = CASES["uniform"]()
    params, grid, st = _forward(vox, cloud)
    w = _weights(st, 10)
    w[7] = float("nan")
    got, used = _backward(vox, params, cloud, grid, w, st["pair_list_id"])
    assert not used
    ref, _ = _backward(vox, params, cloud, grid, w, 0)
    for x, y in zip(got, ref):
        assert np.array_equal(x, y, equal_nan=True)
