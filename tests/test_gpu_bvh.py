"""GPU parity of the VoxelBVH Gaussian keys (VoxelBVH.cpp:1225-1435) against the oracle: world box,
sorted keys, Gaussian ids and run headers bit-exact."""
import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


def _cloud(n, seed, spread=1.0, log_scale=-4.0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    logit = rng.standard_normal(n).astype(np.float32)
    logit[::13] = -6.0                        # opacity below 0.01: clamped cutoff radius
    q = rng.standard_normal((n, 4)).astype(np.float32)
    ls = (log_scale + 0.6 * rng.standard_normal((n, 3))).astype(np.float32)
    sh0 = np.zeros((n, 3), np.float32)
    m, op, qn, s, _ = oracle.preprocess(means, logit, q, ls, sh0)
    return m, op, qn, s


def _xf(seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((3, 3))
    qm, _ = np.linalg.qr(a)
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (1.7 * qm).astype(np.float32)
    m[3, :3] = rng.uniform(-2, 2, 3)
    return m


@pytest.mark.parametrize("n,res,seed,with_xf", [(1, 64, 1, False), (2000, 64, 2, False), (2000, 1024, 3, True),
                                                (20000, 4096, 4, False), (5000, 1, 5, False),
                                                (30000, 512, 6, True)])
def test_voxelbvh_keys_match_oracle(vox, n, res, seed, with_xf):
    m, op, q, s = _cloud(n, seed)
    xf = _xf(seed) if with_xf else None
    ek, ei, er, eb = oracle.voxelbvh_ijkl(m, op, q, s, res, xf)
    print(f"voxelbvh case n={n} res={res} transform={with_xf} runs={int((er[:, 1] > 0).sum())}", flush=True)
    dev = [torch.from_numpy(np.ascontiguousarray(a).ravel()).cuda() for a in (m, op, q, s)]
    k, i, r, b = vox.voxelbvh_ijkl(*dev, res, xf)
    assert np.array_equal(b.cpu().numpy().view(np.uint32), eb.view(np.uint32))
    assert np.array_equal(k.cpu().numpy().view(np.uint64), ek)
    assert np.array_equal(i.cpu().numpy().view(np.uint32), ei)
    assert np.array_equal(r.cpu().numpy().view(np.uint32), er)


def test_voxelbvh_large_cloud_properties(vox):
    """2M Gaussians: sortedness, stability and run headers against np.unique (no oracle run)."""
    n = 2_000_000
    rng = np.random.default_rng(11)
    m = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    op = rng.uniform(0.005, 1.0, n).astype(np.float32)
    q = rng.standard_normal((n, 4)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    s = np.exp(rng.normal(-5.0, 0.7, (n, 3))).astype(np.float32)
    dev = [torch.from_numpy(np.ascontiguousarray(a).ravel()).cuda() for a in (m, op, q, s)]
    k, i, r, _ = vox.voxelbvh_ijkl(*dev, 4096)
    k = k.cpu().numpy().view(np.uint64)
    i = i.cpu().numpy().view(np.uint32)
    r = r.cpu().numpy().view(np.uint32)
    assert np.all(k[1:] >= k[:-1])
    same = k[1:] == k[:-1]
    assert np.all(i[1:][same] >= i[:-1][same])
    assert np.all(np.bincount(i[i != 0xFFFFFFFF], minlength=n) >= 1)
    u, first, counts = np.unique(k, return_index=True, return_counts=True)
    assert np.array_equal(r[:u.size, 0], first)
    assert np.array_equal(r[:u.size, 1], first + counts)
    assert not r[u.size:].any()
