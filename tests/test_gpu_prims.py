"""GPU parity of the parallel primitives (radix sort, scans) against numpy."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vox():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    yield v
    v.close()


@pytest.mark.parametrize("n,bits", [(1, 64), (1000, 17), (4096, 32), (4097, 40), (1_000_003, 48), (3_000_000, 64)])
def test_sort_u64_stable(vox, n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64)
    keys ^= rng.integers(0, 2, n, dtype=np.int64).astype(np.uint64) << np.uint64(63)
    vals = np.arange(n, dtype=np.int32)
    kt = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    vt = torch.from_numpy(vals.copy()).cuda()
    vox.radix_sort_u64(kt, vt, 0, bits)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    assert np.array_equal(kt.cpu().numpy().view(np.uint64), keys[order])
    assert np.array_equal(vt.cpu().numpy(), vals[order])


def test_sort_bit_range(vox):
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 2**62, 200_000, dtype=np.int64).astype(np.uint64)
    kt = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    vox.radix_sort_u64(kt, None, 20, 44)
    sub = (keys >> np.uint64(20)) & np.uint64((1 << 24) - 1)
    assert np.array_equal(kt.cpu().numpy().view(np.uint64), keys[np.argsort(sub, kind="stable")])


def test_sort_u32(vox):
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 2**31, 300_001, dtype=np.int64).astype(np.int32)
    vals = np.arange(keys.size, dtype=np.int32)
    kt = torch.from_numpy(keys.copy()).cuda()
    vt = torch.from_numpy(vals.copy()).cuda()
    vox.radix_sort_u32(kt, vt, 31)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(kt.cpu().numpy(), keys[order])
    assert np.array_equal(vt.cpu().numpy(), vals[order])


@pytest.mark.parametrize("n", [1, 255, 2048, 2049, 1_000_000])
def test_scans(vox, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1000, n, dtype=np.int64).astype(np.int32)
    got = vox.inclusive_scan(torch.from_numpy(a).cuda(), "add").cpu().numpy()
    assert np.array_equal(got.view(np.uint32), np.cumsum(a.astype(np.uint32), dtype=np.uint32))
    b = rng.integers(0, 2**40, n, dtype=np.int64)
    got = vox.inclusive_scan(torch.from_numpy(b).cuda(), "add").cpu().numpy()
    assert np.array_equal(got, np.cumsum(b))
    got = vox.inclusive_scan(torch.from_numpy(a).cuda(), "max").cpu().numpy()
    assert np.array_equal(got, np.maximum.accumulate(a))
