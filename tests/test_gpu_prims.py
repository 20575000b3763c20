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


SORT_SHAPE_THREADS = (512, 256, 1024, 256, 512, 1024, 768)


@pytest.mark.parametrize("digit_bits", [6, 7, 8, 9])
def test_sort_every_shape(digit_bits):
    """Every onesweep tile shape at every digit width gives the stable order (9-bit digits need
    a thread per digit: the 256-thread shapes refuse them)."""
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    try:
        rng = np.random.default_rng(digit_bits)
        for variant in range(7):
            if digit_bits == 9 and SORT_SHAPE_THREADS[variant] < 512:
                with pytest.raises(RuntimeError):
                    v.set_sort_config(0, variant, digit_bits)
                continue
            v.set_sort_config(0, variant, digit_bits)
            for n, bits in ((70_001, 45), (1_200_007, 64)):
                keys = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64)
                keys ^= rng.integers(0, 2, n, dtype=np.int64).astype(np.uint64) << np.uint64(63)
                keys &= np.uint64(0xFFFFFFFFFFFFFFFF >> (64 - bits))
                keys[::7] = keys[0]                      # long runs of equal keys exercise stability
                vals = np.arange(n, dtype=np.int32)
                kt = torch.from_numpy(keys.view(np.int64).copy()).cuda()
                vt = torch.from_numpy(vals.copy()).cuda()
                v.radix_sort_u64(kt, vt, 0, bits)
                order = np.argsort(keys, kind="stable")
                assert np.array_equal(kt.cpu().numpy().view(np.uint64), keys[order]), (variant, n)
                assert np.array_equal(vt.cpu().numpy(), vals[order]), (variant, n)
                # keys only (the shapes without payload room are used only here)
                ko = torch.from_numpy(keys.view(np.int64).copy()).cuda()
                v.radix_sort_u64(ko, None, 0, bits)
                assert np.array_equal(ko.cpu().numpy().view(np.uint64), keys[order]), (variant, n, "keys only")
    finally:
        v.close()


def test_autotune_sort():
    from nanovdb_editor_amd.device import Voxelizer
    v = Voxelizer(0)
    try:
        best, rows = v.autotune_sort(1 << 20, 40)
        expect = {(i, w) for i in range(7) for w in (8, 7)} | {(i, 9) for i in range(7) if SORT_SHAPE_THREADS[i] >= 512}
        assert len(rows) == len(expect) == 19
        assert {(r["variant"], r["digit_bits"]) for r in rows} == expect
        assert best["ms"] == min(r["ms"] for r in rows) and best["ms"] > 0
        # the chosen configuration sorts correctly
        rng = np.random.default_rng(3)
        keys = rng.integers(0, 2**40, 1_500_000, dtype=np.int64)
        kt = torch.from_numpy(keys.copy()).cuda()
        v.radix_sort_u64(kt, None, 0, 40)
        assert np.array_equal(kt.cpu().numpy(), np.sort(keys))
        with pytest.raises(RuntimeError):
            v.set_sort_config(0, 9, 8)
    finally:
        v.close()
