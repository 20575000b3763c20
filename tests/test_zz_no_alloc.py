"""The allocation-free sort entry point (vdbr_radix_sort_u64_noalloc) against the allocating one
(vdbr_radix_sort_u64), both through the C ABI: identical keys and payloads on two calls in a row
(the second smaller, so stale scratch from the first must not leak in), keys-only and with a
payload, even and odd pass counts; sizes beyond the scratch maximum fail with a message; the
context's workspace stays empty."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 22


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()


@pytest.fixture(scope="module")
def lib_ctx():
    from nanovdb_editor_amd._lib import lib
    L = lib()
    ctx = L.vdbr_context_create(0)
    assert ctx
    stream = torch.cuda.current_stream(0)
    assert L.vdbr_context_set_stream(ctx, ctypes.c_void_p(stream.cuda_stream)) == 0
    yield L, ctx
    L.vdbr_context_destroy(ctx)


def _keys(n, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64)
    k ^= rng.integers(0, 2, n, dtype=np.int64).astype(np.uint64) << np.uint64(63)
    k[: n // 3] &= np.uint64(0xFFFF00000000FFFF)  # many equal digits: stability matters
    return k


def _sort_both(L, ctx, keys, with_vals, bit_begin, bit_end):
    n = keys.size
    ka = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    kb = ka.clone()
    va = torch.arange(n, dtype=torch.int32, device="cuda") if with_vals else None
    vb = va.clone() if with_vals else None
    torch.cuda.synchronize()
    rc = L.vdbr_radix_sort_u64_noalloc(ctx, _ptr(ka), _ptr(va), n, bit_begin, bit_end)
    assert rc == 0, L.vdbr_last_error(ctx)
    assert L.vdbr_synchronize(ctx) == 0
    assert L.vdbr_workspace_bytes(ctx) == 0      # nothing came from the context's allocator
    # the allocating entry point on a second context
    ctx2 = L.vdbr_context_create(0)
    try:
        assert L.vdbr_radix_sort_u64(ctx2, _ptr(kb), _ptr(vb), n, bit_begin, bit_end) == 0, L.vdbr_last_error(ctx2)
    finally:
        L.vdbr_context_destroy(ctx2)
    return ka, va, kb, vb


@pytest.mark.parametrize("with_vals,bits", [(True, (0, 64)), (False, (3, 43)), (True, (16, 37))])
def test_noalloc_matches_allocating_sort(lib_ctx, with_vals, bits):
    L, ctx = lib_ctx
    for n, seed in ((3_000_001, 11), (100_003, 12)):
        keys = _keys(n, seed + bits[1])
        ka, va, kb, vb = _sort_both(L, ctx, keys, with_vals, *bits)
        assert torch.equal(ka, kb)
        if with_vals:
            assert torch.equal(va, vb)
        # and against numpy's stable sort of the same bit field
        width = bits[1] - bits[0]
        mask = np.uint64((1 << width) - 1) if width < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
        order = np.argsort((keys >> np.uint64(bits[0])) & mask, kind="stable")
        assert np.array_equal(ka.cpu().numpy().view(np.uint64), keys[order])
        if with_vals:
            assert np.array_equal(va.cpu().numpy(), order.astype(np.int32))


def test_noalloc_limits(lib_ctx):
    L, ctx = lib_ctx
    # the maximum size works
    keys = _keys(MAX_KEYS, 5)
    ka, _, kb, _ = _sort_both(L, ctx, keys, False, 0, 64)
    assert torch.equal(ka, kb)
    # one key more is rejected before any work is queued
    big = torch.zeros(MAX_KEYS + 1, dtype=torch.int64, device="cuda")
    assert L.vdbr_radix_sort_u64_noalloc(ctx, _ptr(big), None, MAX_KEYS + 1, 0, 64) != 0
    assert b"exceed" in L.vdbr_last_error(ctx)
    assert L.vdbr_radix_sort_u64_noalloc(ctx, _ptr(big), None, 10, 0, 65) != 0
    # trivial sizes are no-ops
    one = torch.tensor([7], dtype=torch.int64, device="cuda")
    assert L.vdbr_radix_sort_u64_noalloc(ctx, _ptr(one), None, 1, 0, 64) == 0
    assert L.vdbr_synchronize(ctx) == 0 and int(one.item()) == 7
