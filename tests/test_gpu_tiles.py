"""GPU parity of the Raster2D scan/sort consumer: the dual-key sort against a stable numpy sort
and the tile binning against the oracle (bit-exact)."""
import ctypes

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


def _dual_expected(lo, hi, va, lbits, hbits):
    ml = np.uint64((1 << lbits) - 1)
    mh = np.uint64((1 << hbits) - 1)
    order = np.lexsort(((lo.astype(np.uint64) & ml), (hi.astype(np.uint64) & mh)))  # stable, last key primary
    return lo[order], hi[order], va[order]


@pytest.mark.parametrize("n,lbits,hbits", [(1, 32, 32), (1000, 32, 13), (70_001, 20, 7), (300_000, 32, 32),
                                           (65_536, 0, 9), (5_000, 11, 0)])
def test_dual_key_sort(vox, n, lbits, hbits):
    rng = np.random.default_rng(n + lbits)
    lo = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    hi = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    hi[::3] = hi[0]
    lo[::5] = lo[0]
    va = np.arange(n, dtype=np.uint32)
    tl, th, tv = (torch.from_numpy(a.view(np.int32).copy()).cuda() for a in (lo, hi, va))
    vox.radix_sort_dual_key(tl, th, tv, lbits, hbits)
    el, eh, ev = _dual_expected(lo, hi, va, lbits, hbits)
    assert np.array_equal(tl.cpu().numpy().view(np.uint32), el)
    assert np.array_equal(th.cpu().numpy().view(np.uint32), eh)
    assert np.array_equal(tv.cpu().numpy().view(np.uint32), ev)


def test_dual_key_sort_through_function_table():
    from nanovdb_editor_amd import _lib
    L = _lib.lib()
    pp = ctypes.cast(L.pnanovdb_get_parallel_primitives(), ctypes.POINTER(_lib.PnanovdbParallelPrimitives)).contents
    fn_t = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                            ctypes.c_uint32)
    create_t = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
    destroy_t = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
    ctx = create_t(pp.create_context)(None, None)
    assert ctx
    rng = np.random.default_rng(9)
    n = 50_000
    lo = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    hi = rng.integers(0, 600, n).astype(np.uint32)
    va = rng.integers(0, 2**31, n).astype(np.uint32)
    ts = [torch.from_numpy(a.view(np.int32).copy()).cuda() for a in (lo, hi, va)]
    bufs = [_lib.VdbrBuffer(ctypes.c_void_p(t.data_ptr()), t.numel() * 4) for t in ts]
    torch.cuda.synchronize()
    fn_t(pp.radix_sort_dual_key)(None, None, ctx, ctypes.addressof(bufs[0]), ctypes.addressof(bufs[1]),
                                 ctypes.addressof(bufs[2]), n, n, 32, 10)
    destroy_t(pp.destroy_context)(None, None, ctx)
    el, eh, ev = _dual_expected(lo, hi, va, 32, 10)
    for t, e in zip(ts, (el, eh, ev)):
        assert np.array_equal(t.cpu().numpy().view(np.uint32), e)


def _projected(n, w, h, seed, radius_max=40):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-60, w + 60, n), rng.uniform(-60, h + 60, n)], 1).astype(np.float32)
    radii = rng.integers(-3, radius_max, n).astype(np.int32)
    depths = rng.uniform(0.1, 50.0, n).astype(np.float32)
    if n:
        depths[::7] = depths[0]
        means[::11] = np.round(means[::11] / 16) * 16   # centres on tile edges
    return means, radii, depths


@pytest.mark.parametrize("n,w,h,ts,seed", [(0, 64, 64, 16, 1), (1, 640, 480, 16, 2), (3_000, 640, 480, 16, 3),
                                           (5_000, 1000, 37, 8, 4), (200_000, 1920, 1080, 16, 5),
                                           (40_000, 4096, 4096, 16, 6), (10, 16, 16, 16, 7)])
def test_tile_bin_matches_oracle(vox, n, w, h, ts, seed):
    means, radii, depths = _projected(n, w, h, seed)
    if n >= 4:
        means[1] = (np.nan, 3.0)
        means[2] = (-1e9, 5.0)
        means[3] = (1e9, 1e9)
    lo, hi, va, off = oracle.tile_bin(means, radii, depths, w, h, ts)
    print(f"tile_bin case n={n} image={w}x{h} tile={ts} isects={lo.size}", flush=True)
    g = vox.tile_bin(torch.from_numpy(means).cuda(), torch.from_numpy(radii).cuda(), torch.from_numpy(depths).cuda(),
                     w, h, ts)
    got = [t.cpu().numpy().view(np.uint32) for t in g]
    assert got[0].size == lo.size
    assert np.array_equal(got[0], lo)
    assert np.array_equal(got[1], hi)
    assert np.array_equal(got[2], va)
    assert np.array_equal(got[3], off)
