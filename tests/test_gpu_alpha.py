"""Fast RGBA8 alpha decision (hardware exp2/log2 with a propagated error bound) against the
correctly rounded chain of gaussian_frag_color.slang:197-212 on 2^30 hashed inputs per seed:
no input may differ, and the fast path must decide the large majority."""
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 2])
def test_fast_alpha_matches_correctly_rounded(seed):
    from nanovdb_editor_amd.device import Voxelizer
    vox = Voxelizer(0)
    count = 1 << 30
    mism, fast = vox.selftest_alpha8(count, seed)
    vox.close()
    print(f"alpha self-test seed={seed}: {fast} of {count} decided fast, {mism} mismatches", flush=True)
    assert mism == 0
    assert fast > 0.9 * count
