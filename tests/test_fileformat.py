"""Gaussian file ingest (fileformat/FileFormat.cpp:51-300, 754-790) through the library's
pnanovdb_fileformat_t table: PLY and NPZ files written by numpy load back bit-exact."""
import numpy as np
import pytest

from gaussian_files import write_ply
from nanovdb_editor_amd.fileformat import FileFormat


def _cloud(n, seed, rest=0):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, 3)).astype(np.float32)
    op = rng.standard_normal(n).astype(np.float32)
    q = rng.standard_normal((n, 4)).astype(np.float32)
    s = rng.standard_normal((n, 3)).astype(np.float32)
    sh0 = rng.standard_normal((n, 3)).astype(np.float32)
    shr = rng.standard_normal((n, rest)).astype(np.float32) if rest else None
    return m, op, q, s, sh0, shr


@pytest.fixture(scope="module")
def ff():
    return FileFormat()


@pytest.mark.parametrize("big_endian,rest,extra", [(False, 45, 0), (True, 45, 2), (False, 0, 1), (False, 9, 0)])
def test_ply_round_trip(ff, tmp_path, big_endian, rest, extra):
    m, op, q, s, sh0, shr = _cloud(1000, 3, rest)
    path = str(tmp_path / "g.ply")
    write_ply(path, m, op, q, s, sh0, shr, big_endian=big_endian, extra_props=extra)
    assert ff.can_load_file(path)
    got = ff.load_numpy(path)
    assert np.array_equal(got[0], m.ravel())
    assert np.array_equal(got[1], op)
    assert np.array_equal(got[2], q.ravel())
    assert np.array_equal(got[3], s.ravel())
    assert np.array_equal(got[4], sh0.ravel())
    if rest:
        assert np.array_equal(got[5], shr.ravel())
    else:
        assert got[5] is None          # PLY without f_rest: sh_n missing, the load still succeeds


def test_ply_ascii_rejected_and_unknown_extension(ff, tmp_path):
    m, op, q, s, sh0, _ = _cloud(4, 1)
    path = str(tmp_path / "a.ply")
    write_ply(path, m, op, q, s, sh0, fmt="ascii")
    with pytest.raises(RuntimeError):
        ff.load_file(path)
    assert not ff.can_load_file(str(tmp_path / "a.xyz"))
    with pytest.raises(RuntimeError):
        ff.load_file(str(tmp_path / "missing.ply"))


@pytest.mark.parametrize("compressed", [False, True])
def test_npz_named_arrays(ff, tmp_path, compressed):
    m, op, q, s, sh0, shr = _cloud(777, 5, 45)
    path = str(tmp_path / "g.npz")
    save = np.savez_compressed if compressed else np.savez
    save(path, means=m, opacities=op, quaternions=q, scales=s, sh_0=sh0, sh_n=shr.reshape(777, 15, 3))
    got = ff.load_numpy(path)
    for g, e in zip(got, (m, op, q, s, sh0, shr)):
        assert np.array_equal(g, e.ravel())


def test_npz_sh_alias(ff, tmp_path):
    # "sh" [N, K, 3] stands for sh_0 (row 0 of every Gaussian) and sh_n (rows 1..K-1)
    m, op, q, s, _, _ = _cloud(300, 6)
    sh = np.random.default_rng(2).standard_normal((300, 16, 3)).astype(np.float32)
    path = str(tmp_path / "alias.npz")
    np.savez_compressed(path, means=m, opacities=op, quaternions=q, scales=s, sh=sh)
    got = ff.load_numpy(path)
    assert np.array_equal(got[4], sh[:, 0, :].ravel())
    assert np.array_equal(got[5], sh[:, 1:, :].ravel())


def test_npz_missing_array_fails(ff, tmp_path):
    m, op, q, s, sh0, _ = _cloud(10, 7)
    path = str(tmp_path / "partial.npz")
    np.savez(path, means=m, opacities=op, quaternions=q, scales=s, sh_0=sh0)
    with pytest.raises(RuntimeError):
        ff.load_file(path)
    # a request for the present arrays only succeeds
    got = ff.load_numpy(path, ("means", "scales"))
    assert np.array_equal(got[1], s.ravel())


def test_npz_zip64_member(ff, tmp_path):
    """Members written with zip64 extra fields (numpy forces them for large arrays) parse too."""
    import zipfile
    m, op, q, s, sh0, shr = _cloud(50, 8, 45)
    path = str(tmp_path / "z64.npz")
    with zipfile.ZipFile(path, "w", compression=zipfile.ZIP_DEFLATED, allowZip64=True) as z:
        for name, arr in (("means", m), ("opacities", op), ("quaternions", q), ("scales", s), ("sh_0", sh0),
                          ("sh_n", shr)):
            with z.open(name + ".npy", "w", force_zip64=True) as f:
                np.lib.format.write_array(f, arr)
    got = ff.load_numpy(path)
    assert np.array_equal(got[0], m.ravel()) and np.array_equal(got[5], shr.ravel())
