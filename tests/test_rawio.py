"""Raw data files of the TRTT scripts (yoricktrtt_amd.rawio), CPU only: the
reader follows Yorick's raw_read (first dimension fastest, header offset) and
the three reference call sites."""
import numpy as np
import pytest

from yoricktrtt_amd import rawio


def test_raw_read_is_first_dimension_fastest(tmp_path):
    # a Yorick array a(i, j, k) = 100 i + 10 j + k, written as Yorick stores it
    d1, d2, d3 = 4, 3, 2
    vals = np.array([[[100 * i + 10 * j + k for i in range(d1)] for j in range(d2)]
                     for k in range(d3)], dtype=np.float32)
    p = tmp_path / "a.raw"
    with open(p, "wb") as f:
        f.write(b"\xff" * 512)
        # d1 fastest in memory: that is C order of the [d3][d2][d1] array
        f.write(vals.tobytes(order="C"))
    a = rawio.raw_read(p, "float", (d1, d2, d3), offset=512)
    assert a.shape == (d3, d2, d1)
    assert a[1, 2, 3] == 100 * 3 + 10 * 2 + 1


def test_reference_readers(tmp_path):
    rng = np.random.default_rng(0)
    nu, nv, nst = 6, 5, 12
    proj = rng.uniform(size=(nst, nv, nu)).astype(np.float32)
    rawio.raw_write(tmp_path / "p.raw", proj, offset=512)
    got = rawio.read_projections_3d(tmp_path / "p.raw", nu, nv, nst, each=3, ndata=3)
    assert got.dtype == np.float64 and got.shape == (3, nv, nu)
    assert np.array_equal(got, proj[::3][:3].astype(np.float64))
    vol = rng.uniform(size=(4, 3, 2)).astype(np.float32)
    rawio.raw_write(tmp_path / "v.raw", vol, offset=512)
    assert np.array_equal(rawio.read_volume(tmp_path / "v.raw", 2, 3, 4), vol)
    img = rng.integers(-2 ** 15, 2 ** 15, size=(8, 8), dtype=np.int16)
    rawio.raw_write(tmp_path / "s.raw", img, offset=100)
    r = rawio.read_projection_2d(tmp_path / "s.raw", n=8)
    assert np.array_equal(r, img.astype(np.float64) + 2.0 ** 15)
    assert r.min() >= 0
    rt = rawio.read_projection_2d(tmp_path / "s.raw", n=8, shift=2.0 ** 16 - 1, transpose=True)
    assert np.array_equal(rt, img.T.astype(np.float64) + 2.0 ** 16 - 1)


def test_short_file_is_refused(tmp_path):
    rawio.raw_write(tmp_path / "t.raw", np.zeros(10, np.float32), offset=512)
    with pytest.raises(ValueError):
        rawio.raw_read(tmp_path / "t.raw", "float", (4, 4), offset=512)
