"""Raw data files of the TRTT scripts (yoricktrtt_amd.rawio), CPU only: the
reader follows Yorick's raw_read (first dimension fastest, header offset) and
the three reference call sites, and the reference's simulated ball files
(a committed slice, and the files themselves when the reference is there)."""
import os

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


BOULE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "boule_slice.npz")
SIMUL = "/root/reference/Tomography_3D/Simul_Data"


def _boule_standin(tmp_path, fx):
    """A file of proj_boule.data's length and header holding the fixture's
    views at their positions (zeros elsewhere), and boule.data rebuilt."""
    proj = np.zeros((720, 32, 32), np.float32)
    proj[fx["views"]] = fx["proj"]
    p, v = tmp_path / "proj_boule.data", tmp_path / "boule.data"
    with open(p, "wb") as f:
        f.write(fx["proj_header"].tobytes())
        f.write(proj.tobytes())
    with open(v, "wb") as f:
        f.write(fx["vol_header"].tobytes())
        f.write(fx["vol"].tobytes())
    assert os.path.getsize(p) == int(fx["proj_size"]) and os.path.getsize(v) == int(fx["vol_size"])
    return p, v


def test_boule_slice_readers(tmp_path):
    """The reference's simulated ball (proj_boule.data, boule.data): the
    script's call, every 180th of 720 views, in stored and in reference
    orientation (u reversed, then transposed), and the volume."""
    fx = dict(np.load(BOULE))
    p, v = _boule_standin(tmp_path, fx)
    got = rawio.read_projections_3d(p, 32, 32, 720, each=180, ndata=3)
    assert got.shape == (3, 32, 32) and got.dtype == np.float64
    for k, view in enumerate((0, 180, 360)):
        assert np.array_equal(got[k], fx["proj"][list(fx["views"]).index(view)])
    ro = rawio.read_projections_3d(p, 32, 32, 720, each=180, ndata=3, reference_orientation=True)
    j, i = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    for k in range(3):
        assert np.array_equal(ro[k], got[k][i, 31 - j])
    # the ball: one centred disk, the same in every view; the orientation
    # moves its centroid by mirroring the half-pixel offset of the data
    for a in (got, ro):
        w = a[0]
        cu = (w * np.arange(32)[None, :]).sum() / w.sum()
        cv = (w * np.arange(32)[:, None]).sum() / w.sum()
        assert abs(cu - 15.5) < 0.1 and abs(cv - 15.5) < 0.1
    assert np.abs(fx["proj"] - fx["proj"][0]).max() < 1e-2 * fx["proj"].max()
    vol = rawio.read_volume(v, 16, 16, 16)
    assert np.array_equal(vol, fx["vol"].astype(np.float64))
    assert (vol > 0).sum() == 432 and np.array_equal(vol, vol[::-1, ::-1, ::-1])
    with pytest.raises(ValueError, match="nu == nv"):
        rawio.read_projections_3d(p, 64, 16, 720, reference_orientation=True)


@pytest.mark.skipif(not os.path.isdir(SIMUL), reason="reference data files absent")
def test_boule_reference_files_match_slice():
    fx = dict(np.load(BOULE))
    proj = rawio.read_projections_3d(os.path.join(SIMUL, "proj_boule.data"), 32, 32, 720)
    assert proj.shape == (720, 32, 32)
    assert np.array_equal(proj[fx["views"]], fx["proj"].astype(np.float64))
    vol = rawio.read_volume(os.path.join(SIMUL, "boule.data"), 16, 16, 16)
    assert np.array_equal(vol, fx["vol"].astype(np.float64))
