"""The C ABI library: it loads, exports every function include/trtt_hip.h
declares, and its host-side helpers (XForm3D, mvmult) agree with the
reference transform code.  No GPU compute is issued here."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from yoricktrtt_amd import _lib
from yoricktrtt_amd.projectors import (TRTT_ERR_BAD_GEOMETRY, TRTT_ERR_UNSUPPORTED_JOB,
                                       TRTT_PARALLEL_BEAM, SingleProjector, TomData, TomObj,
                                       TrttError)

REFERENCE_SYMBOLS = ["trtt_PB2D", "trtt_FB2D", "trtt_PB2D_cubic", "trtt_FB2D_cubic", "trtt_PB3D",
                     "trtt_CB3D", "trtt_PB3D_cubic", "trtt_CB3D_cubic", "lf2d_direct",
                     "lf2d_adjoint"]


def test_library_exports_header():
    lib = _lib.load()
    names = _lib.header_functions()
    for n in REFERENCE_SYMBOLS:
        assert n in names
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


def _xf(lib):
    return lib.trtt_xform3d_new()


def _coefs(lib, h):
    c = np.zeros(12)
    lib.trtt_xform3d_get_coefficients(c.ctypes.data_as(_lib.dp), h)
    return c


def test_host_xform3d_matches_reference(golden_xform):
    lib = _lib.load()
    g = golden_xform
    for n in range(6):
        q, t, sc = g["params"][n][:4], g["params"][n][4:7], g["params"][n][7:]
        h = _xf(lib)
        lib.trtt_xform3d_move(h, *t, h)
        assert lib.trtt_xform3d_rotate(h, *q, h) == 0
        lib.trtt_xform3d_scale(h, *sc, h)
        assert np.abs(_coefs(lib, h) - g[f"a{n}"]).max() < 1e-14 * max(1, np.abs(g[f"a{n}"]).max())
        a, b, c = _xf(lib), _xf(lib), _xf(lib)
        lib.trtt_xform3d_set_coefficients(a, g[f"a{n}"].ctypes.data_as(_lib.dp))
        lib.trtt_xform3d_set_coefficients(b, g[f"b{n}"].ctypes.data_as(_lib.dp))
        lib.trtt_xform3d_combine(c, a, b)
        assert np.abs(_coefs(lib, c) - g[f"comb{n}"]).max() < 1e-13 * max(1, np.abs(g[f"comb{n}"]).max())
        assert lib.trtt_xform3d_invert(a, c) == 0
        assert np.abs(_coefs(lib, a) - g[f"inv{n}"]).max() < 1e-12 * max(1, np.abs(g[f"inv{n}"]).max())
        p = np.array([0.3, -1.2, 2.5])
        r = np.zeros(3)
        lib.trtt_xform3d_apply(r.ctypes.data_as(_lib.dp), c, p.ctypes.data_as(_lib.dp))
        m = g[f"comb{n}"].reshape(3, 4)
        assert np.allclose(r, m[:, :3] @ p + m[:, 3], rtol=1e-14, atol=1e-12)
        for x in (h, a, b, c):
            lib.trtt_xform3d_delete(x)
    # a zero quaternion cannot be normalised
    h = _xf(lib)
    assert lib.trtt_xform3d_rotate(h, 0.0, 0.0, 0.0, 0.0, h) == -1
    lib.trtt_xform3d_delete(h)


def test_mvmult():
    lib = _lib.load()
    rng = np.random.default_rng(1)
    a = rng.normal(size=7)
    i = np.array([1, 3, 3, 2, 4, 1, 4], dtype=np.int64)
    j = np.array([2, 1, 5, 5, 3, 4, 1], dtype=np.int64)
    x = rng.normal(size=5)
    out = np.zeros(4)
    lib.trtt_mvmult(out.ctypes.data_as(_lib.dp), x.ctypes.data_as(_lib.dp), 7,
                    a.ctypes.data_as(_lib.dp), i.ctypes.data_as(_lib.lp), j.ctypes.data_as(_lib.lp))
    ref = np.zeros(4)
    np.add.at(ref, i - 1, a * x[j - 1])
    assert np.allclose(out, ref, rtol=1e-15, atol=0)


def test_operator_layer_errors():
    """Yorick-layer status codes: unsupported job, bad geometry."""
    X = TomObj(8, 8, 8, size_footprint=100)
    Yk = TomData(10, 12)
    p = SingleProjector(X, Yk, TRTT_PARALLEL_BEAM)
    with pytest.raises(TrttError) as e:
        p(np.zeros((8, 8, 8)), job=2)
    assert e.value.code == TRTT_ERR_UNSUPPORTED_JOB
    with pytest.raises(TrttError) as e:
        SingleProjector(X, Yk, 99)
    assert e.value.code == TRTT_ERR_BAD_GEOMETRY
