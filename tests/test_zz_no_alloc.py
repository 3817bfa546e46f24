"""The allocation-free drop-in trtt_hip_dropin_noalloc against the drop-in
symbols it stands for, through the C ABI.

Each case makes two calls in a row, the second on a smaller problem, into
host outputs that already hold values (the drop-ins accumulate).  Bar: 1e-12
relative in FP64, 1e-5 in FP32 (the FP32 drop-ins of a closed-form 3-D plan
may run the axial kernels, the allocation-free entry point always runs the
general ones).  A problem larger than the static array is refused.
"""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = {"f32": 1e-5, "f64": 1e-12}


@pytest.fixture(scope="module")
def lib():
    from yoricktrtt_amd import _lib

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _lib.load()


def _p(a):
    from yoricktrtt_amd import _lib

    return a.ctypes.data_as(_lib.dp)


def _geom(kind, big):
    from yoricktrtt_amd.geometry import (cone_geometry, fan_geometry_2d, parallel_geometry_2d,
                                         parallel_geometry_3d)

    th = np.array([0.3, 2.2])
    if kind in ("cone", "parallel3d"):
        nx, ny, nz, nu, nv = (48, 40, 36, 56, 64) if big else (20, 18, 16, 30, 26)
        if kind == "cone":
            return cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 90.0, 140.0)
        return parallel_geometry_3d(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th)
    nx, ny, nv = (96, 80, 130) if big else (40, 30, 57)
    if kind == "fan":
        return fan_geometry_2d(nx, ny, 1.0, nv, 1.0, th, 150.0, 230.0)
    return parallel_geometry_2d(nx, ny, 1.0, nv, 1.0, th)


KIND = {"parallel3d": 0, "cone": 1, "parallel2d": 2, "fan": 3}
NAME = {"parallel3d": "PB3D", "cone": "CB3D", "parallel2d": "PB2D", "fan": "FB2D"}


def _reference_call(lib, kind, model, out, inp, g, v, tab, job):
    """The existing drop-in symbol for view v."""
    xos, xsd = np.ascontiguousarray(g.xos[v]), np.ascontiguousarray(g.xsd[v])
    name = "trtt_" + NAME[kind] + ("_cubic" if model == "cubic" else "")
    f = getattr(lib, name)
    if g.nz > 1 or kind in ("cone", "parallel3d"):
        if model == "cubic":
            r = f(_p(out), _p(inp), g.nx, g.ny, g.nz, 1.0, 3, g.nu, g.nv, 1.0, 1.0, _p(xos), _p(xsd),
                  job)
        else:
            fp, cum, coord, step = tab
            r = f(_p(out), _p(inp), g.nx, g.ny, g.nz, 1.0, 3, _p(fp), _p(cum), cum.size, step,
                  _p(coord), g.nu, g.nv, 1.0, 1.0, _p(xos), _p(xsd), job)
    elif model == "cubic":
        r = f(_p(out), _p(inp), g.nx, g.ny, 1.0, 3, g.nv, 1.0, _p(xos), _p(xsd), job)
    else:
        fp, cum, coord, step = tab
        r = f(_p(out), _p(inp), g.nx, g.ny, 1.0, 3, _p(fp), _p(cum), cum.size, step, _p(coord),
              g.nv, 1.0, _p(xos), _p(xsd), job)
    assert r == 0


def _noalloc_call(lib, kind, model, out, inp, g, v, tab, job):
    xos, xsd = np.ascontiguousarray(g.xos[v]), np.ascontiguousarray(g.xsd[v])
    if model == "cubic":
        cum, m, step, c0 = None, 0, 0.0, 0.0
    else:
        _, cum_a, coord, step = tab
        cum, m, c0 = _p(cum_a), cum_a.size, float(coord[0])
    return lib.trtt_hip_dropin_noalloc(KIND[kind], _p(out), _p(inp), g.nx, g.ny, g.nz, 1.0, 3, cum,
                                       m, step, c0, g.nu, g.nv, 1.0, 1.0, _p(xos), _p(xsd), job)


@pytest.mark.parametrize("kind,model,prec", [
    ("cone", "cubic", "f32"), ("cone", "cubic", "f64"), ("cone", "table", "f32"),
    ("parallel3d", "cubic", "f64"), ("fan", "table", "f64"), ("fan", "cubic", "f32"),
    ("parallel2d", "cubic", "f64"), ("parallel2d", "table", "f32")])
def test_noalloc_matches_dropin(lib, kind, model, prec):
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.geometry import footprint_table

    print(f"[gpu] noalloc {kind} {model} {prec}", flush=True)
    tab = footprint_table(10000) if model == "table" else None
    lib.trtt_hip_set_precision(_lib.F64 if prec == "f64" else _lib.F32)
    try:
        for big in (True, False):
            g = _geom(kind, big)
            rng = np.random.default_rng(7 if big else 8)
            is3d = kind in ("cone", "parallel3d")
            vshape = (g.nz, g.ny, g.nx) if is3d else (g.ny, g.nx)
            dshape = (g.nv, g.nu) if is3d else (g.nv,)
            vol = rng.uniform(size=vshape)
            det = rng.uniform(size=dshape)
            for job, inp, oshape in ((0, vol, dshape), (1, det, vshape)):
                for v in range(g.nviews):
                    base = rng.uniform(size=oshape)
                    ref = base.copy()
                    got = base.copy()
                    _reference_call(lib, kind, model, ref, inp, g, v, tab, job)
                    r = _noalloc_call(lib, kind, model, got, inp, g, v, tab, job)
                    assert r == 0, _lib.last_error()
                    err = rel_err(got - base, ref - base)
                    assert err < TOL[prec], (big, job, v, err)
    finally:
        lib.trtt_hip_set_precision(_lib.F32)


def test_noalloc_refuses_oversize(lib):
    """An FP64 256^3 volume (128 MiB of input alone) is refused before any
    device work, with a message naming the limit."""
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.geometry import cone_geometry

    g = cone_geometry(256, 256, 256, 1.0, 384, 512, 1.0, 1.0, np.array([0.0]), 1000.0, 1500.0)
    lib.trtt_hip_set_precision(_lib.F64)
    try:
        vol = np.zeros((256, 256, 256))
        out = np.zeros((512, 384))
        xos, xsd = np.ascontiguousarray(g.xos[0]), np.ascontiguousarray(g.xsd[0])
        r = lib.trtt_hip_dropin_noalloc(1, _p(out), _p(vol), 256, 256, 256, 1.0, 3, None, 0, 0.0,
                                        0.0, 384, 512, 1.0, 1.0, _p(xos), _p(xsd), 0)
        assert r == -1
        assert "static array" in _lib.last_error()
    finally:
        lib.trtt_hip_set_precision(_lib.F32)
