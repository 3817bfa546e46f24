"""The drop-in boundary: every symbol the unchanged Yorick glue links, and the
host-side per-pixel column symbols.

* trtt_PB2D_sparser / trtt_FB2D_sparser equal the reference's own output
  (tests/golden/sparse2d.npz, written from the reference C by
  tests/golden/make_golden.py) bit for bit.
* lf2d_sparser / lf2d_line_integ_sparser equal the oracle's projection of a
  one-hot image (the reference LF source needs yapi.h, absent here, so it is
  not built; the oracle's LF models are pinned by the mathematics below and
  in test_oracle.py).
* The oracle's line-integral LF model (lf2d_line_integ_*): in the parallel
  limit (source at 1e9) the point-sampled trapezoid times its amplitude is
  the exact chord length of the ray through the square pixel, computed here
  by clipping the ray against the square; direct and adjoint satisfy the
  dot-product identity.
* A plugin object that declares every extern of plugin/ywrap.c and
  tomography-LongFessler-model/ywrap.c (and the XForm3D calls of
  plugin/Ytrtt.c) links against libtrtt_hip.so with -Wl,--no-undefined.
  LU_* stay with the plugin's own Ygauss.o (gauss.c, a generic solver off
  the projector path), so the object does not reference them.

These tests call host code only (no GPU).
"""
import ctypes as C
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import CASES2D, REPO, rel_err


@pytest.fixture(scope="module")
def lib():
    from yoricktrtt_amd import _lib

    return _lib.load()


def _p(a):
    from yoricktrtt_amd import _lib

    return a.ctypes.data_as(_lib.dp)


@pytest.mark.parametrize("name", CASES2D)
def test_sparser_2d_equals_reference(lib, golden2d, golden_sparse2d, name):
    gs, g2 = golden_sparse2d, golden2d
    cum, coord, step = g2["cum"], g2["coord"], float(g2["step"])
    fp = np.zeros_like(cum)  # the footprint samples are not read (as in the reference)
    nx, ny, s = int(gs["nx"]), int(gs["ny"]), float(gs["s"])
    cols = gs[f"{name}/cols"]
    nv, vs = int(gs[f"{name}/nv"]), float(gs[f"{name}/vs"])
    f = lib.trtt_FB2D_sparser if name.startswith("fan") else lib.trtt_PB2D_sparser
    for v in range(cols.shape[0]):
        xos = np.ascontiguousarray(gs[f"{name}/xos"][v])
        xsd = np.ascontiguousarray(gs[f"{name}/xsd"][v])
        got = np.zeros((ny, nx, nv))
        for j in range(ny):
            for i in range(nx):
                assert f(_p(got[j, i]), i, j, nx, ny, s, 3, _p(fp), _p(cum), cum.size, step,
                         _p(coord), nv, vs, _p(xos), _p(xsd)) == 0
        assert np.array_equal(got, cols[v]), (name, v)


def test_sparser_accumulates(lib, golden2d, golden_sparse2d):
    gs, g2 = golden_sparse2d, golden2d
    cum, coord, step = g2["cum"], g2["coord"], float(g2["step"])
    nv, vs = int(gs["fan_a/nv"]), float(gs["fan_a/vs"])
    xos, xsd = np.ascontiguousarray(gs["fan_a/xos"][0]), np.ascontiguousarray(gs["fan_a/xsd"][0])
    out = np.full(nv, 2.0)
    assert lib.trtt_FB2D_sparser(_p(out), 5, 7, 17, 15, float(gs["s"]), 3, _p(cum), _p(cum),
                                 cum.size, step, _p(coord), nv, vs, _p(xos), _p(xsd)) == 0
    assert np.array_equal(out, 2.0 + gs["fan_a/cols"][0, 7, 5])


def _lf(n1=13, n2=11, nd=90, ps=0.9e-3, pst=1.0e-3, vs=1.5e-3, R=541e-3, D=408e-3, o1=1e-3,
        o2=-2e-3):
    from oracle.oracle import LF2DParams
    from yoricktrtt_amd import _lib

    args = (ps, pst, vs, R, D, o1, o2, n1, n2, nd)
    return LF2DParams(*args), _lib.LF2DParameters(*args)


@pytest.mark.parametrize("line", [False, True])
def test_lf_sparser_equals_oracle_columns(lib, oracle, line):
    po, pl = _lf()
    n1, n2, nd = po.objectDimension1, po.objectDimension2, po.detectorDimension
    f = lib.lf2d_line_integ_sparser if line else lib.lf2d_sparser
    for beta, off in ((0.0, 0.0), (0.61, 2e-3), (2.3, -1e-3), (4.4, 0.5e-3)):
        worst = 0.0
        for i2 in range(n2):
            for i1 in range(n1):
                img = np.zeros((n2, n1))
                img[i2, i1] = 1.0
                ref = oracle.lf2d(po, beta, off, img, np.zeros(nd), 1, 0, line=line)
                got = np.full(nd, 7.0)
                f(C.byref(pl), beta, off, i1, i2, _p(got), 1)  # clr: zeroed first
                worst = max(worst, rel_err(got, ref))
                acc = np.full(nd, 1.0)
                f(C.byref(pl), beta, off, i1, i2, _p(acc), 0)
                assert np.allclose(acc - 1.0, got, rtol=0, atol=1e-15)
        assert worst < 1e-14, (beta, worst)


def _chord(x, y, vs, n, s):
    """Length of the line {p : p.n = s} inside the square of side vs centred
    at (x, y), by clipping the parametrised line against both slabs."""
    d = np.array([-n[1], n[0]])
    base = s * np.asarray(n)
    lo, hi = -np.inf, np.inf
    for ax, c in ((0, x), (1, y)):
        a, b = c - 0.5 * vs, c + 0.5 * vs
        if abs(d[ax]) < 1e-300:
            if not (a <= base[ax] <= b):
                return 0.0
            continue
        u1, u2 = (a - base[ax]) / d[ax], (b - base[ax]) / d[ax]
        lo, hi = max(lo, min(u1, u2)), min(hi, max(u1, u2))
    return max(0.0, hi - lo)


@pytest.mark.parametrize("beta", [0.0, 0.29, 0.785398, 1.9, 3.6, 5.5])
def test_oracle_line_model_is_the_chord_length(oracle, beta):
    """Parallel limit (source at 1e9, detector through the centre): the
    line-integral weight of bin k is the chord of the ray at abscissa
    (k - offset) * pixelStep through the pixel square."""
    from oracle.oracle import LF2DParams

    vs, pst, nd = 2.0e-3, 0.7e-3, 64
    p = LF2DParams(pst, pst, vs, 1.0e9, 0.0, 0.0, 0.0, 7, 6, nd)
    cb, sb = math.cos(beta), math.sin(beta)
    off_px = 0.5 * (nd - 1)
    for i1, i2 in ((0, 0), (3, 2), (6, 5), (2, 4)):
        img = np.zeros((6, 7))
        img[i2, i1] = 1.0
        got = oracle.lf2d(p, beta, 0.0, img, np.zeros(nd), 1, 0, line=True)
        x = (i1 - 0.5 * (7 - 1)) * vs
        y = (i2 - 0.5 * (6 - 1)) * vs
        expect = np.array([_chord(x, y, vs, (cb, sb), (k - off_px) * pst) for k in range(nd)])
        assert expect.max() > 0.5 * vs
        assert np.abs(got - expect).max() < 1e-6 * vs, (i1, i2)


def test_oracle_line_model_adjoint(oracle):
    po, _ = _lf()
    rng = np.random.default_rng(57)
    for beta in (0.2, 1.7, 3.3):
        x = rng.normal(size=(po.objectDimension2, po.objectDimension1))
        y = rng.normal(size=po.detectorDimension)
        ax = oracle.lf2d(po, beta, 1e-3, x, np.zeros(po.detectorDimension), 1, 0, line=True)
        aty = oracle.lf2d(po, beta, 1e-3, y, np.zeros_like(x), 1, 1, line=True)
        s1, s2 = float(np.dot(ax, y)), float(np.sum(x * aty))
        assert abs(s1 - s2) <= 1e-12 * abs(s1)


# Every C function the unchanged glue references, with its prototype as the
# glue declares it (plugin/ywrap.c:8-403, tomography-LongFessler-model/
# ywrap.c:9-77) and the XForm3D calls of plugin/Ytrtt.c.  LU_* are defined
# by the plugin's own Ygauss.o and are not referenced here.
GLUE_DECLS = """
typedef struct _trtt_xform3d trtt_xform3d_t;
extern trtt_xform3d_t *trtt_xform3d_new(void);
extern void trtt_xform3d_delete(trtt_xform3d_t *);
extern int trtt_xform3d_reset(trtt_xform3d_t *);
extern int trtt_xform3d_copy(trtt_xform3d_t *, const trtt_xform3d_t *);
extern int trtt_xform3d_combine(trtt_xform3d_t *, const trtt_xform3d_t *, const trtt_xform3d_t *);
extern int trtt_xform3d_move(trtt_xform3d_t *, double, double, double, const trtt_xform3d_t *);
extern int trtt_xform3d_rotate(trtt_xform3d_t *, double, double, double, double,
                               const trtt_xform3d_t *);
extern int trtt_xform3d_scale(trtt_xform3d_t *, double, double, double, const trtt_xform3d_t *);
extern int trtt_xform3d_invert(trtt_xform3d_t *, const trtt_xform3d_t *);
extern int trtt_xform3d_get_coefficients(double *, const trtt_xform3d_t *);
extern int trtt_xform3d_set_coefficients(trtt_xform3d_t *, const double *);
extern int trtt_mvmult(double *, double *, long, double *, long *, long *);
extern int trtt_PB2D_cubic(double *, double *, long, long, double, int, long, double, double *,
                           double *, int);
extern int trtt_FB2D_cubic(double *, double *, long, long, double, int, long, double, double *,
                           double *, int);
extern int trtt_PB2D(double *, double *, long, long, double, int, double *, double *, long,
                     double, double *, long, double, double *, double *, int);
extern int trtt_FB2D(double *, double *, long, long, double, int, double *, double *, long,
                     double, double *, long, double, double *, double *, int);
extern int trtt_PB2D_sparser(double *, long, long, long, long, double, int, double *, double *,
                             long, double, double *, long, double, double *, double *);
extern int trtt_FB2D_sparser(double *, long, long, long, long, double, int, double *, double *,
                             long, double, double *, long, double, double *, double *);
extern int trtt_PB3D(double *, double *, long, long, long, double, int, double *, double *, long,
                     double, double *, long, long, double, double, double *, double *, int);
extern int trtt_CB3D(double *, double *, long, long, long, double, int, double *, double *, long,
                     double, double *, long, long, double, double, double *, double *, int);
extern int trtt_DD_PB3D(double *, double *, long, long, long, double, long, long, double, double,
                        double *, double *, int);
extern int trtt_PB3D_cubic(double *, double *, long, long, long, double, int, long, long, double,
                           double, double *, double *, int);
extern int trtt_CB3D_cubic(double *, double *, long, long, long, double, int, long, long, double,
                           double, double *, double *, int);
extern void lf2d_direct(void *, double, double, double *, double *, int);
extern void lf2d_line_integ_direct(void *, double, double, double *, double *, int);
extern void lf2d_adjoint(void *, double, double, double *, double *, int);
extern void lf2d_line_integ_adjoint(void *, double, double, double *, double *, int);
extern void lf2d_sparser(void *, double, double, long, long, double *, int);
extern void lf2d_line_integ_sparser(void *, double, double, long, long, double *, int);
"""

GLUE_NAMES = [
    "trtt_xform3d_new", "trtt_xform3d_delete", "trtt_xform3d_reset", "trtt_xform3d_copy",
    "trtt_xform3d_combine", "trtt_xform3d_move", "trtt_xform3d_rotate", "trtt_xform3d_scale",
    "trtt_xform3d_invert", "trtt_xform3d_get_coefficients", "trtt_xform3d_set_coefficients",
    "trtt_mvmult", "trtt_PB2D_cubic", "trtt_FB2D_cubic", "trtt_PB2D", "trtt_FB2D",
    "trtt_PB2D_sparser", "trtt_FB2D_sparser", "trtt_PB3D", "trtt_CB3D", "trtt_DD_PB3D",
    "trtt_PB3D_cubic", "trtt_CB3D_cubic", "lf2d_direct", "lf2d_line_integ_direct",
    "lf2d_adjoint", "lf2d_line_integ_adjoint", "lf2d_sparser", "lf2d_line_integ_sparser"]


def test_glue_links_without_undefined_symbols(tmp_path):
    """A plugin shared object referencing every glue-called symbol links
    against libtrtt_hip.so with no undefined symbol left."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc absent")
    libdir = os.path.join(REPO, "yoricktrtt_amd")
    assert os.path.exists(os.path.join(libdir, "libtrtt_hip.so"))
    table = ",\n  ".join(f"(void *)&{n}" for n in GLUE_NAMES)
    src = tmp_path / "glue.c"
    src.write_text(GLUE_DECLS + "\nvoid *glue_symbols[] = {\n  " + table + "\n};\n")
    so = tmp_path / "glue.so"
    r = subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(so), str(src), f"-L{libdir}",
                        "-ltrtt_hip", "-Wl,--no-undefined"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # the check has teeth: one symbol the library does not export fails it
    bad = tmp_path / "bad.c"
    bad.write_text("extern int trtt_not_exported(void);\n"
                   "void *glue_symbols[] = {(void *)&trtt_not_exported};\n")
    r = subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(tmp_path / "bad.so"), str(bad),
                        f"-L{libdir}", "-ltrtt_hip", "-Wl,--no-undefined"], capture_output=True,
                       text=True)
    assert r.returncode != 0 and "trtt_not_exported" in r.stderr
    # and the dynamic symbol table of the library holds each of them
    from yoricktrtt_amd import _lib

    lib = C.CDLL(os.path.join(libdir, "libtrtt_hip.so"))
    for n in GLUE_NAMES:
        assert hasattr(lib, n), n
    assert _lib.load() is not None
