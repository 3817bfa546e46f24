"""Write the golden fixtures of the footprint projectors.

Runs in the build container only: it loads the reference C plugin compiled
from its own sources by ``oracle/Makefile`` into ``oracle/_ref/libtrtt_ref.so``
and records inputs and outputs of every projector on small seeded cases:

* ``xform.npz``     transform sequences built with the reference trtt_xform3d_*
* ``footprint.npz`` reference ``_trtt_cumsum`` of cubic B-spline samples
* ``proj3d.npz``    trtt_{PB,CB}3D[_cubic], job 0 and 1, four geometries
* ``proj2d.npz``    trtt_{PB,FB}2D[_cubic], job 0 and 1, three geometries
* ``sparse2d.npz``  trtt_{PB,FB}2D_sparser: the coefficient column of every
                    pixel on the views of the 2-D geometries
* ``dd3d.npz``      trtt_DD_PB3D, job 0 and 1, parallel geometries with pixels
                    smaller and larger than the voxels
* ``deg2.npz``      trtt_{PB,CB}3D and trtt_{PB,FB}2D with s_deg = 2 and a
                    degree-2 table (10^3 and 10^4 entries) on the geometries
                    and inputs of proj3d / proj2d, and the ``_cubic`` forms
                    called with s_deg = 2 (keys ``cubic/...``)

The per-view XOS/XSD pairs are composed with the reference transform code in
the same sequence as the Yorick system builders
(Tomography_3D/Ytrtt3D_system.i:436-480, Tomography_2D/Ytrtt2D_system.i:529-589,
Ytrtt_3D_projectors.i:104-110), so the fixtures pin the host geometry too.

Usage:  make -C oracle && python tests/golden/make_golden.py [fixture ...]
(default: every fixture)
"""
from __future__ import annotations

import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libtrtt_ref.so")

dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(dp)


class RefXForm:
    """Thin handle over the reference trtt_xform3d_t object."""

    def __init__(self, lib):
        self.lib = lib
        self.h = lib.trtt_xform3d_new()

    def move(self, tx, ty, tz):
        self.lib.trtt_xform3d_move(self.h, C.c_double(tx), C.c_double(ty), C.c_double(tz), self.h)
        return self

    def scale(self, sx, sy, sz):
        self.lib.trtt_xform3d_scale(self.h, C.c_double(sx), C.c_double(sy), C.c_double(sz), self.h)
        return self

    def rotate(self, qw, qx, qy, qz):
        r = self.lib.trtt_xform3d_rotate(self.h, C.c_double(qw), C.c_double(qx), C.c_double(qy),
                                         C.c_double(qz), self.h)
        assert r == 0
        return self

    def combine(self, left, right):
        out = RefXForm(self.lib)
        self.lib.trtt_xform3d_combine(out.h, left.h, right.h)
        return out

    def invert(self):
        out = RefXForm(self.lib)
        assert self.lib.trtt_xform3d_invert(out.h, self.h) == 0
        return out

    @property
    def a(self):
        c = np.zeros(12)
        self.lib.trtt_xform3d_get_coefficients(_p(c), self.h)
        return c


def load():
    lib = C.CDLL(LIB)
    lib.trtt_xform3d_new.restype = C.c_void_p
    for name in ("move", "scale"):
        getattr(lib, "trtt_xform3d_" + name).argtypes = [C.c_void_p, C.c_double, C.c_double,
                                                         C.c_double, C.c_void_p]
    lib.trtt_xform3d_rotate.argtypes = [C.c_void_p] + [C.c_double] * 4 + [C.c_void_p]
    lib.trtt_xform3d_combine.argtypes = [C.c_void_p] * 3
    lib.trtt_xform3d_invert.argtypes = [C.c_void_p] * 2
    lib.trtt_xform3d_get_coefficients.argtypes = [dp, C.c_void_p]
    L, I = C.c_long, C.c_int
    d = C.c_double
    lib._trtt_cumsum.argtypes = [dp, dp, L, d]
    lib._trtt_cumsum.restype = None
    tab3 = [dp, dp, L, L, L, d, I, dp, dp, L, d, dp, L, L, d, d, dp, dp, I]
    cub3 = [dp, dp, L, L, L, d, I, L, L, d, d, dp, dp, I]
    tab2 = [dp, dp, L, L, d, I, dp, dp, L, d, dp, L, d, dp, dp, I]
    cub2 = [dp, dp, L, L, d, I, L, d, dp, dp, I]
    lib.trtt_DD_PB3D.argtypes = [dp, dp, L, L, L, d, L, L, d, d, dp, dp, I]
    lib.trtt_DD_PB3D.restype = I
    sparser = [dp, L, L, L, L, d, I, dp, dp, L, d, dp, L, d, dp, dp]
    for n in ("trtt_PB2D_sparser", "trtt_FB2D_sparser"):
        getattr(lib, n).argtypes = sparser
        getattr(lib, n).restype = I
    for n, a in (("trtt_PB3D", tab3), ("trtt_CB3D", tab3), ("trtt_PB3D_cubic", cub3),
                 ("trtt_CB3D_cubic", cub3), ("trtt_PB2D", tab2), ("trtt_FB2D", tab2),
                 ("trtt_PB2D_cubic", cub2), ("trtt_FB2D_cubic", cub2)):
        getattr(lib, n).argtypes = a
        getattr(lib, n).restype = I
    return lib


def ref_geometry_3d(lib, nx, ny, nz, s, nu, nv, us, vs, theta, Rsc, Rsd, offs):
    """XOS/XSD per view with the reference transform code."""
    oxi = RefXForm(lib).move(-0.5 * (nx - 1), -0.5 * (ny - 1), -0.5 * (nz - 1)).scale(s, s, s)
    dxi = RefXForm(lib).move(0.0, -0.5 * nv, -0.5 * nu).scale(1.0, vs, us)
    dxi_inv = dxi.invert()
    XOS, XSD = [], []
    for k, th in enumerate(theta):
        t = 0.5 * (math.pi - th)
        xos = RefXForm(lib).rotate(math.cos(t), 0.0, 0.0, math.sin(t)).move(Rsc, 0.0, 0.0)
        xsd = RefXForm(lib).move(-1.0 * Rsd, 0.0, 0.0)
        u_off, v_off, x_off, y_off, z_off = [o[k] for o in offs]
        if v_off:
            xsd.move(0.0, -1.0 * v_off, 0.0)
        if u_off:
            xsd.move(0.0, 0.0, -1.0 * u_off)
        xot = RefXForm(lib)
        if x_off:
            xot.move(x_off, 0.0, 0.0)
        if y_off:
            xot.move(0.0, y_off, 0.0)
        if z_off:
            xot.move(0.0, 0.0, z_off)
        xp = RefXForm(lib).combine(xot, oxi)
        xp = RefXForm(lib).combine(xos, xp)
        sp = RefXForm(lib).combine(dxi_inv, xsd)
        XOS.append(xp.a)
        XSD.append(sp.a)
    return np.array(XOS), np.array(XSD)


def ref_geometry_2d(lib, nx, ny, s, nv, vs, theta, Rsc, Rsd, voff, xoff, yoff):
    oxi = RefXForm(lib).move(-0.5 * (nx - 1), -0.5 * (ny - 1), 0.0).scale(s, s, s)
    dxi = RefXForm(lib).move(0.0, -0.5 * nv, -0.5).scale(1.0, vs, 1.0)
    dxi_inv = dxi.invert()
    XOS, XSD = [], []
    for k, th in enumerate(theta):
        t = 0.5 * (math.pi - th)
        xos = RefXForm(lib).rotate(math.cos(t), 0.0, 0.0, math.sin(t))
        if Rsc is not None:
            xos.move(Rsc, 0.0, 0.0)
        xsd = RefXForm(lib)
        if Rsd is not None:
            xsd.move(-1.0 * Rsd, 0.0, 0.0)
        if voff[k]:
            xsd.move(0.0, voff[k], 0.0)
        xot = RefXForm(lib)
        if xoff[k]:
            xot.move(xoff[k], 0.0, 0.0)
        if yoff[k]:
            xot.move(0.0, yoff[k], 0.0)
        xp = RefXForm(lib).combine(xos, RefXForm(lib).combine(xot, oxi))
        sp = RefXForm(lib).combine(dxi_inv, xsd)
        XOS.append(xp.a)
        XSD.append(sp.a)
    return np.array(XOS), np.array(XSD)


def bspline3(x):
    t = np.abs(x)
    out = np.where(t < 1.0, 2.0 / 3.0 - t * t + 0.5 * t ** 3, 0.0)
    return np.where((t >= 1.0) & (t < 2.0), (2.0 - t) ** 3 / 6.0, out)


def ref_table(lib, m):
    step = 8.0 / (m - 1)
    coord = np.arange(m) * step - 4.0
    fp = bspline3(coord)
    cum = np.zeros(m)
    lib._trtt_cumsum(_p(cum), _p(fp), m, step)
    return fp, cum, coord, step


def bspline2(x):
    t = np.abs(x)
    out = np.where(t < 0.5, 0.75 - t * t, 0.0)
    return np.where((t >= 0.5) & (t < 1.5), 0.5 * (1.5 - t) ** 2, out)


def ref_table2(lib, m):
    """Degree-2 footprint table: samples of the quadratic B-spline on [-3, 3]
    (n = 2 s_deg + 3 nodes, Ytrtt_tomobj.i:140-181) summed by the reference
    _trtt_cumsum."""
    step = 6.0 / (m - 1)
    coord = np.arange(m) * step - 3.0
    fp = bspline2(coord)
    cum = np.zeros(m)
    lib._trtt_cumsum(_p(cum), _p(fp), m, step)
    return fp, cum, coord, step


def deg2_cases(lib, proj3d, proj2d):
    """s_deg = 2 with a degree-2 table (the 4-D benchmark script's setting,
    Tomography_4D/Ytrtt4D_script_computation_time_benchmarking.i:19-42) on
    the geometries and inputs of proj3d / proj2d: trtt_{PB,CB}3D and
    trtt_{PB,FB}2D, job 0 and 1, tables of 10^3 and 10^4 entries."""
    out = {}
    for m in (1000, 10000):
        fp, cum, coord, step = ref_table2(lib, m)
        out[f"m{m}/cum"] = cum
        out[f"m{m}/coord"] = coord
        out[f"m{m}/step"] = np.array(step)
        vol = np.ascontiguousarray(proj3d["vol"])
        nz, ny, nx = vol.shape
        s = float(proj3d["s"])
        for name in ("cone_a", "cone_edge", "par_a"):
            nu, nv = (int(v) for v in proj3d[f"{name}/dims"])
            us, vs = (float(v) for v in proj3d[f"{name}/scl"])
            xos, xsd = proj3d[f"{name}/xos"], proj3d[f"{name}/xsd"]
            sino_in = proj3d[f"{name}/sino_in"]
            nview = xos.shape[0]
            fn = getattr(lib, "trtt_CB3D" if name.startswith("cone") else "trtt_PB3D")
            fwd = np.zeros((nview, nv, nu))
            bwd = np.zeros((nview, nz, ny, nx))
            for v in range(nview):
                for job in (0, 1):
                    o = fwd[v] if job == 0 else bwd[v]
                    i = np.ascontiguousarray(vol if job == 0 else sino_in[v])
                    r = fn(_p(o), _p(i), nx, ny, nz, s, 2, _p(fp), _p(cum), m, step, _p(coord),
                           nu, nv, us, vs, _p(np.ascontiguousarray(xos[v])),
                           _p(np.ascontiguousarray(xsd[v])), job)
                    assert r == 0
            out[f"m{m}/3d/{name}/fwd"] = fwd
            out[f"m{m}/3d/{name}/bwd"] = bwd
        img = np.ascontiguousarray(proj2d["img"])
        ny2, nx2 = img.shape
        for name in ("fan_a", "fan_edge", "par_a"):
            nv = int(proj2d[f"{name}/nv"])
            vs = float(proj2d[f"{name}/vs"])
            xos, xsd = proj2d[f"{name}/xos"], proj2d[f"{name}/xsd"]
            sino_in = proj2d[f"{name}/sino_in"]
            nview = xos.shape[0]
            fn = getattr(lib, "trtt_FB2D" if name.startswith("fan") else "trtt_PB2D")
            fwd = np.zeros((nview, nv))
            bwd = np.zeros((nview, ny2, nx2))
            for v in range(nview):
                for job in (0, 1):
                    o = fwd[v] if job == 0 else bwd[v]
                    i = np.ascontiguousarray(img if job == 0 else sino_in[v])
                    r = fn(_p(o), _p(i), nx2, ny2, float(proj2d["s"]), 2, _p(fp), _p(cum), m,
                           step, _p(coord), nv, vs, _p(np.ascontiguousarray(xos[v])),
                           _p(np.ascontiguousarray(xsd[v])), job)
                    assert r == 0
            out[f"m{m}/2d/{name}/fwd"] = fwd
            out[f"m{m}/2d/{name}/bwd"] = bwd
    # the closed-form cubic symbols called with s_deg = 2: the cubic CDF on
    # the pixel window of a 1.5-voxel half support (a truncated footprint)
    vol = np.ascontiguousarray(proj3d["vol"])
    nz, ny, nx = vol.shape
    s = float(proj3d["s"])
    for name in ("cone_a", "cone_edge", "par_a"):
        nu, nv = (int(v) for v in proj3d[f"{name}/dims"])
        us, vs = (float(v) for v in proj3d[f"{name}/scl"])
        xos, xsd = proj3d[f"{name}/xos"], proj3d[f"{name}/xsd"]
        sino_in = proj3d[f"{name}/sino_in"]
        nview = xos.shape[0]
        fn = getattr(lib, "trtt_CB3D_cubic" if name.startswith("cone") else "trtt_PB3D_cubic")
        fwd = np.zeros((nview, nv, nu))
        bwd = np.zeros((nview, nz, ny, nx))
        for v in range(nview):
            for job in (0, 1):
                o = fwd[v] if job == 0 else bwd[v]
                i = np.ascontiguousarray(vol if job == 0 else sino_in[v])
                r = fn(_p(o), _p(i), nx, ny, nz, s, 2, nu, nv, us, vs,
                       _p(np.ascontiguousarray(xos[v])), _p(np.ascontiguousarray(xsd[v])), job)
                assert r == 0
        out[f"cubic/3d/{name}/fwd"] = fwd
        out[f"cubic/3d/{name}/bwd"] = bwd
    img = np.ascontiguousarray(proj2d["img"])
    ny2, nx2 = img.shape
    for name in ("fan_a", "fan_edge", "par_a"):
        nv = int(proj2d[f"{name}/nv"])
        vs = float(proj2d[f"{name}/vs"])
        xos, xsd = proj2d[f"{name}/xos"], proj2d[f"{name}/xsd"]
        sino_in = proj2d[f"{name}/sino_in"]
        nview = xos.shape[0]
        fn = getattr(lib, "trtt_FB2D_cubic" if name.startswith("fan") else "trtt_PB2D_cubic")
        fwd = np.zeros((nview, nv))
        bwd = np.zeros((nview, ny2, nx2))
        for v in range(nview):
            for job in (0, 1):
                o = fwd[v] if job == 0 else bwd[v]
                i = np.ascontiguousarray(img if job == 0 else sino_in[v])
                r = fn(_p(o), _p(i), nx2, ny2, float(proj2d["s"]), 2, nv, vs,
                       _p(np.ascontiguousarray(xos[v])), _p(np.ascontiguousarray(xsd[v])), job)
                assert r == 0
        out[f"cubic/2d/{name}/fwd"] = fwd
        out[f"cubic/2d/{name}/bwd"] = bwd
    return out


def xform_cases(lib):
    out = {}
    rng = np.random.default_rng(11)
    seq = []
    for n in range(6):
        q = rng.normal(size=4)
        if n == 0:
            q = q / np.linalg.norm(q)
        t = rng.normal(size=3) * 10
        sc = rng.uniform(0.5, 2.0, size=3)
        a = RefXForm(lib).move(*t).rotate(*q).scale(*sc)
        b = RefXForm(lib).rotate(*rng.normal(size=4)).move(*rng.normal(size=3))
        c = RefXForm(lib).combine(a, b)
        seq.append(np.concatenate([q, t, sc]))
        out[f"a{n}"] = a.a
        out[f"b{n}"] = b.a
        out[f"comb{n}"] = c.a
        out[f"inv{n}"] = c.invert().a
    out["params"] = np.array(seq)
    return out


def proj3d_cases(lib):
    nx, ny, nz = 12, 10, 9
    s = 1.0
    rng = np.random.default_rng(3)
    vol = rng.uniform(0.0, 1.0, size=(nz, ny, nx))
    fp, cum, coord, step = ref_table(lib, 1000)
    cases = {
        # centred cone, footprints inside the detector
        "cone_a": dict(cone=True, nu=20, nv=24, us=1.0, vs=1.0, Rsc=40.0, Rsd=60.0,
                       theta=np.arange(8) * (2 * math.pi / 8), offs=None),
        # small detector: footprints straddle every edge
        "cone_edge": dict(cone=True, nu=9, nv=11, us=1.3, vs=1.1, Rsc=30.0, Rsd=50.0,
                          theta=np.array([0.1, 1.3, 2.9]), offs=None),
        # detector / object offsets (u_off, v_off, x_off, y_off, z_off)
        "cone_off": dict(cone=True, nu=22, nv=26, us=0.9, vs=1.2, Rsc=35.0, Rsd=55.0,
                         theta=np.array([0.2, 2.2, 4.1]),
                         offs=(np.array([1.5, -2.0, 0.7]), np.array([-1.0, 2.5, 0.3]),
                               np.array([0.5, -0.3, 1.1]), np.array([-0.8, 0.4, 0.0]),
                               np.array([0.6, 0.0, -0.9]))),
        "par_a": dict(cone=False, nu=14, nv=18, us=1.0, vs=1.0, Rsc=0.0, Rsd=0.0,
                      theta=np.arange(6) * (math.pi / 6), offs=None),
        "par_edge": dict(cone=False, nu=7, nv=9, us=0.8, vs=1.4, Rsc=0.0, Rsd=0.0,
                         theta=np.array([0.3, 1.9]),
                         offs=(np.array([0.4, -1.2]), np.array([0.9, 0.0]),
                               np.zeros(2), np.zeros(2), np.zeros(2))),
    }
    out = {"vol": vol, "s": np.array(s), "cum": cum, "coord": coord, "step": np.array(step)}
    for name, c in cases.items():
        nview = len(c["theta"])
        offs = c["offs"] or tuple(np.zeros(nview) for _ in range(5))
        xos, xsd = ref_geometry_3d(lib, nx, ny, nz, s, c["nu"], c["nv"], c["us"], c["vs"],
                                   c["theta"], c["Rsc"], c["Rsd"], offs)
        sino_in = rng.uniform(0.0, 1.0, size=(nview, c["nv"], c["nu"]))
        out[f"{name}/xos"] = xos
        out[f"{name}/xsd"] = xsd
        out[f"{name}/dims"] = np.array([c["nu"], c["nv"]])
        out[f"{name}/scl"] = np.array([c["us"], c["vs"]])
        out[f"{name}/theta"] = c["theta"]
        out[f"{name}/geom"] = np.array([c["Rsc"], c["Rsd"]])
        out[f"{name}/offs"] = np.array(offs)
        out[f"{name}/sino_in"] = sino_in
        pre = "CB3D" if c["cone"] else "PB3D"
        for mode in ("cubic", "table"):
            fwd = np.zeros((nview, c["nv"], c["nu"]))
            bwd = np.zeros((nview, nz, ny, nx))
            for v in range(nview):
                for job in (0, 1):
                    o = fwd[v] if job == 0 else bwd[v]
                    i = vol if job == 0 else sino_in[v]
                    i = np.ascontiguousarray(i)
                    if mode == "cubic":
                        r = getattr(lib, f"trtt_{pre}_cubic")(
                            _p(o), _p(i), nx, ny, nz, s, 3, c["nu"], c["nv"], c["us"], c["vs"],
                            _p(np.ascontiguousarray(xos[v])), _p(np.ascontiguousarray(xsd[v])), job)
                    else:
                        r = getattr(lib, f"trtt_{pre}")(
                            _p(o), _p(i), nx, ny, nz, s, 3, _p(fp), _p(cum), len(cum), step,
                            _p(coord), c["nu"], c["nv"], c["us"], c["vs"],
                            _p(np.ascontiguousarray(xos[v])), _p(np.ascontiguousarray(xsd[v])), job)
                    assert r == 0, (name, mode, v, job)
            out[f"{name}/{mode}/fwd"] = fwd
            out[f"{name}/{mode}/bwd"] = bwd
    return out


def proj2d_cases(lib):
    nx, ny = 17, 15
    s = 1.0
    rng = np.random.default_rng(2)
    img = rng.uniform(0.0, 1.0, size=(ny, nx))
    fp, cum, coord, step = ref_table(lib, 1000)
    cases = {
        "fan_a": dict(fan=True, nv=31, vs=1.0, Rsc=50.0, Rsd=80.0,
                      theta=np.arange(12) * (2 * math.pi / 12), voff=None),
        "fan_edge": dict(fan=True, nv=13, vs=1.2, Rsc=30.0, Rsd=45.0,
                         theta=np.array([0.4, 2.0, 3.7]), voff=np.array([2.0, -1.5, 0.5]),
                         xoff=np.array([0.5, 0.0, -0.7]), yoff=np.array([0.0, 0.9, 0.3])),
        "par_a": dict(fan=False, nv=27, vs=1.0, Rsc=None, Rsd=None,
                      theta=np.arange(10) * (math.pi / 10), voff=None),
        "par_edge": dict(fan=False, nv=11, vs=0.9, Rsc=None, Rsd=None,
                         theta=np.array([0.25, 1.5]), voff=np.array([1.0, -2.0])),
    }
    out = {"img": img, "s": np.array(s), "cum": cum, "coord": coord, "step": np.array(step)}
    for name, c in cases.items():
        nview = len(c["theta"])
        z = np.zeros(nview)
        voff = c.get("voff") if c.get("voff") is not None else z
        xoff = c.get("xoff", z)
        yoff = c.get("yoff", z)
        xos, xsd = ref_geometry_2d(lib, nx, ny, s, c["nv"], c["vs"], c["theta"], c["Rsc"],
                                   c["Rsd"], voff, xoff, yoff)
        sino_in = rng.uniform(0.0, 1.0, size=(nview, c["nv"]))
        out[f"{name}/xos"] = xos
        out[f"{name}/xsd"] = xsd
        out[f"{name}/nv"] = np.array(c["nv"])
        out[f"{name}/vs"] = np.array(c["vs"])
        out[f"{name}/theta"] = c["theta"]
        out[f"{name}/geom"] = np.array([c["Rsc"] or 0.0, c["Rsd"] or 0.0])
        out[f"{name}/offs"] = np.array([voff, xoff, yoff])
        out[f"{name}/sino_in"] = sino_in
        pre = "FB2D" if c["fan"] else "PB2D"
        for mode in ("cubic", "table"):
            fwd = np.zeros((nview, c["nv"]))
            bwd = np.zeros((nview, ny, nx))
            for v in range(nview):
                for job in (0, 1):
                    o = fwd[v] if job == 0 else bwd[v]
                    i = np.ascontiguousarray(img if job == 0 else sino_in[v])
                    xo = np.ascontiguousarray(xos[v])
                    xs = np.ascontiguousarray(xsd[v])
                    if mode == "cubic":
                        r = getattr(lib, f"trtt_{pre}_cubic")(_p(o), _p(i), nx, ny, s, 3, c["nv"],
                                                            c["vs"], _p(xo), _p(xs), job)
                    else:
                        r = getattr(lib, f"trtt_{pre}")(_p(o), _p(i), nx, ny, s, 3, _p(fp), _p(cum),
                                                      len(cum), step, _p(coord), c["nv"], c["vs"],
                                                      _p(xo), _p(xs), job)
                    assert r == 0
            out[f"{name}/{mode}/fwd"] = fwd
            out[f"{name}/{mode}/bwd"] = bwd
    return out


def sparse2d_cases(lib, proj2d):
    """Reference sparser columns on the geometries of proj2d.npz (at most
    four views each): col[v, j, i] = trtt_{PB,FB}2D_sparser(i, j) on view v."""
    g = proj2d
    ny, nx = g["img"].shape
    s = float(g["s"])
    fp, cum, coord, step = ref_table(lib, 1000)
    out = {"nx": np.array(nx), "ny": np.array(ny), "s": np.array(s)}
    for name in ("fan_a", "fan_edge", "par_a", "par_edge"):
        nv, vs = int(g[f"{name}/nv"]), float(g[f"{name}/vs"])
        xos, xsd = g[f"{name}/xos"][:4], g[f"{name}/xsd"][:4]
        fn = getattr(lib, "trtt_FB2D_sparser" if name.startswith("fan") else "trtt_PB2D_sparser")
        cols = np.zeros((xos.shape[0], ny, nx, nv))
        for v in range(xos.shape[0]):
            xo, xs = np.ascontiguousarray(xos[v]), np.ascontiguousarray(xsd[v])
            for j in range(ny):
                for i in range(nx):
                    o = cols[v, j, i]
                    assert fn(_p(o), i, j, nx, ny, s, 3, _p(fp), _p(cum), len(cum), step,
                              _p(coord), nv, vs, _p(xo), _p(xs)) == 0
        out[f"{name}/xos"] = xos
        out[f"{name}/xsd"] = xsd
        out[f"{name}/nv"] = np.array(nv)
        out[f"{name}/vs"] = np.array(vs)
        out[f"{name}/cols"] = cols
    return out


def dd3d_cases(lib):
    """trtt_DD_PB3D on parallel geometries: pixels smaller than the voxels
    (footprints over several pixels, which exercises the transpose's v
    loop), larger ones, and a small detector that clamps the rectangles."""
    nx, ny, nz = 11, 9, 8
    s = 1.0
    rng = np.random.default_rng(8)
    vol = rng.uniform(0.0, 1.0, size=(nz, ny, nx))
    cases = {
        "fine": dict(nu=30, nv=34, us=0.45, vs=0.4, theta=np.array([0.0, 0.35, 1.2, 2.6])),
        "coarse": dict(nu=9, nv=12, us=1.6, vs=1.3, theta=np.array([0.5, 1.9])),
        "edge": dict(nu=10, nv=11, us=0.7, vs=0.6, theta=np.array([0.2, 2.3, 4.0])),
    }
    out = {"vol": vol, "s": np.array(s)}
    for name, c in cases.items():
        nview = len(c["theta"])
        offs = tuple(np.zeros(nview) for _ in range(5))
        xos, xsd = ref_geometry_3d(lib, nx, ny, nz, s, c["nu"], c["nv"], c["us"], c["vs"],
                                   c["theta"], 0.0, 0.0, offs)
        sino_in = rng.uniform(0.0, 1.0, size=(nview, c["nv"], c["nu"]))
        fwd = np.zeros((nview, c["nv"], c["nu"]))
        bwd = np.zeros((nview, nz, ny, nx))
        for v in range(nview):
            xo, xs = np.ascontiguousarray(xos[v]), np.ascontiguousarray(xsd[v])
            assert lib.trtt_DD_PB3D(_p(fwd[v]), _p(vol), nx, ny, nz, s, c["nu"], c["nv"], c["us"],
                                    c["vs"], _p(xo), _p(xs), 0) == 0
            si = np.ascontiguousarray(sino_in[v])
            assert lib.trtt_DD_PB3D(_p(bwd[v]), _p(si), nx, ny, nz, s, c["nu"], c["nv"], c["us"],
                                    c["vs"], _p(xo), _p(xs), 1) == 0
        out[f"{name}/xos"] = xos
        out[f"{name}/xsd"] = xsd
        out[f"{name}/dims"] = np.array([c["nu"], c["nv"]])
        out[f"{name}/scl"] = np.array([c["us"], c["vs"]])
        out[f"{name}/sino_in"] = sino_in
        out[f"{name}/fwd"] = fwd
        out[f"{name}/bwd"] = bwd
    return out


def main(names=None):
    if not os.path.exists(LIB):
        sys.exit(f"{LIB} missing: run `make -C oracle` where the reference sources exist")
    lib = load()
    names = names or ["xform", "footprint", "proj3d", "proj2d", "sparse2d", "dd3d", "deg2"]
    if "xform" in names:
        np.savez_compressed(os.path.join(HERE, "xform.npz"), **xform_cases(lib))
    if "footprint" in names:
        fp1, cum1, coord1, step1 = ref_table(lib, 1000)
        fp2, cum2, coord2, step2 = ref_table(lib, 100000)
        np.savez_compressed(os.path.join(HERE, "footprint.npz"), cum1000=cum1,
                            step1000=np.array(step1), cum100000_sub=cum2[::997],
                            cum100000_last=np.array(cum2[-1]), step100000=np.array(step2))
    if "proj3d" in names:
        np.savez_compressed(os.path.join(HERE, "proj3d.npz"), **proj3d_cases(lib))
    if "proj2d" in names:
        np.savez_compressed(os.path.join(HERE, "proj2d.npz"), **proj2d_cases(lib))
    if "sparse2d" in names:
        proj2d = np.load(os.path.join(HERE, "proj2d.npz"))
        np.savez_compressed(os.path.join(HERE, "sparse2d.npz"), **sparse2d_cases(lib, proj2d))
    if "dd3d" in names:
        np.savez_compressed(os.path.join(HERE, "dd3d.npz"), **dd3d_cases(lib))
    if "deg2" in names:
        proj3d = np.load(os.path.join(HERE, "proj3d.npz"))
        proj2d = np.load(os.path.join(HERE, "proj2d.npz"))
        np.savez_compressed(os.path.join(HERE, "deg2.npz"), **deg2_cases(lib, proj3d, proj2d))
    for f in names:
        print(f, os.path.getsize(os.path.join(HERE, f + ".npz")), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
