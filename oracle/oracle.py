"""ctypes front-end of the CPU checkers (test infrastructure only).

``Oracle`` wraps ``oracle/liboracle.so`` (the FP64 restatement in
trtt_oracle.c); ``Reference`` wraps ``oracle/_ref/libtrtt_ref.so`` (the
reference C plugin compiled from its own sources by oracle/Makefile) when that
library has been built.  Only tests, ``__graft_entry__.smoke()`` and the CPU
baseline of ``bench.py`` import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtrtt_ref.so")
REF_O3_SO = os.path.join(HERE, "_ref", "libtrtt_ref_o3.so")

dp = C.POINTER(C.c_double)
L, I, D = C.c_long, C.c_int, C.c_double

FP_CUBIC, FP_TABLE3D, FP_TABLE2D = 0, 1, 2


def _p(a):
    return a.ctypes.data_as(dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference where its sources exist)."""
    if force or not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


class LF2DParams(C.Structure):
    """Field order of lf2d_parameters_t (LongFessler2D.h:16-28)."""

    _fields_ = [("pixelSize", D), ("pixelStep", D), ("voxelSize", D), ("sourceDistance", D),
                ("detectorDistance", D), ("objectOffset1", D), ("objectOffset2", D),
                ("objectDimension1", L), ("objectDimension2", L), ("detectorDimension", L)]


class Oracle:
    """FP64 restatement of the projectors (trtt_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.orc_bspline3_cdf.argtypes = [D]
        lib.orc_bspline3_cdf.restype = D
        lib.orc_bspline3.argtypes = [D]
        lib.orc_bspline3.restype = D
        lib.orc_footprint_table.argtypes = [L, I, dp, dp, dp, dp]
        lib.orc_footprint_table.restype = None
        lib.orc_sb3d_ex.argtypes = [dp, dp, L, L, L, D, I, I, dp, L, D, D, L, L, D, D, dp, dp, I, I]
        lib.orc_sb3d_ex.restype = I
        lib.orc_sb2d_ex.argtypes = [dp, dp, L, L, D, I, I, dp, L, D, D, L, D, dp, dp, I, I]
        lib.orc_sb2d_ex.restype = I
        lib.orc_sb3d_views_fwd.argtypes = [dp, dp, L, L, L, D, I, I, dp, L, D, D, L, L, D, D, dp, dp,
                                           L, I]
        lib.orc_sb3d_views_fwd.restype = I
        lib.orc_sb3d_views_bwd.argtypes = [dp, dp, L, L, L, D, I, I, dp, L, D, D, L, L, D, D, dp, dp,
                                           L, I, L, L]
        lib.orc_sb3d_views_bwd.restype = I
        lib.orc_sb2d_views.argtypes = [dp, dp, L, L, D, I, I, dp, L, D, D, L, D, dp, dp, L, I, I]
        lib.orc_sb2d_views.restype = I
        lib.orc_lf2d_views.argtypes = [C.POINTER(LF2DParams), dp, dp, dp, dp, L, I, I]
        lib.orc_lf2d_ex.argtypes = [C.POINTER(LF2DParams), D, D, dp, dp, I, I, I]
        lib.orc_lf2d_ex.restype = None
        lib.orc_lf2d_views.restype = I
        lib.orc_sb3d_count_pairs.argtypes = [L, L, L, D, I, L, L, D, D, dp, dp, I]
        lib.orc_sb3d_count_pairs.restype = D
        lib.orc_lf2d.argtypes = [C.POINTER(LF2DParams), D, D, dp, dp, I, I]
        lib.orc_dd_pb3d.argtypes = [dp, dp, L, L, L, D, L, L, D, D, dp, dp, I, I]
        lib.orc_dd_pb3d.restype = I
        lib.orc_lf2d.restype = None
        self.lib = lib

    # -- scalar helpers -------------------------------------------------------
    def cdf(self, x: float) -> float:
        return self.lib.orc_bspline3_cdf(float(x))

    def bspline3(self, x: float) -> float:
        return self.lib.orc_bspline3(float(x))

    def footprint_table(self, m: int, deg: int = 3):
        fp, cum, coord, step = np.zeros(m), np.zeros(m), np.zeros(m), np.zeros(1)
        self.lib.orc_footprint_table(m, deg, _p(fp), _p(cum), _p(coord), _p(step))
        return fp, cum, coord, float(step[0])

    # -- projectors -----------------------------------------------------------
    def sb3d(self, out, inp, nx, ny, nz, s, nu, nv, us, vs, xos, xsd, cone, job,
             table=None, deg=3):
        """One view of trtt_{PB,CB}3D[_cubic]; accumulates into ``out``."""
        kind, cum, m, step, c0 = self._fp(table, FP_TABLE3D)
        xos, xsd, inp = _f64(xos), _f64(xsd), _f64(inp)
        assert out.dtype == np.float64 and out.flags.c_contiguous
        r = self.lib.orc_sb3d_ex(_p(out), _p(inp), nx, ny, nz, s, deg, kind, _p(cum), m, step, c0,
                                 nu, nv, us, vs, _p(xos), _p(xsd), int(cone), int(job))
        if r != 0:
            raise RuntimeError("oracle projector failure")
        return out

    def sb2d(self, out, inp, nx, ny, s, nv, vs, xos, xsd, fan, job, table=None, deg=3):
        """One view of trtt_{PB,FB}2D[_cubic]; accumulates into ``out``."""
        kind, cum, m, step, c0 = self._fp(table, FP_TABLE2D)
        xos, xsd, inp = _f64(xos), _f64(xsd), _f64(inp)
        assert out.dtype == np.float64 and out.flags.c_contiguous
        r = self.lib.orc_sb2d_ex(_p(out), _p(inp), nx, ny, s, deg, kind, _p(cum), m, step, c0, nv, vs,
                                 _p(xos), _p(xsd), int(fan), int(job))
        if r != 0:
            raise RuntimeError("oracle projector failure")
        return out

    def views_fwd3d(self, vol, nx, ny, nz, s, nu, nv, us, vs, xos, xsd, cone, table=None, deg=3):
        """Forward projection of every view (OpenMP over views)."""
        kind, cum, m, step, c0 = self._fp(table, FP_TABLE3D)
        nview = xos.shape[0]
        sino = np.zeros((nview, nv, nu))
        r = self.lib.orc_sb3d_views_fwd(_p(sino), _p(_f64(vol)), nx, ny, nz, s, deg, kind, _p(cum), m,
                                        step, c0, nu, nv, us, vs, _p(_f64(xos)), _p(_f64(xsd)),
                                        nview, int(cone))
        if r != 0:
            raise RuntimeError("oracle projector failure")
        return sino

    def views_bwd3d(self, sino, nx, ny, nz, s, nu, nv, us, vs, xos, xsd, cone, table=None, deg=3,
                    k0=0, k1=None):
        """Back projection of every view, restricted to the slices
        k0 <= k < k1 (OpenMP over views): returns the slab [k1-k0][ny][nx]."""
        kind, cum, m, step, c0 = self._fp(table, FP_TABLE3D)
        k1 = nz if k1 is None else k1
        nview = xos.shape[0]
        sino = _f64(sino)
        assert sino.size == nview * nv * nu
        vol = np.zeros((k1 - k0, ny, nx))
        r = self.lib.orc_sb3d_views_bwd(_p(vol), _p(sino), nx, ny, nz, s, deg, kind, _p(cum), m,
                                        step, c0, nu, nv, us, vs, _p(_f64(xos)), _p(_f64(xsd)),
                                        nview, int(cone), int(k0), int(k1))
        if r != 0:
            raise RuntimeError("oracle projector failure")
        return vol

    def views2d(self, inp, nx, ny, s, nv, vs, xos, xsd, fan, job, table=None, deg=3):
        """2-D projector over every view (OpenMP over views): job 0 returns
        the sinogram [view][nv] of image ``inp``; job 1 the back projection
        of sinogram ``inp``."""
        kind, cum, m, step, c0 = self._fp(table, FP_TABLE2D)
        nview = xos.shape[0]
        inp = _f64(inp)
        out = np.zeros((nview, nv)) if job == 0 else np.zeros((ny, nx))
        r = self.lib.orc_sb2d_views(_p(out), _p(inp), nx, ny, s, deg, kind, _p(cum), m, step, c0,
                                    nv, vs, _p(_f64(xos)), _p(_f64(xsd)), nview, int(fan),
                                    int(job))
        if r != 0:
            raise RuntimeError("oracle projector failure")
        return out

    def lf2d_views(self, params: LF2DParams, beta, offset, src, job, line=False):
        """Long & Fessler projector over every view: job 0 returns the
        sinogram [view][nd] of image ``src``; job 1 the back projection."""
        beta, offset = _f64(beta), _f64(offset)
        nview, nd = beta.size, params.detectorDimension
        out = (np.zeros((nview, nd)) if job == 0
               else np.zeros((params.objectDimension2, params.objectDimension1)))
        r = self.lib.orc_lf2d_views(C.byref(params), _p(beta), _p(offset), _p(_f64(src)), _p(out),
                                    nview, int(job), int(line))
        if r != 0:
            raise RuntimeError("oracle projector failure")
        return out

    def count_pairs(self, nx, ny, nz, s, nu, nv, us, vs, xos, xsd, cone, deg=3) -> float:
        return self.lib.orc_sb3d_count_pairs(nx, ny, nz, s, deg, nu, nv, us, vs, _p(_f64(xos)),
                                             _p(_f64(xsd)), int(cone))

    def dd_pb3d(self, out, inp, nx, ny, nz, s, nu, nv, us, vs, xos, xsd, job, adjoint=False):
        """One view of trtt_DD_PB3D; accumulates into ``out``.  job != 0 is
        the reference's transpose, or the exact adjoint with ``adjoint``."""
        xos, xsd, inp = _f64(xos), _f64(xsd), _f64(inp)
        assert out.dtype == np.float64 and out.flags.c_contiguous
        self.lib.orc_dd_pb3d(_p(out), _p(inp), nx, ny, nz, s, nu, nv, us, vs, _p(xos), _p(xsd),
                             int(job), int(adjoint))
        return out

    def lf2d(self, params: LF2DParams, beta, offset, src, dst, clr, job, line=False):
        """lf2d_direct / lf2d_adjoint (job 0 / 1); ``line`` selects the
        line-integral model (lf2d_line_integ_direct / _adjoint)."""
        src = _f64(src)
        assert dst.dtype == np.float64 and dst.flags.c_contiguous
        self.lib.orc_lf2d_ex(C.byref(params), float(beta), float(offset), _p(src), _p(dst),
                             int(clr), int(job), int(line))
        return dst

    @staticmethod
    def sparse2d_coo(column, nviews, nx, ny, decimals=6):
        """Sparse system matrix assembled as Ytrtt_2D_projectors.i:252-290
        and :326-350 do it: for every view, pixel (j, then i) and bin, the
        non-zero coefficients of ``column(view, i, j)`` (the nv coefficients
        of one pixel, e.g. a projection of a one-hot image), written as "%f"
        text and read back when ``decimals`` is 6.  Rows (view, bin) and
        columns j nx + i are 1-based.  Pure Python: small cases only."""
        coefs, rows, cols = [], [], []
        for v in range(nviews):
            for j in range(ny):
                for i in range(nx):
                    c = np.asarray(column(v, i, j), dtype=np.float64)
                    nz = np.nonzero(c != 0.0)[0]
                    vals = c[nz]
                    if decimals == 6:
                        vals = np.array([float("%f" % x) for x in vals])
                    coefs.append(vals)
                    rows.append(v * c.size + nz + 1)
                    cols.append(np.full(nz.size, j * nx + i + 1))
        cat = (lambda a, t: np.concatenate(a).astype(t) if a else np.zeros(0, t))
        return cat(coefs, np.float64), cat(rows, np.int64), cat(cols, np.int64)

    @staticmethod
    def _fp(table, kind):
        if table is None:
            return FP_CUBIC, np.zeros(1), 1, 1.0, 0.0
        cum, coord, step = table
        cum = _f64(cum)
        return kind, cum, cum.size, float(step), float(coord[0])


class Reference:
    """The reference C plugin built from its sources (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        tab3 = [dp, dp, L, L, L, D, I, dp, dp, L, D, dp, L, L, D, D, dp, dp, I]
        cub3 = [dp, dp, L, L, L, D, I, L, L, D, D, dp, dp, I]
        tab2 = [dp, dp, L, L, D, I, dp, dp, L, D, dp, L, D, dp, dp, I]
        cub2 = [dp, dp, L, L, D, I, L, D, dp, dp, I]
        for n, a in (("trtt_PB3D", tab3), ("trtt_CB3D", tab3), ("trtt_PB3D_cubic", cub3),
                     ("trtt_CB3D_cubic", cub3), ("trtt_PB2D", tab2), ("trtt_FB2D", tab2),
                     ("trtt_PB2D_cubic", cub2), ("trtt_FB2D_cubic", cub2)):
            getattr(lib, n).argtypes = a
            getattr(lib, n).restype = I
        self.lib = lib

    @staticmethod
    def available(path: str = REF_SO) -> bool:
        return os.path.exists(path)

    def CB3D_cubic(self, out, inp, nx, ny, nz, s, nu, nv, us, vs, xos, xsd, job, deg=3):
        return self.lib.trtt_CB3D_cubic(_p(out), _p(_f64(inp)), nx, ny, nz, s, deg, nu, nv, us, vs,
                                        _p(_f64(xos)), _p(_f64(xsd)), job)

    def PB3D_cubic(self, out, inp, nx, ny, nz, s, nu, nv, us, vs, xos, xsd, job, deg=3):
        return self.lib.trtt_PB3D_cubic(_p(out), _p(_f64(inp)), nx, ny, nz, s, deg, nu, nv, us, vs,
                                        _p(_f64(xos)), _p(_f64(xsd)), job)

    def PB2D_cubic(self, out, inp, nx, ny, s, nv, vs, xos, xsd, job, deg=3):
        return self.lib.trtt_PB2D_cubic(_p(out), _p(_f64(inp)), nx, ny, s, deg, nv, vs,
                                        _p(_f64(xos)), _p(_f64(xsd)), job)

    def FB2D_cubic(self, out, inp, nx, ny, s, nv, vs, xos, xsd, job, deg=3):
        return self.lib.trtt_FB2D_cubic(_p(out), _p(_f64(inp)), nx, ny, s, deg, nv, vs,
                                        _p(_f64(xos)), _p(_f64(xsd)), job)
