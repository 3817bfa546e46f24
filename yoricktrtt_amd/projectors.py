"""Host mirror of the TRTT operator layer over the gfx950 C ABI.

Two levels, both calling ``libtrtt_hip.so`` only:

* :class:`SingleProjector` mirrors the Yorick evaluators
  ``trtt_single_3D_projector`` / ``_trtt_single_3D_projector``
  (Ytrtt_3D_projectors.i:51-187), ``trtt_single_2D_projector``
  (Ytrtt_2D_projectors.i:376-510) and ``trtt_single_Long_Fessler_2D_projector``
  (Ytrtt_2D_projectors_Long_Fessler.i:134-280): one view, host FP64 arrays,
  ``job`` 0 (direct) or 1 (transpose), a fresh output array per call, and
  the same error codes (Ytrtt_error.i:47-62).
* :class:`Projector` is the device-resident batched operator: every view of
  every phase in one launch, on torch tensors already in HBM.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .geometry import (Geometry, XForm3D, detector_xi, footprint_table, object_xi,
                       projector_coefs)

# Geometry and model constants (Ytrtt_common.i:62-77)
TRTT_SPLINE_DRIVEN = 1
TRTT_DISTANCE_DRIVEN = 2
TRTT_LONG_FESSLER = 3
TRTT_PARALLEL_BEAM = 3
TRTT_CONE_BEAM = 4
TRTT_FAN_BEAM = 5

# Error codes (Ytrtt_error.i:47-62)
TRTT_ERR_OUT_OF_RANGE_INDEX = 5
TRTT_ERR_NOT_TOMOBJ = 7
TRTT_ERR_NOT_TOMDATA = 8
TRTT_ERR_PROJECTOR_FAILURE = 11
TRTT_ERR_UNSUPPORTED_JOB = 12
TRTT_ERR_BAD_GEOMETRY = 13


class TrttError(RuntimeError):
    """A TRTT status error carrying the reference error code."""

    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"TRTT error {code}: {msg}")
        self.code = code


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_lib.dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class TomObj:
    """Tomographic object (Ytrtt_tomobj.i:341-465): sizes, B-spline degree,
    footprint table and the voxel->metric transform ``xi``."""

    nx: int
    ny: int
    nz: int = 1
    nt: int = 1
    s_scl: float = 1.0
    s_deg: int = 3
    size_footprint: int = 10000
    footprint: np.ndarray = field(init=False)
    cum_footprint: np.ndarray = field(init=False)
    coord_footprint: np.ndarray = field(init=False)
    step_footprint: float = field(init=False)
    xi: XForm3D = field(init=False)

    def __post_init__(self):
        fp, cum, coord, step = footprint_table(self.size_footprint, self.s_deg)
        self.footprint, self.cum_footprint, self.coord_footprint = fp, cum, coord
        self.step_footprint = step
        self.xi = object_xi(self.nx, self.ny, self.nz, self.s_scl)


@dataclass
class TomData:
    """One projection (Ytrtt_tomdata.i:210-288): detector sizes and scales, the
    time index and the xos/xotilt/xsd transforms of its view."""

    nu: int
    nv: int
    u_scl: float = 1.0
    v_scl: float = 1.0
    t_index: float = 0.0
    xos: XForm3D = field(default_factory=XForm3D)
    xotilt: XForm3D = field(default_factory=XForm3D)
    xsd: XForm3D = field(default_factory=XForm3D)
    theta: float = 0.0
    Rsc: float | None = None
    Rsd: float | None = None
    voff: float = 0.0
    xoff: float = 0.0
    yoff: float = 0.0

    @property
    def xi(self) -> XForm3D:
        return detector_xi(self.nu, self.nv, self.u_scl, self.v_scl)


class SingleProjector:
    """Evaluator of one view, calling the reference-named C ABI symbols.

    ``model`` selects the footprint: ``"table"`` (trtt_PB/CB/FB*D, what the
    Yorick layer uses) or ``"cubic"`` (the closed-form ``*_cubic`` variants).
    """

    def __init__(self, X: TomObj, Yk: TomData, mode: int, model: str = "table",
                 is3d: bool | None = None, lf: bool = False):
        if not isinstance(X, TomObj):
            raise TrttError(TRTT_ERR_NOT_TOMOBJ)
        if not isinstance(Yk, TomData):
            raise TrttError(TRTT_ERR_NOT_TOMDATA)
        if mode not in (TRTT_PARALLEL_BEAM, TRTT_FAN_BEAM, TRTT_CONE_BEAM):
            raise TrttError(TRTT_ERR_BAD_GEOMETRY)
        if not (0 <= Yk.t_index <= X.nt):
            raise TrttError(TRTT_ERR_OUT_OF_RANGE_INDEX,
                            "Incompatible TOMDATA 't_index' with dimension 'nt' of the TOMOBJ.")
        self.X, self.Yk, self.mode, self.model, self.lf = X, Yk, mode, model, lf
        self.is3d = (Yk.nu > 1) if is3d is None else is3d
        self.xos_coefs, self.xsd_coefs = projector_coefs(X.xi, Yk.xi, Yk.xos, Yk.xotilt, Yk.xsd)
        self.lib = _lib.load()
        if lf:
            self.lf_param = _lib.LF2DParameters(Yk.v_scl, Yk.v_scl, X.s_scl, Yk.Rsc, Yk.Rsd - Yk.Rsc,
                                                Yk.xoff, Yk.yoff, X.nx, X.ny, Yk.nv)
            self.beta = Yk.theta - math.pi / 2
            self.offset = Yk.voff

    def __call__(self, x, job: int = 0) -> np.ndarray:
        X, Yk = self.X, self.Yk
        if job not in (0, 1):
            raise TrttError(TRTT_ERR_UNSUPPORTED_JOB)
        x = _f64(x)
        if self.lf:
            out = np.zeros(Yk.nv if job == 0 else (X.ny, X.nx))
            f = self.lib.lf2d_direct if job == 0 else self.lib.lf2d_adjoint
            f(C.byref(self.lf_param), self.beta, self.offset, _ptr(x), _ptr(out), 1)
            err = _lib.last_error()
            if err:
                raise TrttError(TRTT_ERR_PROJECTOR_FAILURE, err)
            return out
        if self.is3d:
            out = np.zeros((Yk.nv, Yk.nu) if job == 0 else (X.nz, X.ny, X.nx))
            pre = "CB3D" if self.mode != TRTT_PARALLEL_BEAM else "PB3D"
            if self.model == "cubic":
                r = getattr(self.lib, f"trtt_{pre}_cubic")(
                    _ptr(out), _ptr(x), X.nx, X.ny, X.nz, X.s_scl, X.s_deg, Yk.nu, Yk.nv, Yk.u_scl,
                    Yk.v_scl, _ptr(self.xos_coefs), _ptr(self.xsd_coefs), job)
            else:
                r = getattr(self.lib, f"trtt_{pre}")(
                    _ptr(out), _ptr(x), X.nx, X.ny, X.nz, X.s_scl, X.s_deg, _ptr(X.footprint),
                    _ptr(X.cum_footprint), X.size_footprint, X.step_footprint,
                    _ptr(X.coord_footprint), Yk.nu, Yk.nv, Yk.u_scl, Yk.v_scl,
                    _ptr(self.xos_coefs), _ptr(self.xsd_coefs), job)
        else:
            out = np.zeros(Yk.nv if job == 0 else (X.ny, X.nx))
            pre = "FB2D" if self.mode != TRTT_PARALLEL_BEAM else "PB2D"
            if self.model == "cubic":
                r = getattr(self.lib, f"trtt_{pre}_cubic")(
                    _ptr(out), _ptr(x), X.nx, X.ny, X.s_scl, X.s_deg, Yk.nv, Yk.v_scl,
                    _ptr(self.xos_coefs), _ptr(self.xsd_coefs), job)
            else:
                r = getattr(self.lib, f"trtt_{pre}")(
                    _ptr(out), _ptr(x), X.nx, X.ny, X.s_scl, X.s_deg, _ptr(X.footprint),
                    _ptr(X.cum_footprint), X.size_footprint, X.step_footprint,
                    _ptr(X.coord_footprint), Yk.nv, Yk.v_scl, _ptr(self.xos_coefs),
                    _ptr(self.xsd_coefs), job)
        if r == _lib.TRTT_FAILURE:
            raise TrttError(TRTT_ERR_PROJECTOR_FAILURE, _lib.last_error())
        return out


def set_precision(fp64: bool) -> None:
    """Precision of the reference-named symbols (default FP32)."""
    if _lib.load().trtt_hip_set_precision(_lib.F64 if fp64 else _lib.F32) != 0:
        raise RuntimeError(_lib.last_error())


_KIND = {"parallel3d": _lib.PB3D, "cone": _lib.CB3D, "parallel2d": _lib.PB2D, "fan": _lib.FB2D}


class Projector:
    """Device-resident batched projector over every view of a geometry.

    ``forward(vol, sino)`` accumulates ``sino[k] += A_k vol`` and
    ``backward(sino, vol)`` accumulates ``vol += sum_k A_k' sino[k]`` on torch
    tensors in HBM (float32 or float64 per ``precision``).  For a 4-D series
    give ``nt`` and per-view ``t0, t1, w0, w1``: view k sees
    ``w0[k] x[t0[k]] + w1[k] x[t1[k]]`` (Ytrtt_3D_projectors.i:254-375).
    """

    def __init__(self, geom: Geometry, model: str = "cubic", precision: str = "f32", nt: int = 1,
                 t0=None, t1=None, w0=None, w1=None, size_footprint: int = 10000, table=None):
        import torch  # device memory and streams only

        self.torch = torch
        self.geom = geom
        self.lib = _lib.load()
        self.precision = _lib.F64 if precision == "f64" else _lib.F32
        self.dtype = torch.float64 if self.precision == _lib.F64 else torch.float32
        self.nt = int(nt)
        kind = _KIND[geom.kind]
        self.kind = kind
        xos, xsd = _f64(geom.xos), _f64(geom.xsd)
        self._keep = [xos, xsd]
        if model == "table":
            # the caller's (cum, coord, step), or the table of Ytrtt_tomobj.i:140-196
            if table is not None:
                cum, coord, step = table
                cum, coord, step = _f64(cum), _f64(coord), float(step)
            else:
                _, cum, coord, step = footprint_table(size_footprint, geom.s_deg)
            self._keep.append(cum)
            cum_p, m, c0 = _ptr(cum), cum.size, float(coord[0])
        else:
            cum_p, m, step, c0 = None, 0, 0.0, 0.0

        def iarr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.int32)
            self._keep.append(a)
            return a.ctypes.data_as(_lib.ip)

        def darr(a):
            if a is None:
                return None
            a = _f64(a)
            self._keep.append(a)
            return _ptr(a)

        self.handle = self.lib.trtt_hip_plan_create(
            kind, self.precision, geom.nx, geom.ny, geom.nz, self.nt, geom.s_scl, geom.s_deg,
            geom.nu, geom.nv, geom.u_scl, geom.v_scl, _ptr(xos), _ptr(xsd), geom.nviews, cum_p, m,
            step, c0, iarr(t0), iarr(t1), darr(w0), darr(w1))
        if not self.handle:
            raise RuntimeError("plan creation failed: " + _lib.last_error())

    @property
    def vol_shape(self):
        g = self.geom
        base = (g.nz, g.ny, g.nx) if g.is3d else (g.ny, g.nx)
        return ((self.nt,) + base) if self.nt > 1 else base

    @property
    def sino_shape(self):
        g = self.geom
        return (g.nviews, g.nv, g.nu) if g.is3d else (g.nviews, g.nv)

    def _check(self, t, shape):
        torch = self.torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError("expected a device tensor")
        if t.dtype != self.dtype or not t.is_contiguous() or tuple(t.shape) != tuple(shape):
            raise ValueError(f"expected contiguous {self.dtype} of shape {shape}, got "
                             f"{t.dtype} {tuple(t.shape)}")

    def _stream(self, stream):
        s = stream if stream is not None else self.torch.cuda.current_stream()
        return C.c_void_p(s.cuda_stream)

    def forward(self, vol, sino=None, stream=None):
        if sino is None:
            sino = self.torch.zeros(self.sino_shape, dtype=self.dtype, device=vol.device)
        self._check(vol, self.vol_shape)
        self._check(sino, self.sino_shape)
        if self.lib.trtt_hip_forward(self.handle, C.c_void_p(vol.data_ptr()),
                                     C.c_void_p(sino.data_ptr()), self._stream(stream)) != 0:
            raise RuntimeError(_lib.last_error())
        return sino

    def backward(self, sino, vol=None, stream=None):
        if vol is None:
            vol = self.torch.zeros(self.vol_shape, dtype=self.dtype, device=sino.device)
        self._check(vol, self.vol_shape)
        self._check(sino, self.sino_shape)
        if self.lib.trtt_hip_backward(self.handle, C.c_void_p(sino.data_ptr()),
                                      C.c_void_p(vol.data_ptr()), self._stream(stream)) != 0:
            raise RuntimeError(_lib.last_error())
        return vol

    def _check_fx(self, fx):
        torch = self.torch
        if (not isinstance(fx, torch.Tensor) or not fx.is_cuda or fx.dtype != torch.float64
                or fx.numel() != 1):
            raise ValueError("fx must be a 1-element float64 device tensor")

    def weighted_residual(self, sino, y, w, fx, stream=None):
        """sino <- w*(sino-y) and fx += 0.5 sum w (sino-y)^2 (fx: 1-element float64).
        sino, y and w (optional) are contiguous device tensors of the plan's
        type and sinogram shape."""
        self._check(sino, self.sino_shape)
        self._check(y, self.sino_shape)
        if w is not None:
            self._check(w, self.sino_shape)
        self._check_fx(fx)
        wp = C.c_void_p(w.data_ptr()) if w is not None else None
        if self.lib.trtt_hip_weighted_residual(self.handle, C.c_void_p(sino.data_ptr()),
                                               C.c_void_p(y.data_ptr()), wp,
                                               C.c_void_p(fx.data_ptr()),
                                               self._stream(stream)) != 0:
            raise RuntimeError(_lib.last_error())

    def eval_set_data(self, y, w=None) -> None:
        """Upload host data y (and weights w) for :meth:`eval_host`
        (trtt_hip_eval_set_data; forms WCyT = sum_k A_k' W_k y_k)."""
        y = _f64(y).reshape(-1)
        if y.size != int(np.prod(self.sino_shape)):
            raise ValueError(f"y must hold {self.sino_shape} values")
        self._eval_keep = [y]
        wp = None
        if w is not None:
            w = _f64(w).reshape(-1)
            if w.size != y.size:
                raise ValueError("w must have the shape of y")
            self._eval_keep.append(w)
            wp = _ptr(w)
        if self.lib.trtt_hip_eval_set_data(self.handle, _ptr(y), wp) != 0:
            raise RuntimeError(_lib.last_error())

    def eval_host(self, x, split: bool = False):
        """f(x) and g(x) of the data term on host FP64 arrays (trtt_hip_eval):
        g = sum A'W(Ax - y), or sum A'WAx - WCyT with ``split`` (the
        reference's form, Ytrtt4D_optim.i:437-446)."""
        x = _f64(x)
        if x.shape != tuple(self.vol_shape):
            raise ValueError(f"x must have shape {self.vol_shape}")
        g = np.zeros_like(x)
        fx = C.c_double(0.0)
        if self.lib.trtt_hip_eval(self.handle, C.byref(fx), _ptr(g), _ptr(x), int(split)) != 0:
            raise RuntimeError(_lib.last_error())
        return fx.value, g

    def set_timing(self, on: bool = True) -> None:
        """Record HIP events around the main axial kernels of each call
        (not under graph capture)."""
        if self.lib.trtt_hip_plan_set_timing(self.handle, int(on)) != 0:
            raise RuntimeError(_lib.last_error())

    def kernel_time(self, job: int):
        """(ms, launches) of the main kernel of the last forward (job 0) or
        backward (job 1) call, from the plan's HIP events."""
        ms = C.c_double(0.0)
        n = C.c_int(0)
        if self.lib.trtt_hip_plan_kernel_time(self.handle, int(job), C.byref(ms), C.byref(n)) != 0:
            raise RuntimeError(_lib.last_error())
        return ms.value, n.value

    @property
    def path(self) -> int:
        """0: general kernels; W > 0: axial closed-form kernels (width bound W)."""
        return self.lib.trtt_hip_plan_path(self.handle)

    def error_flag(self) -> int:
        return self.lib.trtt_hip_plan_error_flag(self.handle)

    def sparse_matrix(self, decimals: int = 6) -> "SparseMatrix":
        """System matrix of this 2-D plan on the device (_trtt_2D_sparse_matrix,
        Ytrtt_2D_projectors.i:295-369); ``decimals=6`` keeps the reference's
        6-decimal coefficients, ``-1`` exact ones."""
        return SparseMatrix(self, decimals)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.trtt_hip_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LFProjector(Projector):
    """Batched Long & Fessler projector (LongFessler2D.c) over many views.
    ``line_integ`` selects the line-integral model (the trapezoid sampled at
    the bin centres, lf2d_line_integ_direct / _adjoint,
    LongFessler2D.c:143-470) instead of the bin-integrated one."""

    def __init__(self, params: _lib.LF2DParameters, beta, offset=None, precision: str = "f32",
                 line_integ: bool = False):
        import torch

        self.torch = torch
        self.lib = _lib.load()
        self.precision = _lib.F64 if precision == "f64" else _lib.F32
        self.dtype = torch.float64 if self.precision == _lib.F64 else torch.float32
        self.params = params
        beta = _f64(beta)
        offset = _f64(np.zeros_like(beta) if offset is None else offset)
        self._keep = [beta, offset]
        self.nt = 1
        self.nviews = beta.size
        self.handle = self.lib.trtt_hip_plan_create_lf2d(self.precision, C.byref(params), _ptr(beta),
                                                         _ptr(offset), beta.size)
        if not self.handle:
            raise RuntimeError("plan creation failed: " + _lib.last_error())
        if line_integ and self.lib.trtt_hip_plan_set_lf_line(self.handle, 1) != 0:
            raise RuntimeError(_lib.last_error())

    @property
    def vol_shape(self):
        return (self.params.objectDimension2, self.params.objectDimension1)

    @property
    def sino_shape(self):
        return (self.nviews, self.params.detectorDimension)


class SparseMatrix:
    """Device-resident sparse system matrix of a 2-D plan.

    Rows are (view, bin) = view * nv + bin, columns pixels j * nx + i.
    ``mvmult(img)`` returns ``M img`` as a sinogram and ``mvmult(sino, 1)``
    returns ``M' sino`` as an image, like Yeti's ``mvmult(M, x[, 1])`` on the
    matrix of Ytrtt_2D_projectors.i:362-366; ``forward`` / ``backward``
    accumulate into given tensors.  ``coo()`` returns the entries in the
    reference's order with 1-based indices (MCops_coefs, MCops_row_indices,
    MCops_col_indices).
    """

    def __init__(self, proj: Projector, decimals: int = 6):
        self.proj = proj
        self.lib = proj.lib
        self.torch = proj.torch
        self.handle = self.lib.trtt_hip_sparse_create(proj.handle, int(decimals))
        if not self.handle:
            raise RuntimeError("sparse matrix creation failed: " + _lib.last_error())

    @property
    def nnz(self) -> int:
        return int(self.lib.trtt_hip_sparse_nnz(self.handle))

    # operator interface of the plan (the data term runs on either)
    @property
    def vol_shape(self):
        return self.proj.vol_shape

    @property
    def sino_shape(self):
        return self.proj.sino_shape

    @property
    def dtype(self):
        return self.proj.dtype

    def weighted_residual(self, sino, y, w, fx, stream=None):
        self.proj.weighted_residual(sino, y, w, fx, stream)

    def error_flag(self) -> int:
        return self.proj.error_flag()

    def coo(self):
        n = self.nnz
        coefs = np.empty(n, dtype=np.float64)
        rows = np.empty(n, dtype=np.int64)
        cols = np.empty(n, dtype=np.int64)
        if self.lib.trtt_hip_sparse_coo(self.handle, _ptr(coefs), rows.ctypes.data_as(_lib.lp),
                                        cols.ctypes.data_as(_lib.lp)) != 0:
            raise RuntimeError(_lib.last_error())
        return coefs, rows, cols

    def forward(self, img, sino=None, stream=None):
        P = self.proj
        if sino is None:
            sino = self.torch.zeros(P.sino_shape, dtype=P.dtype, device=img.device)
        P._check(img, P.vol_shape)
        P._check(sino, P.sino_shape)
        if self.lib.trtt_hip_sparse_forward(self.handle, C.c_void_p(img.data_ptr()),
                                            C.c_void_p(sino.data_ptr()), P._stream(stream)) != 0:
            raise RuntimeError(_lib.last_error())
        return sino

    def backward(self, sino, img=None, stream=None):
        P = self.proj
        if img is None:
            img = self.torch.zeros(P.vol_shape, dtype=P.dtype, device=sino.device)
        P._check(img, P.vol_shape)
        P._check(sino, P.sino_shape)
        if self.lib.trtt_hip_sparse_backward(self.handle, C.c_void_p(sino.data_ptr()),
                                             C.c_void_p(img.data_ptr()), P._stream(stream)) != 0:
            raise RuntimeError(_lib.last_error())
        return img

    def mvmult(self, x, transpose: int = 0):
        return self.backward(x) if transpose else self.forward(x)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.trtt_hip_sparse_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def temporal_weights(t_index, nt: int, t_deg: int = 1):
    """Taps of the B_k operators: periodic B-spline interpolation in time.

    View k at time ``t_index[k]`` (phase units) reads phases
    j = j0, ..., j0 + t_deg (modulo nt), j0 = floor(t - (t_deg + 1)/2) + 1,
    with weights beta^t_deg(t - j).  Degree 1 is the linear interpolation of
    Ytrtt_3D_projectors.i:301-311 (phases floor(t) and floor(t) + 1 modulo
    nt).  Degrees 2 and 3 take the place of SPL's
    spl_interp_coefs_new_spline (Ytrtt_2D_projectors.i:747), whose source is
    not available: periodic boundaries here.  Returns idx (int32) and w
    (float64), both [nviews][t_deg + 1]."""
    from .geometry import bspline

    t = np.asarray(t_index, dtype=np.float64).reshape(-1)
    if t_deg < 1 or t_deg > 3:
        raise ValueError("t_deg must be 1, 2 or 3")
    if np.any(t < 0) or np.any(t > nt):
        raise ValueError("time index out of [0, nt]")
    j0 = np.floor(t - 0.5 * (t_deg + 1)).astype(np.int64) + 1
    j = j0[:, None] + np.arange(t_deg + 1)[None, :]
    w = bspline(t[:, None] - j, t_deg)
    return np.ascontiguousarray(j % nt, dtype=np.int32), np.ascontiguousarray(w)


class TimeInterpolator:
    """B_k of every view on the device (trtt_hip_tinterp_*): ``forward(x)``
    maps a series [nt, ...] to one frame per view [nviews, ...];
    ``backward(frames)`` is the transpose (_trtt_2D_time_interpolator,
    Ytrtt_2D_projectors.i:761-814)."""

    def __init__(self, frame_shape, nt: int, idx, w, precision: str = "f32"):
        import torch

        self.torch = torch
        self.lib = _lib.load()
        self.precision = _lib.F64 if precision == "f64" else _lib.F32
        self.dtype = torch.float64 if self.precision == _lib.F64 else torch.float32
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        w = _f64(w)
        self.frame_shape = tuple(frame_shape)
        self.nt, (self.nviews, self.taps) = int(nt), idx.shape
        npix = int(np.prod(self.frame_shape))
        self.handle = self.lib.trtt_hip_tinterp_create(self.precision, npix, self.nt, self.nviews,
                                                       self.taps, idx.ctypes.data_as(_lib.ip),
                                                       _ptr(w))
        if not self.handle:
            raise RuntimeError("temporal interpolator creation failed: " + _lib.last_error())

    def _apply(self, job, src, dst, stream):
        s = stream if stream is not None else self.torch.cuda.current_stream()
        if self.lib.trtt_hip_tinterp_apply(self.handle, job, C.c_void_p(src.data_ptr()),
                                           C.c_void_p(dst.data_ptr()),
                                           C.c_void_p(s.cuda_stream)) != 0:
            raise RuntimeError(_lib.last_error())
        return dst

    def forward(self, x, frames=None, stream=None):
        if frames is None:
            frames = self.torch.zeros((self.nviews,) + self.frame_shape, dtype=self.dtype,
                                      device=x.device)
        return self._apply(0, x, frames, stream)

    def backward(self, frames, x=None, stream=None):
        if x is None:
            x = self.torch.zeros((self.nt,) + self.frame_shape, dtype=self.dtype,
                                 device=frames.device)
        return self._apply(1, frames, x, stream)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.trtt_hip_tinterp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DynamicProjector:
    """C_k = A_k B_k over every view of a geometry for a series of nt
    frames (Ytrtt_2D_projectors.i:695-814, Tomography_2D_dyn): the frames of
    the views are interpolated by a :class:`TimeInterpolator` and projected
    by a plan with one frame per view; the transpose runs back the same way.
    Linear interpolation (``t_deg`` 1) needs no frame buffer: the plan's own
    two-phase weights carry it."""

    def __init__(self, geom: Geometry, nt: int, t_index, t_deg: int = 1, model: str = "cubic",
                 precision: str = "f32"):
        import torch

        self.torch = torch
        self.geom = geom
        self.nt = int(nt)
        idx, w = temporal_weights(t_index, nt, t_deg)
        self.linear = t_deg == 1
        V = geom.nviews
        if self.linear:
            self.P = Projector(geom, model=model, precision=precision, nt=nt, t0=idx[:, 0],
                               t1=idx[:, 1], w0=w[:, 0], w1=w[:, 1])
            self.B = None
        else:
            one = np.arange(V, dtype=np.int32)
            self.P = Projector(geom, model=model, precision=precision, nt=V, t0=one, t1=one,
                               w0=np.ones(V), w1=np.zeros(V))
            frame = self.P.vol_shape[1:]
            self.B = TimeInterpolator(frame, nt, idx, w, precision)
            self.frames = torch.zeros((V,) + tuple(frame), dtype=self.P.dtype, device="cuda")
        self.dtype = self.P.dtype

    @property
    def vol_shape(self):
        g = self.geom
        return (self.nt,) + ((g.nz, g.ny, g.nx) if g.is3d else (g.ny, g.nx))

    @property
    def sino_shape(self):
        return self.P.sino_shape

    def _zero_frames(self, stream):
        if stream is None:
            self.frames.zero_()
        else:
            with self.torch.cuda.stream(stream):
                self.frames.zero_()

    def forward(self, x, sino=None, stream=None):
        if self.linear:
            return self.P.forward(x, sino, stream)
        self._zero_frames(stream)
        self.B.forward(x, self.frames, stream)
        return self.P.forward(self.frames, sino, stream)

    def backward(self, sino, x=None, stream=None):
        if self.linear:
            return self.P.backward(sino, x, stream)
        if x is None:
            x = self.torch.zeros(self.vol_shape, dtype=self.dtype, device=sino.device)
        self._zero_frames(stream)
        self.P.backward(sino, self.frames, stream)
        return self.B.backward(self.frames, x, stream)

    def weighted_residual(self, sino, y, w, fx, stream=None):
        self.P.weighted_residual(sino, y, w, fx, stream)

    def error_flag(self) -> int:
        return self.P.error_flag()

    def close(self):
        self.P.close()
        if self.B is not None:
            self.B.close()


class DDProjector:
    """Distance-driven parallel-beam projector over every view of a
    ``parallel3d`` geometry on the device (batched trtt_DD_PB3D,
    trtt_3D_projectors.c:1077-1173).  ``backward(..., exact=True)`` is the
    adjoint of ``forward``; ``exact=False`` reproduces the reference's
    transpose, which is not (trtt_3D_projectors.c:909-916)."""

    def __init__(self, geom: Geometry, precision: str = "f32"):
        import torch

        if geom.kind != "parallel3d":
            raise ValueError("the distance-driven projector is parallel-beam 3-D")
        self.torch = torch
        self.geom = geom
        self.lib = _lib.load()
        self.precision = _lib.F64 if precision == "f64" else _lib.F32
        self.dtype = torch.float64 if self.precision == _lib.F64 else torch.float32
        xos, xsd = _f64(geom.xos), _f64(geom.xsd)
        self.handle = self.lib.trtt_hip_dd_create(self.precision, geom.nx, geom.ny, geom.nz,
                                                  geom.s_scl, geom.nu, geom.nv, geom.u_scl,
                                                  geom.v_scl, _ptr(xos), _ptr(xsd), geom.nviews)
        if not self.handle:
            raise RuntimeError("distance-driven plan creation failed: " + _lib.last_error())

    @property
    def vol_shape(self):
        g = self.geom
        return (g.nz, g.ny, g.nx)

    @property
    def sino_shape(self):
        g = self.geom
        return (g.nviews, g.nv, g.nu)

    def _stream(self, stream):
        s = stream if stream is not None else self.torch.cuda.current_stream()
        return C.c_void_p(s.cuda_stream)

    def forward(self, vol, sino=None, stream=None):
        if sino is None:
            sino = self.torch.zeros(self.sino_shape, dtype=self.dtype, device=vol.device)
        if self.lib.trtt_hip_dd_forward(self.handle, C.c_void_p(vol.data_ptr()),
                                        C.c_void_p(sino.data_ptr()), self._stream(stream)) != 0:
            raise RuntimeError(_lib.last_error())
        return sino

    def backward(self, sino, vol=None, exact: bool = True, stream=None):
        if vol is None:
            vol = self.torch.zeros(self.vol_shape, dtype=self.dtype, device=sino.device)
        if self.lib.trtt_hip_dd_backward(self.handle, C.c_void_p(sino.data_ptr()),
                                         C.c_void_p(vol.data_ptr()), int(exact),
                                         self._stream(stream)) != 0:
            raise RuntimeError(_lib.last_error())
        return vol

    def close(self):
        if getattr(self, "handle", None):
            self.lib.trtt_hip_dd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
