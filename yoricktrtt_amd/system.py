"""Whole tomographic systems: build, save and load.

A system is the reference's ``Htomo`` table (Tomography_4D/Ytrtt4D_system.i:
190-237): the tomographic object X (sizes, scales, B-spline degrees,
footprint table, voxel->metric transform, optionally its coefficients), the
data set Y (one TomData per view: detector sizes and scales, time index,
angle, distances, the xi/xotilt/xos/xsd transforms, the measured view and its
weights), an optional reference object Xref, and the projection mode from
which the projector set Cops is built.

* :func:`create_simu_system_4d` follows trtt_4D_create_whole_simu_system,
  trtt_4D_create_simu_tomdata_set and trtt_4D_create_tomobj
  (Ytrtt4D_system.i:190-237, :400-547, :549-591): view times
  t = k tacq / (ndata - 1), time index (t mod tcycl) nt / tcycl, angles
  k pi / ndata unless given, the xos/xsd/xotilt moves of each view.
* :func:`save_system` / :func:`load_system` take the place of
  trtt_4D_save_simu_system / trtt_4D_load_simu_system (:241-399), which
  store the hash table with Yeti's yhd_save after deleting what cannot be
  serialised (the XForm3D objects and spline interpolators), keeping their
  ``*_coefs``, and rebuild them on load.  Yeti's YHD container is not
  available here: the same fields go into one numpy ``.npz`` archive of
  plain arrays (no pickled objects; loaded with allow_pickle=False), per-view
  fields stacked along a leading view axis.  Loading rebuilds every XForm3D
  from its coefficients, and the projector set is built from the loaded
  fields on request (:meth:`TomoSystem.projectors`), as the reference
  re-hangs its evaluators and rebuilds the time interpolation from t_index.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

from .geometry import Geometry, XForm3D, projector_coefs, view_transforms_3d
from .projectors import (TRTT_CONE_BEAM, TRTT_FAN_BEAM, TRTT_PARALLEL_BEAM, TomData, TomObj,
                         TrttError)

FORMAT = "trtt-system-1"

# Error codes (Ytrtt_error.i:47-62)
TRTT_ERR_BAD_NDIMS = 2
TRTT_ERR_BAD_GEOMETRY = 13
TRTT_ERR_MISSING_PARAMETER = 15


@dataclass
class TomoSystem:
    """X, Y, Xref and the projection mode of one reconstruction problem.

    ``data`` is [nviews][nv][nu] (the view k of Y[k]), ``weights`` the same
    shape or None, ``x`` the object coefficients [nt][nz][ny][nx] or None,
    ``xref`` the reference object samples of the same shape or None."""

    X: TomObj
    Y: list
    mode: int
    data: np.ndarray
    t_deg: int = 1
    t_scl: float = 1.0
    weights: np.ndarray | None = None
    x: np.ndarray | None = None
    xref: np.ndarray | None = None

    @property
    def nviews(self) -> int:
        return len(self.Y)

    def geometry(self) -> Geometry:
        """Per-view XOS/XSD coefficients from the transforms of X and Y
        (Ytrtt_3D_projectors.i:104-110)."""
        X, Y = self.X, self.Y
        if not Y:
            raise ValueError("empty data set")
        nu, nv, us, vs = Y[0].nu, Y[0].nv, Y[0].u_scl, Y[0].v_scl
        for Yk in Y:
            if (Yk.nu, Yk.nv, Yk.u_scl, Yk.v_scl) != (nu, nv, us, vs):
                raise ValueError("the views of a system share one detector")
        coefs = [projector_coefs(X.xi, Yk.xi, Yk.xos, Yk.xotilt, Yk.xsd) for Yk in Y]
        xos = np.stack([c[0] for c in coefs])
        xsd = np.stack([c[1] for c in coefs])
        kind = "parallel3d" if self.mode == TRTT_PARALLEL_BEAM else "cone"
        return Geometry(kind, X.nx, X.ny, X.nz, nu, nv, X.s_scl, us, vs, xos, xsd, X.s_deg,
                        np.array([Yk.theta for Yk in Y]))

    def t_index(self) -> np.ndarray:
        return np.array([Yk.t_index for Yk in self.Y], dtype=np.float64)

    def projectors(self, model: str = "cubic", precision: str = "f32"):
        """The projector set C_k = A_k B_k over every view on the GPU: a
        :class:`~yoricktrtt_amd.projectors.Projector` for one phase, else a
        :class:`~yoricktrtt_amd.projectors.DynamicProjector` whose time
        interpolation is rebuilt from the views' t_index and t_deg."""
        from .projectors import DynamicProjector, Projector

        geom = self.geometry()
        X = self.X
        table = None
        if model == "table":
            table = (X.cum_footprint, X.coord_footprint, X.step_footprint)
        if X.nt == 1:
            return Projector(geom, model=model, precision=precision, table=table)
        if model != "cubic":
            raise ValueError("a multi-phase projector set uses the closed-form cubic model")
        return DynamicProjector(geom, X.nt, self.t_index(), t_deg=self.t_deg, model=model,
                                precision=precision)


def create_tomobj_4d(nx, ny, nz, nt, s_scl, tcycl, s_deg=3, t_deg=1, size_footprint=10000,
                     voxels_ref=None):
    """X of a 4-D system (trtt_4D_create_tomobj, Ytrtt4D_system.i:549-591):
    t_scl = tcycl / nt; returns (X, t_scl, t_deg, xref)."""
    X = TomObj(int(nx), int(ny), int(nz), int(nt), float(s_scl), int(s_deg), int(size_footprint))
    xref = None
    if voxels_ref is not None:
        xref = np.asarray(voxels_ref, dtype=np.float64)
        if xref.shape != (nt, nz, ny, nx):
            raise TrttError(TRTT_ERR_BAD_NDIMS, "voxels_ref must be [nt][nz][ny][nx]")
    return X, float(tcycl) / nt, int(t_deg), xref


def create_tomdata_set_4d(data, nu, nv, u_scl, v_scl, ndata, nt, tcycl, tacq, mode, dates=None,
                          x_off=None, y_off=None, z_off=None, u_off=None, v_off=None,
                          theta=None, Rsc=None, Rsd=None):
    """Y of a 4-D system (trtt_4D_create_simu_tomdata_set,
    Ytrtt4D_system.i:400-547): view times from tacq or ``dates``, time
    index (t mod tcycl) nt / tcycl, angles k pi / ndata unless given, and
    the per-view transform moves (xos: rotation by (pi - theta)/2 about z
    and Rsc along x; xsd: -Rsd along x, then +v_off along y and +u_off
    along z, :497-510; xotilt: the x/y/z offsets).  ``data`` is
    [ndata][nv][nu]."""
    if mode not in (TRTT_PARALLEL_BEAM, TRTT_FAN_BEAM, TRTT_CONE_BEAM):
        raise TrttError(TRTT_ERR_BAD_GEOMETRY)
    if mode != TRTT_PARALLEL_BEAM and (Rsc is None or Rsd is None):
        raise TrttError(TRTT_ERR_MISSING_PARAMETER, "You must specify the source/center and "
                        "source/detector distances in fan beam mode.")
    ndata = int(ndata)
    if dates is None:
        t = np.arange(ndata) * (tacq / (ndata - 1.0))
    else:
        t = np.asarray(dates, dtype=np.float64)
        if t.size != ndata:
            raise TrttError(TRTT_ERR_BAD_NDIMS, "There must be as DATES as the number of "
                            "projections NDATA.")
    t_index = np.fmod(t, tcycl) * (nt / tcycl)
    if theta is None:
        theta = np.arange(ndata) * (math.pi / ndata)
    theta = np.asarray(theta, dtype=np.float64)
    for name, off in (("V_OFF", v_off), ("U_OFF", u_off)):
        if off is not None and np.size(off) not in (1, ndata):
            raise TrttError(TRTT_ERR_BAD_NDIMS, f"Offset parameter {name} has bad number of "
                            "elements.")

    def pick(v, k):
        if v is None:
            return 0.0
        v = np.asarray(v, dtype=np.float64).reshape(-1)
        return float(v[0] if v.size == 1 else v[k])

    data = np.asarray(data, dtype=np.float64)
    if data.shape != (ndata, nv, nu):
        raise TrttError(TRTT_ERR_BAD_NDIMS, "data must be [ndata][nv][nu]")
    par = mode == TRTT_PARALLEL_BEAM
    Y = []
    for k in range(ndata):
        # the 4-D builder moves xsd by +v_off along y and +u_off along z,
        # the 3-D one (view_transforms_3d) by -v_off and -u_off
        xos, xotilt, xsd = view_transforms_3d(
            float(theta[k]), 0.0 if par else float(Rsc), 0.0 if par else float(Rsd),
            -pick(u_off, k), -pick(v_off, k), pick(x_off, k), pick(y_off, k), pick(z_off, k))
        Y.append(TomData(int(nu), int(nv), float(u_scl), float(v_scl), float(t_index[k]), xos,
                         xotilt, xsd, float(theta[k]), None if par else float(Rsc),
                         None if par else float(Rsd), pick(v_off, k), pick(x_off, k),
                         pick(y_off, k)))
    return Y


def create_simu_system_4d(data, nx, ny, nz, nt, s_scl, nu, nv, u_scl, v_scl, ndata, tcycl, tacq,
                          mode, s_deg=3, t_deg=1, size_footprint=10000, weights=None,
                          voxels_ref=None, **view_kw) -> TomoSystem:
    """A whole 4-D system (trtt_4D_create_whole_simu_system,
    Ytrtt4D_system.i:190-237) from the data and the acquisition parameters;
    ``view_kw`` are the keywords of :func:`create_tomdata_set_4d`."""
    X, t_scl, t_deg, xref = create_tomobj_4d(nx, ny, nz, nt, s_scl, tcycl, s_deg, t_deg,
                                             size_footprint, voxels_ref)
    Y = create_tomdata_set_4d(data, nu, nv, u_scl, v_scl, ndata, nt, tcycl, tacq, mode, **view_kw)
    w = None if weights is None else np.asarray(weights, dtype=np.float64)
    if w is not None and w.shape != (ndata, nv, nu):
        raise TrttError(TRTT_ERR_BAD_NDIMS, "weights must be [ndata][nv][nu]")
    return TomoSystem(X, Y, mode, np.asarray(data, dtype=np.float64), t_deg, t_scl, w, None, xref)


def _opt(v):
    return np.nan if v is None else float(v)


def save_system(path, system: TomoSystem, overwrite: bool = False) -> None:
    """Write ``system`` to ``path`` (a ``.npz`` archive of plain arrays).

    As with trtt_4D_save_simu_system (Ytrtt4D_system.i:357-399), an
    existing file is an error unless ``overwrite`` is set; the transforms
    go in as their 12 coefficients (X.xi -> X_xi_coefs; per view
    xi/xotilt/xos/xsd -> Y_*_coefs), and nothing that cannot be rebuilt is
    lost.  The archive is written to a temporary name and renamed, so a
    crash leaves the previous file whole."""
    path = os.fspath(path)
    if not path.endswith(".npz"):
        path += ".npz"
    if os.path.exists(path) and not overwrite:
        raise FileExistsError(f"{path} exists (give overwrite=True to replace it)")
    X, Y = system.X, system.Y
    a = dict(
        format=np.array(FORMAT),
        mode=np.int64(system.mode), t_deg=np.int64(system.t_deg), t_scl=np.float64(system.t_scl),
        X_dims=np.array([X.nx, X.ny, X.nz, X.nt], dtype=np.int64),
        X_s_scl=np.float64(X.s_scl), X_s_deg=np.int64(X.s_deg),
        X_size_footprint=np.int64(X.size_footprint), X_footprint=X.footprint,
        X_cum_footprint=X.cum_footprint, X_coord_footprint=X.coord_footprint,
        X_step_footprint=np.float64(X.step_footprint), X_xi_coefs=X.xi.a,
        Y_det=np.array([[Yk.nu, Yk.nv] for Yk in Y], dtype=np.int64).reshape(-1, 2),
        Y_scl=np.array([[Yk.u_scl, Yk.v_scl] for Yk in Y], dtype=np.float64).reshape(-1, 2),
        Y_t_index=np.array([Yk.t_index for Yk in Y], dtype=np.float64),
        Y_theta=np.array([Yk.theta for Yk in Y], dtype=np.float64),
        Y_dist=np.array([[_opt(Yk.Rsc), _opt(Yk.Rsd)] for Yk in Y], dtype=np.float64).reshape(-1, 2),
        Y_off=np.array([[Yk.voff, Yk.xoff, Yk.yoff] for Yk in Y], dtype=np.float64).reshape(-1, 3),
        Y_xi_coefs=np.array([Yk.xi.a for Yk in Y]).reshape(-1, 12),
        Y_xotilt_coefs=np.array([Yk.xotilt.a for Yk in Y]).reshape(-1, 12),
        Y_xos_coefs=np.array([Yk.xos.a for Yk in Y]).reshape(-1, 12),
        Y_xsd_coefs=np.array([Yk.xsd.a for Yk in Y]).reshape(-1, 12),
        data=np.asarray(system.data),
    )
    for name in ("weights", "x", "xref"):
        v = getattr(system, name)
        if v is not None:
            a[name] = np.asarray(v)
    tmp = path[:-4] + ".tmp.npz"
    np.savez(tmp, **a)
    os.replace(tmp, path)


def load_system(path) -> TomoSystem:
    """Read a system written by :func:`save_system` and rebuild its
    transforms from their coefficients (trtt_4D_load_simu_system,
    Ytrtt4D_system.i:241-355).  The detector transform of a view is a
    function of its sizes and scales; the stored coefficients must agree
    with it."""
    path = os.fspath(path)
    with np.load(path, allow_pickle=False) as z:
        a = {k: z[k] for k in z.files}
    if str(a.get("format")) != FORMAT:
        raise ValueError(f"{path}: not a {FORMAT} archive")
    nx, ny, nz, nt = (int(v) for v in a["X_dims"])
    X = TomObj(nx, ny, nz, nt, float(a["X_s_scl"]), int(a["X_s_deg"]),
               int(a["X_size_footprint"]))
    X.footprint = a["X_footprint"]
    X.cum_footprint = a["X_cum_footprint"]
    X.coord_footprint = a["X_coord_footprint"]
    X.step_footprint = float(a["X_step_footprint"])
    X.xi = XForm3D(a["X_xi_coefs"])
    Y = []
    for k in range(a["Y_t_index"].size):
        Rsc, Rsd = (None if math.isnan(v) else float(v) for v in a["Y_dist"][k])
        voff, xoff, yoff = (float(v) for v in a["Y_off"][k])
        Yk = TomData(int(a["Y_det"][k, 0]), int(a["Y_det"][k, 1]), float(a["Y_scl"][k, 0]),
                     float(a["Y_scl"][k, 1]), float(a["Y_t_index"][k]),
                     XForm3D(a["Y_xos_coefs"][k]), XForm3D(a["Y_xotilt_coefs"][k]),
                     XForm3D(a["Y_xsd_coefs"][k]), float(a["Y_theta"][k]), Rsc, Rsd, voff, xoff,
                     yoff)
        if not np.array_equal(Yk.xi.a, a["Y_xi_coefs"][k]):
            raise ValueError(f"{path}: view {k}: detector transform does not match its sizes")
        Y.append(Yk)
    return TomoSystem(X, Y, int(a["mode"]), a["data"], int(a["t_deg"]), float(a["t_scl"]),
                      a.get("weights"), a.get("x"), a.get("xref"))
