"""Host-side FP64 geometry: 3-D affine transforms, object/detector frames and
the per-view XOS/XSD coefficient pairs the projectors consume.

The transform is a 3x4 affine map stored as 12 row-major coefficients
``[a0 a1 a2 a3; a4 a5 a6 a7; a8 a9 a10 a11]`` acting as
``x' = a0 x + a1 y + a2 z + a3`` etc. (plugin/trtt_xform3d.c:52, row-major
build; plugin/trtt_xform3d.h).  Operations follow the reference semantics:

* ``move``    adds a translation after the map       (trtt_xform3d.c:307-331)
* ``scale``   scales rows x, y, z after the map      (trtt_xform3d.c:390-425)
* ``rotate``  left-multiplies by a quaternion rotation (trtt_xform3d.c:333-388)
* ``combine`` composes ``left o right``              (trtt_xform3d.c:286-305)
* ``invert``  inverts the affine map                 (trtt_xform3d.c:438-459)

Frames (Yorick layer):

* object ``xi``: voxel index -> metric, centred (Ytrtt_tomobj.i:198-236)
* detector ``xi``: (w, v, u) pixel-bound index -> metric (Ytrtt_tomdata.i:140-185)
* per view ``xos`` (object->source), ``xotilt`` (object shift), ``xsd``
  (source->detector) built as in Tomography_3D/Ytrtt3D_system.i:436-480 and
  Tomography_2D/Ytrtt2D_system.i:529-589
* projector coefficients ``XOS = xos . xotilt . obj_xi`` and
  ``XSD = data_xi^-1 . xsd`` (Ytrtt_3D_projectors.i:104-110).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "XForm3D",
    "object_xi",
    "detector_xi",
    "view_transforms_3d",
    "view_transforms_2d",
    "projector_coefs",
    "cone_geometry",
    "parallel_geometry_3d",
    "fan_geometry_2d",
    "parallel_geometry_2d",
    "footprint_table",
    "bspline",
    "Geometry",
]


class XForm3D:
    """3x4 affine transform with the reference XForm3D operations (FP64)."""

    __slots__ = ("a",)

    def __init__(self, coefs=None):
        if coefs is None:
            self.a = np.array([1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0])
        else:
            a = np.asarray(coefs, dtype=np.float64).reshape(-1)
            if a.size != 12:
                raise ValueError("an XForm3D has 12 coefficients")
            self.a = a.copy()

    def copy(self) -> "XForm3D":
        return XForm3D(self.a)

    def reset(self) -> "XForm3D":
        self.a = XForm3D().a
        return self

    def move(self, tx: float, ty: float, tz: float) -> "XForm3D":
        """Translate after the map (in place, returns self)."""
        self.a[3] += tx
        self.a[7] += ty
        self.a[11] += tz
        return self

    def scale(self, sx: float, sy: float, sz: float) -> "XForm3D":
        """Scale the output axes (rows) after the map (in place)."""
        self.a[0:4] *= sx
        self.a[4:8] *= sy
        self.a[8:12] *= sz
        return self

    @staticmethod
    def _rotation(qw, qx, qy, qz):
        ww, xx, yy, zz = qw * qw, qx * qx, qy * qy, qz * qz
        n2 = (ww + xx) + (yy + zz)
        if not n2 > 0.0:
            raise ValueError("quaternion cannot be normalised")
        if n2 == 1.0:
            d = np.array([(ww + xx) - (yy + zz), (ww + yy) - (xx + zz), (ww + zz) - (xx + yy)])
            s = 2.0
        else:
            inv = 1.0 / n2
            d = inv * np.array([(ww + xx) - (yy + zz), (ww + yy) - (xx + zz), (ww + zz) - (xx + yy)])
            s = inv + inv
        wx, wy, wz = s * qw * qx, s * qw * qy, s * qw * qz
        xy, xz, yz = s * qx * qy, s * qx * qz, s * qy * qz
        return np.array([
            [d[0], xy - wz, xz + wy],
            [xy + wz, d[1], yz - wx],
            [xz - wy, yz + wx, d[2]],
        ])

    def rotate(self, qw: float, qx: float, qy: float, qz: float) -> "XForm3D":
        """Left-multiply by the rotation of quaternion (qw, qx, qy, qz)."""
        r = self._rotation(qw, qx, qy, qz)
        m = self.a.reshape(3, 4)
        self.a = (r @ m).reshape(-1)
        return self

    def matrix(self) -> np.ndarray:
        m = np.eye(4)
        m[:3, :] = self.a.reshape(3, 4)
        return m

    @staticmethod
    def combine(left: "XForm3D", right: "XForm3D") -> "XForm3D":
        """Return ``left o right`` (apply right first)."""
        m = left.matrix() @ right.matrix()
        return XForm3D(m[:3, :].reshape(-1))

    def invert(self) -> "XForm3D":
        """Return the inverse affine map (raises on a singular linear part)."""
        lin = self.a.reshape(3, 4)[:, :3]
        t = self.a.reshape(3, 4)[:, 3]
        if abs(np.linalg.det(lin)) == 0.0:
            raise np.linalg.LinAlgError("singular XForm3D")
        inv = np.linalg.inv(lin)
        out = np.empty((3, 4))
        out[:, :3] = inv
        out[:, 3] = -inv @ t
        return XForm3D(out.reshape(-1))

    def apply(self, p) -> np.ndarray:
        p = np.asarray(p, dtype=np.float64)
        return self.a.reshape(3, 4)[:, :3] @ p + self.a.reshape(3, 4)[:, 3]

    def __repr__(self) -> str:
        return "XForm3D(" + np.array2string(self.a.reshape(3, 4), precision=6) + ")"


def object_xi(nx: int, ny: int, nz: int, s_scl: float) -> XForm3D:
    """Voxel index -> metric, origin at the object centre (Ytrtt_tomobj.i:214-232)."""
    return XForm3D().move(-0.5 * (nx - 1), -0.5 * (ny - 1), -0.5 * (nz - 1)).scale(s_scl, s_scl, s_scl)


def detector_xi(nu: int, nv: int, u_scl: float, v_scl: float) -> XForm3D:
    """Pixel-bound index -> metric on the detector plane (Ytrtt_tomdata.i:165-179)."""
    return XForm3D().move(0.0, -0.5 * nv, -0.5 * nu).scale(1.0, v_scl, u_scl)


def view_transforms_3d(theta, Rsc, Rsd, u_off=0.0, v_off=0.0, x_off=0.0, y_off=0.0, z_off=0.0):
    """xos, xotilt, xsd of one 3-D view (Tomography_3D/Ytrtt3D_system.i:436-480)."""
    th = 0.5 * (math.pi - theta)
    xos = XForm3D().rotate(math.cos(th), 0.0, 0.0, math.sin(th)).move(Rsc, 0.0, 0.0)
    xsd = XForm3D().move(-1.0 * Rsd, 0.0, 0.0)
    if v_off:
        xsd.move(0.0, -1.0 * v_off, 0.0)
    if u_off:
        xsd.move(0.0, 0.0, -1.0 * u_off)
    xotilt = XForm3D()
    if x_off:
        xotilt.move(x_off, 0.0, 0.0)
    if y_off:
        xotilt.move(0.0, y_off, 0.0)
    if z_off:
        xotilt.move(0.0, 0.0, z_off)
    return xos, xotilt, xsd


def view_transforms_2d(theta, Rsc=None, Rsd=None, voff=0.0, xoff=0.0, yoff=0.0):
    """xos, xotilt, xsd of one 2-D view (Tomography_2D/Ytrtt2D_system.i:529-589).

    Unlike 3-D, the 2-D builder moves ``xsd`` by ``+voff``.
    """
    th = 0.5 * (math.pi - theta)
    xos = XForm3D().rotate(math.cos(th), 0.0, 0.0, math.sin(th))
    if Rsc is not None:
        xos.move(Rsc, 0.0, 0.0)
    xsd = XForm3D()
    if Rsd is not None:
        xsd.move(-1.0 * Rsd, 0.0, 0.0)
    if voff:
        xsd.move(0.0, voff, 0.0)
    xotilt = XForm3D()
    if xoff:
        xotilt.move(xoff, 0.0, 0.0)
    if yoff:
        xotilt.move(0.0, yoff, 0.0)
    return xos, xotilt, xsd


def projector_coefs(obj_xi: XForm3D, data_xi: XForm3D, xos: XForm3D, xotilt: XForm3D, xsd: XForm3D):
    """(XOS, XSD) coefficient pair of one view (Ytrtt_3D_projectors.i:104-110)."""
    xos_plus = XForm3D.combine(xos, XForm3D.combine(xotilt, obj_xi))
    xsd_plus = XForm3D.combine(data_xi.invert(), xsd)
    return xos_plus.a.copy(), xsd_plus.a.copy()


def default_angles(nviews: int, span: float = math.pi) -> np.ndarray:
    """Evenly spaced view angles ``k*span/nviews`` (Ytrtt3D_system.i:398-402)."""
    return np.arange(nviews, dtype=np.float64) * (span / nviews)


@dataclass
class Geometry:
    """Sizes and the per-view coefficient tables of one acquisition."""

    kind: str                 # "cone", "parallel3d", "fan", "parallel2d"
    nx: int
    ny: int
    nz: int
    nu: int
    nv: int
    s_scl: float
    u_scl: float
    v_scl: float
    xos: np.ndarray           # (nviews, 12)
    xsd: np.ndarray           # (nviews, 12)
    s_deg: int = 3
    theta: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def nviews(self) -> int:
        return int(self.xos.shape[0])

    @property
    def is3d(self) -> bool:
        return self.kind in ("cone", "parallel3d")

    def subset(self, views) -> "Geometry":
        views = np.asarray(views)
        return Geometry(self.kind, self.nx, self.ny, self.nz, self.nu, self.nv, self.s_scl,
                        self.u_scl, self.v_scl, self.xos[views].copy(), self.xsd[views].copy(),
                        self.s_deg, self.theta[views].copy() if self.theta.size else self.theta)


def _stack(per_view):
    xos = np.stack([p[0] for p in per_view])
    xsd = np.stack([p[1] for p in per_view])
    return xos, xsd


def _per_view(v, k):
    if v is None:
        return 0.0
    v = np.asarray(v, dtype=np.float64)
    return float(v) if v.ndim == 0 else float(v[k])


def cone_geometry(nx, ny, nz, s_scl, nu, nv, u_scl, v_scl, theta, Rsc, Rsd,
                  u_off=None, v_off=None, x_off=None, y_off=None, z_off=None, s_deg=3) -> Geometry:
    """Circular cone-beam orbit with a flat detector (Ytrtt3D_system.i:368-511)."""
    theta = np.asarray(theta, dtype=np.float64)
    oxi = object_xi(nx, ny, nz, s_scl)
    dxi = detector_xi(nu, nv, u_scl, v_scl)
    pv = []
    for k, th in enumerate(theta):
        xos, xot, xsd = view_transforms_3d(th, Rsc, Rsd, _per_view(u_off, k), _per_view(v_off, k),
                                           _per_view(x_off, k), _per_view(y_off, k), _per_view(z_off, k))
        pv.append(projector_coefs(oxi, dxi, xos, xot, xsd))
    xos, xsd = _stack(pv)
    return Geometry("cone", nx, ny, nz, nu, nv, s_scl, u_scl, v_scl, xos, xsd, s_deg, theta)


def parallel_geometry_3d(nx, ny, nz, s_scl, nu, nv, u_scl, v_scl, theta, u_off=None, v_off=None,
                         x_off=None, y_off=None, z_off=None, s_deg=3) -> Geometry:
    """Parallel-beam 3-D geometry: the same frames without source distances."""
    theta = np.asarray(theta, dtype=np.float64)
    oxi = object_xi(nx, ny, nz, s_scl)
    dxi = detector_xi(nu, nv, u_scl, v_scl)
    pv = []
    for k, th in enumerate(theta):
        xos, xot, xsd = view_transforms_3d(th, 0.0, 0.0, _per_view(u_off, k), _per_view(v_off, k),
                                           _per_view(x_off, k), _per_view(y_off, k), _per_view(z_off, k))
        pv.append(projector_coefs(oxi, dxi, xos, xot, xsd))
    xos, xsd = _stack(pv)
    return Geometry("parallel3d", nx, ny, nz, nu, nv, s_scl, u_scl, v_scl, xos, xsd, s_deg, theta)


def fan_geometry_2d(nx, ny, s_scl, nv, v_scl, theta, Rsc, Rsd, voff=None, xoff=None, yoff=None,
                    s_deg=3) -> Geometry:
    """2-D fan beam, flat detector (Tomography_2D/Ytrtt2D_system.i:454-634)."""
    theta = np.asarray(theta, dtype=np.float64)
    oxi = object_xi(nx, ny, 1, s_scl)
    dxi = detector_xi(1, nv, 1.0, v_scl)
    pv = []
    for k, th in enumerate(theta):
        xos, xot, xsd = view_transforms_2d(th, Rsc, Rsd, _per_view(voff, k), _per_view(xoff, k),
                                           _per_view(yoff, k))
        pv.append(projector_coefs(oxi, dxi, xos, xot, xsd))
    xos, xsd = _stack(pv)
    return Geometry("fan", nx, ny, 1, 1, nv, s_scl, 1.0, v_scl, xos, xsd, s_deg, theta)


def parallel_geometry_2d(nx, ny, s_scl, nv, v_scl, theta, voff=None, xoff=None, yoff=None,
                         s_deg=3) -> Geometry:
    """2-D parallel beam (Tomography_2D/Ytrtt2D_system.i:454-634 without Rsc/Rsd)."""
    theta = np.asarray(theta, dtype=np.float64)
    oxi = object_xi(nx, ny, 1, s_scl)
    dxi = detector_xi(1, nv, 1.0, v_scl)
    pv = []
    for k, th in enumerate(theta):
        xos, xot, xsd = view_transforms_2d(th, None, None, _per_view(voff, k), _per_view(xoff, k),
                                           _per_view(yoff, k))
        pv.append(projector_coefs(oxi, dxi, xos, xot, xsd))
    xos, xsd = _stack(pv)
    return Geometry("parallel2d", nx, ny, 1, 1, nv, s_scl, 1.0, v_scl, xos, xsd, s_deg, theta)


def bspline(x, deg: int = 3) -> np.ndarray:
    """Centred B-spline of degree 1..3 evaluated at x (vectorised)."""
    t = np.abs(np.asarray(x, dtype=np.float64))
    if deg == 3:
        out = np.where(t < 1.0, 2.0 / 3.0 - t * t + 0.5 * t ** 3, 0.0)
        return np.where((t >= 1.0) & (t < 2.0), (2.0 - t) ** 3 / 6.0, out)
    if deg == 2:
        out = np.where(t < 0.5, 0.75 - t * t, 0.0)
        return np.where((t >= 0.5) & (t < 1.5), 0.5 * (1.5 - t) ** 2, out)
    if deg == 1:
        return np.where(t < 1.0, 1.0 - t, 0.0)
    raise ValueError("footprint degree must be 1, 2 or 3")


def footprint_table(m: int, deg: int = 3):
    """Footprint samples, trapezoid cumulative sum, coordinates and step.

    ``m`` samples of the centred B-spline on ``[-(n-1)/2, (n-1)/2]`` with
    ``n = 2*deg+3`` (Ytrtt_tomobj.i:140-181, :296-337); cumulative sum
    ``cum[i] = cum[i-1] + step/2 (f[i-1] + f[i])`` (trtt_common.c:134-142).
    """
    n = 2 * deg + 3
    step = float(n - 1) / float(m - 1)
    coord = np.arange(m, dtype=np.float64) * step - 0.5 * (n - 1)
    fp = bspline(coord, deg)
    cum = np.empty(m)
    cum[0] = 0.0
    incr = 0.5 * step * (fp[:-1] + fp[1:])
    # sequential sum in the reference's left-to-right order
    acc = 0.0
    for i in range(1, m):
        acc = acc + incr[i - 1]
        cum[i] = acc
    return fp, cum, coord, step


def cone_rays(geom: Geometry, view: int, uv):
    """Source and ray directions of a cone-beam view in voxel index
    coordinates (i, j, k), derived from the view's own projective
    coefficients (those of trtt_CB3D, trtt_3D_projectors.c:1190-1261): a point
    p projects to uc = c3 - a4 num_u(p)/den(p), vc = b3 - a4 num_v(p)/den(p)
    with num_u, num_v and den affine in p.  The source is their common zero;
    the ray through detector point (u, v) (pixel-bound coordinates) is
    orthogonal to the linear parts of num_u - a den and num_v - b den,
    a = (u - c3)/(-a4), b = (v - b3)/(-a4).  ``uv`` is an (n, 2) array of
    detector points; returns (source (3,), unit directions (n, 3))."""
    xos = np.asarray(geom.xos[view], dtype=np.float64)
    xsd = np.asarray(geom.xsd[view], dtype=np.float64)
    a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12 = xsd
    a1inv = -1.0 / a1
    b1, b2, b3 = a6 + a1inv * a5 * a2, a7 + a1inv * a5 * a3, a8 + a1inv * a5 * a4
    c1, c2, c3 = a10 + a1inv * a9 * a2, a11 + a1inv * a9 * a3, a12 + a1inv * a9 * a4
    o = xos.reshape(3, 4)
    lv = b1 * o[1, :3] + b2 * o[2, :3]
    lu = c1 * o[1, :3] + c2 * o[2, :3]
    ld = a1 * o[0, :3] + a2 * o[1, :3] + a3 * o[2, :3]
    cv = b1 * o[1, 3] + b2 * o[2, 3]
    cu = c1 * o[1, 3] + c2 * o[2, 3]
    cd = a1 * o[0, 3] + a2 * o[1, 3] + a3 * o[2, 3]
    src = np.linalg.solve(np.stack([lu, lv, ld]), -np.array([cu, cv, cd]))
    uv = np.asarray(uv, dtype=np.float64)
    al = (uv[:, 0] - c3) / (-a4)
    be = (uv[:, 1] - b3) / (-a4)
    d = np.cross(lu[None, :] - al[:, None] * ld[None, :], lv[None, :] - be[:, None] * ld[None, :])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return src, d
