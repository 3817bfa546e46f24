"""Seeded synthetic objects and measurement noise.

* :func:`ellipses_2d` -- additive random ellipses on a square image, the
  phantoms of the 2-D scripts (Tomography_2D/Ytrtt2D_simulated_data_generation.i
  builds its objects as sums of ellipses);
* :func:`ellipsoids_3d` -- additive random ellipsoids with random 3-D
  orientations (Tomography_3D/Ytrtt3D_simulated_data_generation.i:46-155
  builds its objects the same way);
* :func:`body_phantom` -- a body ellipsoid with inner structures, the object
  class of the learned warm start;
* :func:`transmission_noise` -- Poisson counts of a transmission measurement
  and the inverse-variance weights of the reference noise model
  (Tomography_4D/Ytrtt4D_system.i:829-843).

Every draw comes from ``numpy.random.default_rng(seed)``; arrays are indexed
[z][y][x] (3-D) or [y][x] (2-D) with the grid centred on the origin, one unit
per voxel.
"""
from __future__ import annotations

import numpy as np


def _rotation(rng):
    """Uniform random rotation (unit quaternion)."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def add_ellipsoid(vol, centre, axes, rot, value):
    """vol += value inside {p : |R' (p - c) / a| <= 1} (p = (x, y, z)),
    evaluated over the bounding box of the ellipsoid only."""
    nz, ny, nx = vol.shape
    half = np.array([(nx - 1) / 2, (ny - 1) / 2, (nz - 1) / 2])
    ext = np.sqrt((np.asarray(rot) ** 2) @ (np.asarray(axes, dtype=np.float64) ** 2))
    lo = np.maximum(np.floor(np.asarray(centre) - ext + half).astype(int), 0)
    hi = np.minimum(np.ceil(np.asarray(centre) + ext + half).astype(int) + 1, [nx, ny, nz])
    if np.any(hi <= lo):
        return vol
    z, y, x = np.meshgrid(*(np.arange(lo[k], hi[k]) - half[k] for k in (2, 1, 0)), indexing="ij")
    d = np.stack([x - centre[0], y - centre[1], z - centre[2]], axis=-1) @ rot
    q = (d[..., 0] / axes[0]) ** 2 + (d[..., 1] / axes[1]) ** 2 + (d[..., 2] / axes[2]) ** 2
    vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += np.where(q <= 1.0, value, 0.0).astype(vol.dtype)
    return vol


def ellipsoids_3d(shape, n_ell, seed, centre_frac=0.45, axes_frac=(0.06, 0.4), dtype=np.float32,
                  params=False):
    """n_ell additive ellipsoids: centres uniform in the central box of half
    width centre_frac * n/2, semi-axes in axes_frac * n/2, values in [0, 1].
    With ``params`` also the list of (centre, axes, rotation, value) in grid
    units centred on the origin (for :func:`cone_projection_ellipsoids`)."""
    rng = np.random.default_rng(seed)
    vol = np.zeros(shape, dtype=dtype)
    half = np.array(shape[::-1], dtype=np.float64) / 2.0  # (x, y, z)
    ells = []
    for _ in range(n_ell):
        c = rng.uniform(-centre_frac, centre_frac, size=3) * half
        a = rng.uniform(axes_frac[0], axes_frac[1], size=3) * half.min()
        rot, val = _rotation(rng), rng.uniform(0.0, 1.0)
        add_ellipsoid(vol, c, a, rot, val)
        ells.append((c, a, rot, val))
    return (vol, ells) if params else vol


def cone_projection_ellipsoids(geom, ells, views=None, device="cuda", quad=2):
    """Analytic cone-beam data of a sum of ellipsoids: for every detector
    pixel the line integral of the object along the ray from the source,
    averaged over the pixel (quad x quad Gauss-Legendre points), scaled to
    the footprint projector's units (s_scl u_scl v_scl times the integral in
    voxel lengths, the measured normalisation of trtt_CB3D).  The chord of a
    ray through an ellipsoid is that of Tomography_3D/Simul_Data/simul.c:
    49-111 (a quadratic in the ray parameter after the ellipsoid's own
    frame).  ``ells`` holds (centre, axes, rotation, value) rows in grid units
    centred on the origin, or one such list per view (moving objects).
    Returns [views][nv][nu] float32 on ``device``."""
    import torch

    from .geometry import cone_rays

    views = range(geom.nviews) if views is None else views
    half = np.array([(geom.nx - 1) / 2, (geom.ny - 1) / 2, (geom.nz - 1) / 2])
    g, w = np.polynomial.legendre.leggauss(quad)
    off = 0.5 + 0.5 * g  # points in the pixel [0, 1]
    # (nv, nu, q, q) detector points, pixel-bound coordinates
    V, U, QA, QB = np.meshgrid(np.arange(geom.nv), np.arange(geom.nu), off, off, indexing="ij")
    uv = np.stack([(U + QA).ravel(), (V + QB).ravel()], axis=1)
    wq = np.outer(w, w).ravel() / 4.0
    scale = geom.s_scl * geom.u_scl * geom.v_scl
    out = torch.zeros((len(views), geom.nv, geom.nu), dtype=torch.float32, device=device)
    wq_t = torch.tensor(wq, dtype=torch.float64, device=device)
    for n, v in enumerate(views):
        src, d = cone_rays(geom, v, uv)
        D = torch.tensor(d, dtype=torch.float64, device=device)
        acc = torch.zeros(D.shape[0], dtype=torch.float64, device=device)
        rows = ells[n] if ells and isinstance(ells[0], list) else ells
        for c, a, rot, val in rows:
            rot = np.asarray(rot, dtype=np.float64)
            ainv = 1.0 / np.asarray(a, dtype=np.float64)
            s0 = (rot.T @ (src - (np.asarray(c, dtype=np.float64) + half))) * ainv
            M = torch.tensor((rot * ainv[None, :]).T, dtype=torch.float64, device=device)
            Dp = D @ M.T
            c2 = (Dp * Dp).sum(1)
            c1 = Dp @ torch.tensor(s0, dtype=torch.float64, device=device)
            c0 = float(s0 @ s0) - 1.0
            det = c1 * c1 - c0 * c2
            acc += val * 2.0 * torch.sqrt(torch.clamp(det, min=0.0)) / c2
        pix = (acc.view(geom.nv, geom.nu, -1) * wq_t.view(1, 1, -1)).sum(-1)
        out[n] = (scale * pix).float()
    return out


def body_phantom(n, seed, dtype=np.float32):
    """A body ellipsoid (value 0.3-0.6) holding 4-12 inner ellipsoids of
    value -0.25..0.6 relative to it, clipped at 0; an n^3 grid."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    vol = np.zeros(shape, dtype=np.float64)
    r = n / 2.0
    body_axes = rng.uniform(0.6, 0.85, size=3) * r
    angle = rng.uniform(0, np.pi)
    c, s = np.cos(angle), np.sin(angle)
    body_rot = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    add_ellipsoid(vol, rng.uniform(-0.05, 0.05, size=3) * r, body_axes, body_rot,
                  rng.uniform(0.3, 0.6))
    for _ in range(int(rng.integers(4, 13))):
        c = rng.uniform(-0.45, 0.45, size=3) * body_axes
        a = rng.uniform(0.05, 0.3, size=3) * r
        add_ellipsoid(vol, c, a, _rotation(rng), rng.uniform(-0.25, 0.6))
    return np.maximum(vol, 0.0).astype(dtype)


def ellipses_2d(n, n_ell, seed, centre_frac=0.8, axes_frac=(0.05, 0.4), dtype=np.float64):
    """n_ell additive ellipses on an n x n image: centres uniform in the disk
    of radius centre_frac * n/2, semi-axes in axes_frac * n/2, angle in
    [0, pi), values in [0, 1].  Returns (image, params) with params rows
    (cx, cy, a, b, angle, value) in pixel units."""
    rng = np.random.default_rng(seed)
    r = n / 2.0
    yy, xx = np.meshgrid(np.arange(n) - (n - 1) / 2, np.arange(n) - (n - 1) / 2, indexing="ij")
    img = np.zeros((n, n), dtype=dtype)
    rows = []
    for _ in range(n_ell):
        rho = centre_frac * r * np.sqrt(rng.uniform())
        phi = rng.uniform(0, 2 * np.pi)
        cx, cy = rho * np.cos(phi), rho * np.sin(phi)
        a, b = rng.uniform(axes_frac[0], axes_frac[1], size=2) * r
        ang = rng.uniform(0, np.pi)
        val = rng.uniform(0, 1)
        ca, sa = np.cos(ang), np.sin(ang)
        u = (xx - cx) * ca + (yy - cy) * sa
        v = -(xx - cx) * sa + (yy - cy) * ca
        img += np.where((u / a) ** 2 + (v / b) ** 2 <= 1.0, val, 0.0).astype(dtype)
        rows.append((cx, cy, a, b, ang, val))
    return img, np.array(rows)


def _band_parallel(ell, theta, lo, hi):
    """Exact integral over detector coordinates [lo, hi] of the parallel
    projection at angle theta of one ellipse row (cx, cy, a, b, alpha, mu):
    the chord profile 2ab/r^2 sqrt(r^2 - (v - v0)^2), r^2 = a^2 cos^2(theta -
    alpha) + b^2 sin^2(theta - alpha), v0 = cx cos(theta) + cy sin(theta)
    (Tomography_2D/Ytrtt2D_simulated_data_generation.i:261-302), integrated
    with its antiderivative instead of Romberg quadrature (:520-530)."""
    cx, cy, a, b, alpha, mu = ell
    tp = theta - alpha
    r2 = (a * np.cos(tp)) ** 2 + (b * np.sin(tp)) ** 2
    r = np.sqrt(r2)
    v0 = cx * np.cos(theta) + cy * np.sin(theta)

    def prim(t):
        t = np.clip(t, -r, r)
        return (a * b / r2) * (t * np.sqrt(np.maximum(r2 - t * t, 0.0)) + r2 * np.arcsin(t / r))

    return mu * (prim(hi - v0) - prim(lo - v0))


def radon_ellipses_parallel(ells, theta, v, v_scl):
    """Band-integrated parallel-beam sinogram [view][bin] of a sum of
    ellipses (rows of :func:`ellipses_2d`'s params, one length unit per
    pixel): bin q covers [v_q - v_scl/2, v_q + v_scl/2]
    (Ytrtt2D_simulated_data_generation.i:609-659)."""
    theta = np.asarray(theta, dtype=np.float64)[:, None]
    v = np.asarray(v, dtype=np.float64)[None, :]
    out = np.zeros((theta.shape[0], v.shape[1]))
    for ell in ells:
        out += _band_parallel(ell, theta, v - 0.5 * v_scl, v + 0.5 * v_scl)
    return out


def radon_ellipses_fan(ells, theta, v, v_scl, Rsc, Rsd, nodes=8):
    """Band-integrated flat-detector fan-beam sinogram: the ray through
    detector coordinate v at view angle theta is the parallel ray of angle
    theta + atan(v/Rsd) at offset Rsc sin(atan(v/Rsd))
    (Ytrtt2D_simulated_data_generation.i:304-350); each bin is integrated
    over its width by Gauss-Legendre quadrature with ``nodes`` points
    (the reference uses Romberg, :595-603)."""
    theta = np.asarray(theta, dtype=np.float64)[:, None, None]
    v = np.asarray(v, dtype=np.float64)[None, :, None]
    g, wq = np.polynomial.legendre.leggauss(nodes)
    vv = v + 0.5 * v_scl * g[None, None, :]
    gam = np.arctan(vv / Rsd)
    beta = theta + gam
    veq = Rsc * np.sin(gam)
    prof = np.zeros(np.broadcast_shapes(beta.shape, veq.shape))
    for cx, cy, a, b, alpha, mu in ells:
        tp = beta - alpha
        r2 = (a * np.cos(tp)) ** 2 + (b * np.sin(tp)) ** 2
        v0 = cx * np.cos(beta) + cy * np.sin(beta)
        prof += mu * (2.0 * a * b / r2) * np.sqrt(np.maximum(r2 - (veq - v0) ** 2, 0.0))
    return (prof * wq[None, None, :]).sum(axis=-1) * (0.5 * v_scl)


def transmission_noise(line_integrals, photons, mu_scale=1.0, read_noise=0.0, generator=None):
    """Noisy line integrals of a transmission measurement and their weights.

    ``line_integrals`` (torch tensor, any device) are in object units times
    length; the attenuation is ``mu_scale`` times that.  Counts are Poisson
    with mean I0 exp(-mu_scale * p); the data returned are back in object
    units, ``-log(max(N, 1) / I0) / mu_scale``, with the weights of the
    reference model, 1/var = z^2 / (z + sigma_r^2) with z the counts, scaled
    by mu_scale^2 to the same units."""
    import torch

    mean = photons * torch.exp(-mu_scale * line_integrals.double())
    counts = torch.poisson(mean, generator=generator).clamp_(min=1.0)
    y = -torch.log(counts / photons) / mu_scale
    w = counts * counts / (counts + read_noise ** 2) * (mu_scale * mu_scale)
    return y.to(line_integrals.dtype), w.to(line_integrals.dtype)
