"""Drivers of the benchmark configurations on one GPU.

Each driver builds its geometry, seeded phantom and data, times one forward
and one back projection over every view (inputs resident in HBM), then runs
VMLM-B with TV in sample space and positivity, and returns a dict of timings
and results.  Configuration 5 (the learned warm start) lives in
:mod:`yoricktrtt_amd.warmstart`.

1. 2-D parallel beam, 256 x 256 cubic B-spline image, 360 views over 180 deg
   (k pi / 360, Ytrtt2D_system.i:501-505), 367 bins at unit pitch, FP64; 10
   random ellipses (seed 1); band-integrated analytic parallel-beam data
   (Ytrtt2D_simulated_data_generation.i:609-659, called with the view angle
   minus pi/2 as in Ytrtt2D_script_reconstructions.i:46); 20 iterations.
2. 2-D fan beam, 512 x 512 pixels of 0.5 mm, 720 views over 360 deg, 736 flat
   detector bins of 0.75 mm (the reference has no curved detector),
   Rsc = 1000 mm, Rsd = 1500 mm, FP32 data with FP64 scalars; 20 ellipses
   (seed 2) of 0.02 mm^-1 per unit; analytic fan-beam data with Poisson
   noise at 1e5 photons and inverse-variance weights; 50 iterations; then the
   Long & Fessler SF projector on the same geometry (timing, adjoint test
   and its own 50 iterations), and the sparse-matrix mode (matrix built on
   the GPU from the tabulated footprint, products timed, its own solve).
3. 3-D cone beam, 256^3 voxels of 1 mm, 360 views over 360 deg, 512 (v,
   transaxial) x 384 (u, axial) flat detector at 1 mm, Rsc = 1000 mm,
   Rsd = 1500 mm, FP32; 30 random ellipsoids (seed 3); data = the
   analytic cone-beam line integrals of the ellipsoids averaged over each
   detector pixel (phantoms.cone_projection_ellipsoids, in the projector's
   units: pitch^2 times the integral in mm) with Poisson noise (1e5
   photons, 0.02 mm^-1 per unit); 30 iterations.
4. 4-D cone beam: 10 phases of 128 x 128 x 64 voxels of 2 mm, 256 (v) x 192
   (u) detector at 2 mm, distances of configuration 3; view counts per phase
   uniform in 30..60 (seed 4), angles evenly spaced over 360 deg and dealt
   to the phases in a seeded random order; a body ellipsoid whose inner
   ellipsoid moves by 10 sin(2 pi p / 10) mm along z; data = the analytic
   cone-beam integrals of each view's phase (as for configuration 3) with
   Poisson noise; all phases in one plan (one launch sequence per
   projection, the data term captured in one HIP graph per evaluation), TV
   along t as well; 30 iterations.
6. 2-D+t fan beam (the dynamic scripts of Tomography_2D_dyn): 16 phases of
   256 x 256 pixels of 1 mm, a moving inner ellipse; 720 views over 360 deg
   acquired during 127 s of a 7 s motion; 512 bins of 1 mm, Rsc = 1000 mm,
   Rsd = 1536 mm; cubic B-spline in time; data = this operator's projection
   with Poisson noise; TV along t as well; 30 iterations.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np


def _torch():
    import torch

    return torch


def _sync():
    _torch().cuda.synchronize()
    return time.perf_counter()


def _time(fn, reps=3):
    """Median wall time of fn() on the GPU, after one warm-up call."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = _sync()
        fn()
        ts.append(_sync() - t0)
    return float(np.median(ts))


def _projection_times(P, x, sino):
    torch = _torch()
    fwd = _time(lambda: P.forward(x, sino.zero_()))
    vol = torch.zeros_like(x)
    bwd = _time(lambda: P.backward(sino, vol.zero_()))
    return fwd, bwd


TUNE_CACHE = None  # set by main(); None: the table shipped with the package


def _tuned(P):
    """Apply the autotuned launch shape to an axial plan (cached per GPU and
    geometry; tuned on the spot when missing)."""
    import os

    from .autotune import autotune

    if P.path > 0:
        cache = TUNE_CACHE or os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning",
                                           "tuning.json")
        autotune(P, cache=cache)
    return P


def _solve(obj, x0, iters, truth=None, sample=None):
    """VMLM-B with positivity; per-iteration wall time (after a short
    warm-up solve) and final RMSE of the B-spline samples against the truth."""
    from .vmlmb import vmlmb

    torch = _torch()
    # first-call costs (kernel loading of every operation the solver uses)
    # outside the timing: a 2-iteration solve from the same start
    vmlmb(obj, x0.clone(), lower=0.0, maxiter=2)
    t0 = _sync()
    r = vmlmb(obj, x0, lower=0.0, maxiter=iters)
    t = _sync() - t0
    obj.data.check()
    out = dict(iters=r.niter, evals=r.neval, status=r.status, solve_s=t,
               iter_ms=1e3 * t / max(r.niter, 1), f0=r.history[0], f=r.f)
    if truth is not None:
        xs = sample(r.x) if sample is not None else r.x
        out["rmse"] = float(torch.sqrt(torch.mean((xs.double() - truth.double()) ** 2)).item())
    return r, out


def _sampler(shape, dtype, device):
    """x -> B-spline samples S x (the filter of the TV term)."""
    import ctypes as C

    from . import _lib

    torch = _torch()
    lib = _lib.load()
    prec = _lib.F64 if dtype == torch.float64 else _lib.F32
    nt = shape[0] if len(shape) == 4 else 1
    nz, ny, nx = ((1,) + tuple(shape))[-3:] if len(shape) == 2 else tuple(shape)[-3:]

    def sample(x):
        out = torch.empty_like(x)
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        if lib.trtt_hip_bspline_filter(C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), nx,
                                       ny, nz, nt, 0, prec, st) != 0:
            raise RuntimeError("bspline filter failed")
        return out

    return sample


def config1_setup():
    """Geometry of configuration 1: 2-D parallel, 256^2, 360 views / 180 deg,
    367 unit bins."""
    from .geometry import parallel_geometry_2d

    n, nv, nviews = 256, 367, 360
    theta = np.arange(nviews) * (math.pi / nviews)
    return parallel_geometry_2d(n, n, 1.0, nv, 1.0, theta)


def config2_setup():
    """Geometry of configuration 2 (2-D fan, 512^2 x 0.5 mm, 720 views /
    360 deg, 736 flat bins x 0.75 mm, Rsc 1000, Rsd 1500) and the
    parameters of its Long & Fessler projector (detector on the rotating
    frame at Rsd - Rsc from the centre, bin pitch = bin size; source angle =
    view angle - pi/2)."""
    from . import _lib
    from .geometry import fan_geometry_2d

    n, s, nv, vs, nviews, Rsc, Rsd = 512, 0.5, 736, 0.75, 720, 1000.0, 1500.0
    theta = np.arange(nviews) * (2 * math.pi / nviews)
    geo = fan_geometry_2d(n, n, s, nv, vs, theta, Rsc, Rsd)
    prm = _lib.LF2DParameters(vs, vs, s, Rsc, Rsd - Rsc, 0.0, 0.0, n, n, nv)
    return geo, prm, theta - 0.5 * math.pi


def config1(iters=20, tv_weight=0.05, tv_eps=0.01):
    from .phantoms import ellipses_2d, radon_ellipses_parallel
    from .projectors import Projector
    from .reconstruction import DataTerm, Objective, TotalVariation

    torch = _torch()
    geo = config1_setup()
    n, nv, theta = geo.nx, geo.nv, geo.theta
    img, ells = ellipses_2d(n, 10, seed=1)
    v = (np.arange(nv) - (nv - 1) / 2) * 1.0
    y = radon_ellipses_parallel(ells, theta - 0.5 * math.pi, v, 1.0)
    dev = torch.device("cuda")
    P = Projector(geo, model="cubic", precision="f64")
    truth = torch.tensor(img, device=dev)
    sino = torch.zeros(P.sino_shape, dtype=torch.float64, device=dev)
    fwd, bwd = _projection_times(P, truth, sino)
    D = DataTerm(P, torch.tensor(y, device=dev))
    tv = TotalVariation((n, n), tv_weight, tv_eps, samples=True, dtype=torch.float64, device=dev)
    D.capture(torch.zeros((n, n), dtype=torch.float64, device=dev),
              torch.zeros((n, n), dtype=torch.float64, device=dev))
    _, sol = _solve(Objective(D, [tv]), torch.zeros((n, n), dtype=torch.float64, device=dev),
                    iters, truth, _sampler((n, n), torch.float64, dev))
    P.close()
    return dict(config=1, workload="2D parallel 256^2, 360 views / 180 deg, 367 bins, FP64",
                fwd_ms=1e3 * fwd, bwd_ms=1e3 * bwd, pairs_per_s=1.0 / (fwd + bwd), **sol)


def config2(iters=50, tv_weight=2.0, tv_eps=0.01, photons=1e5, mu_scale=0.02):
    from .phantoms import ellipses_2d, radon_ellipses_fan, transmission_noise
    from .projectors import LFProjector, Projector
    from .reconstruction import DataTerm, Objective, TotalVariation

    torch = _torch()
    n, s, nv, vs, nviews, Rsc, Rsd = 512, 0.5, 736, 0.75, 720, 1000.0, 1500.0
    geo, prm, beta = config2_setup()
    theta = geo.theta
    img, ells = ellipses_2d(n, 20, seed=2)
    ells_mm = ells.copy()
    ells_mm[:, :4] *= s  # pixel units -> mm
    v = (np.arange(nv) - (nv - 1) / 2) * vs
    band = radon_ellipses_fan(ells_mm, theta - 0.5 * math.pi, v, vs, Rsc, Rsd)
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev).manual_seed(2)
    # noise on the bin-averaged line integral, data and weights back in band units
    yavg, wavg = transmission_noise(torch.tensor(band / vs, device=dev), photons, mu_scale,
                                    generator=gen)
    y = (yavg * vs).float()
    w = (wavg / (vs * vs)).float()
    P = Projector(geo, model="cubic", precision="f32")
    truth = torch.tensor(img, dtype=torch.float32, device=dev)
    sino = torch.zeros(P.sino_shape, dtype=torch.float32, device=dev)
    fwd, bwd = _projection_times(P, truth, sino)
    D = DataTerm(P, y, w)
    tv = TotalVariation((n, n), tv_weight, tv_eps, samples=True, dtype=torch.float32, device=dev)
    D.capture(torch.zeros((n, n), dtype=torch.float32, device=dev),
              torch.zeros((n, n), dtype=torch.float32, device=dev))
    sample = _sampler((n, n), torch.float32, dev)
    _, sol = _solve(Objective(D, [tv]), torch.zeros((n, n), dtype=torch.float32, device=dev),
                    iters, truth, sample)
    P.close()
    out = dict(config=2, workload="2D fan 512^2 x 0.5 mm, 720 views / 360 deg, 736 flat bins "
               "0.75 mm, Rsc 1000, Rsd 1500, FP32, 1e5 photons",
               fwd_ms=1e3 * fwd, bwd_ms=1e3 * bwd, pairs_per_s=1.0 / (fwd + bwd), **sol)
    # Long & Fessler SF projector on the same geometry (config2_setup)
    lf = LFProjector(prm, beta, np.zeros(nviews), precision="f32")
    sino_lf = torch.zeros(lf.sino_shape, dtype=torch.float32, device=dev)
    lf_fwd, lf_bwd = _projection_times(lf, truth, sino_lf)
    xr = torch.rand((n, n), generator=torch.Generator(device=dev).manual_seed(7), device=dev)
    yr = torch.rand(lf.sino_shape, generator=torch.Generator(device=dev).manual_seed(8), device=dev)
    ax = lf.forward(xr)
    aty = lf.backward(yr.contiguous())
    lhs = float(torch.dot(ax.double().reshape(-1), yr.double().reshape(-1)))
    rhs = float(torch.dot(xr.double().reshape(-1), aty.double().reshape(-1)))
    D2 = DataTerm(lf, y, w)
    _, sol2 = _solve(Objective(D2, [tv]), torch.zeros((n, n), dtype=torch.float32, device=dev),
                     iters, truth, sample)
    lf.close()
    out["lf2d"] = dict(fwd_ms=1e3 * lf_fwd, bwd_ms=1e3 * lf_bwd,
                       adjoint_rel=abs(lhs - rhs) / abs(lhs), **sol2)
    out["sparse"] = _config2_sparse(geo, truth, y, w, tv, sample, iters)
    return out


def _config2_sparse(geo, truth, y, w, tv, sample, iters):
    """The reference's sparse-matrix mode on the configuration-2 geometry:
    tabulated footprint, coefficients rounded to 6 decimals
    (Ytrtt_2D_projectors.i:212-369), the matrix built on the GPU, the two
    products timed against the table projector, then the same solve."""
    from .projectors import Projector
    from .reconstruction import DataTerm, Objective

    torch = _torch()
    dev = truth.device
    Pt = Projector(geo, model="table", precision="f32")
    t0 = _sync()
    M = Pt.sparse_matrix(6)
    build = _sync() - t0
    sino = torch.zeros(Pt.sino_shape, dtype=torch.float32, device=dev)
    fwd, bwd = _projection_times(M, truth, sino)
    pfwd, pbwd = _projection_times(Pt, truth, sino)
    yr = torch.rand(Pt.sino_shape, generator=torch.Generator(device=dev).manual_seed(9), device=dev)
    mx, px = M.forward(truth), Pt.forward(truth)
    my, py = M.backward(yr), Pt.backward(yr)
    rel = (float((mx - px).abs().max() / px.abs().max()),
           float((my - py).abs().max() / py.abs().max()))
    # bytes per product: every entry (FP32 value + 32-bit index), the pointers,
    # the gathered vector once and the output read and written
    nnz, nrows, ncols = M.nnz, geo.nviews * geo.nv, geo.nx * geo.ny
    b_fwd = nnz * 8 + (nrows + 1) * 4 + ncols * 4 + nrows * 8
    b_bwd = nnz * 8 + (ncols + 1) * 4 + nrows * 4 + ncols * 8
    D = DataTerm(M, y, w)
    n = geo.nx
    D.capture(torch.zeros((n, n), dtype=torch.float32, device=dev),
              torch.zeros((n, n), dtype=torch.float32, device=dev))
    _, sol = _solve(Objective(D, [tv]), torch.zeros((n, n), dtype=torch.float32, device=dev),
                    iters, truth, sample)
    M.close()
    Pt.close()
    return dict(nnz=nnz, build_s=build, fwd_ms=1e3 * fwd, bwd_ms=1e3 * bwd,
                fwd_GBps=b_fwd / fwd / 1e9, bwd_GBps=b_bwd / bwd / 1e9,
                table_projector_fwd_ms=1e3 * pfwd, table_projector_bwd_ms=1e3 * pbwd,
                rel_vs_table_projector=rel, **sol)


def cone_setup(n=256, nviews=360, nu=384, nv=512, pitch=1.0, Rsc=1000.0, Rsd=1500.0):
    from .geometry import cone_geometry

    theta = np.arange(nviews) * (2 * math.pi / nviews)
    return cone_geometry(n, n, n, 1.0, nu, nv, pitch, pitch, theta, Rsc, Rsd)


def config3(iters=30, tv_weight=5.0, tv_eps=0.05, photons=1e5, mu_scale=0.02):
    from .phantoms import cone_projection_ellipsoids, ellipsoids_3d, transmission_noise
    from .projectors import Projector
    from .reconstruction import DataTerm, Objective, TotalVariation

    torch = _torch()
    n = 256
    geo = cone_setup(n)
    dev = torch.device("cuda")
    vol, ells = ellipsoids_3d((n, n, n), 30, seed=3, params=True)
    truth = torch.from_numpy(vol).to(dev)
    P = _tuned(Projector(geo, model="cubic", precision="f32"))
    sino = torch.zeros(P.sino_shape, dtype=torch.float32, device=dev)
    fwd, bwd = _projection_times(P, truth, sino)
    del sino
    # analytic data of the 30 ellipsoids (pixel-averaged cone-beam line
    # integrals), not this projector's own output
    data = cone_projection_ellipsoids(geo, ells, device=dev)
    gen = torch.Generator(device=dev).manual_seed(3)
    y, w = transmission_noise(data, photons, mu_scale, generator=gen)
    del data
    D = DataTerm(P, y, w)
    tv = TotalVariation((n, n, n), tv_weight, tv_eps, samples=True, dtype=torch.float32,
                        device=dev)
    D.capture(torch.zeros((n, n, n), device=dev), torch.zeros((n, n, n), device=dev))
    _, sol = _solve(Objective(D, [tv]), torch.zeros((n, n, n), device=dev), iters, truth,
                    _sampler((n, n, n), torch.float32, dev))
    P.close()
    return dict(config=3, workload="3D cone 256^3 x 1 mm, 360 views / 360 deg, 512x384 flat 1 mm, "
                "Rsc 1000, Rsd 1500, FP32, analytic ellipsoid data, 1e5 photons",
                fwd_ms=1e3 * fwd, bwd_ms=1e3 * bwd, pairs_per_s=1.0 / (fwd + bwd), **sol)


def respiratory_series(nt=10, shape=(64, 128, 128), voxel=2.0, amplitude=10.0, seed=4,
                       params=False):
    """Body ellipsoid with fixed structures and an inner ellipsoid moving by
    amplitude sin(2 pi p / nt) mm along z, for phases p = 0..nt-1.  With
    ``params`` also the ellipsoids of every phase (centre, axes, rotation,
    value in grid units), for the analytic data."""
    from .phantoms import add_ellipsoid

    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    eye = np.eye(3)
    base = np.zeros(shape, dtype=np.float64)
    rows = [(np.zeros(3), np.array([0.42 * nx, 0.32 * ny, 0.45 * nz]), eye, 0.4)]
    for _ in range(6):
        c = rng.uniform(-0.25, 0.25, size=3) * np.array([nx, ny, nz])
        a = rng.uniform(0.04, 0.1, size=3) * np.array([nx, ny, nz])
        rows.append((c, a, eye, rng.uniform(0.1, 0.5)))
    for c, a, r, v in rows:
        add_ellipsoid(base, c, a, r, v)
    lesion_c = np.array([0.12 * nx, -0.05 * ny, 0.0])
    lesion_a = np.array([0.08 * nx, 0.08 * ny, 0.12 * nz])
    series = np.empty((nt,) + shape, dtype=np.float32)
    phases = []
    for p in range(nt):
        vol = base.copy()
        dz = amplitude * math.sin(2 * math.pi * p / nt) / voxel
        lc = lesion_c + np.array([0.0, 0.0, dz])
        add_ellipsoid(vol, lc, lesion_a, eye, 0.5)
        series[p] = vol
        phases.append(rows + [(lc, lesion_a, eye, 0.5)])
    return (series, phases) if params else series


def config4_setup(nt=10, seed=4):
    """Geometry (views sorted by angle) and the phase of every view."""
    from .geometry import cone_geometry

    rng = np.random.default_rng(seed)
    counts = rng.integers(30, 61, size=nt)
    total = int(counts.sum())
    phase = rng.permutation(np.repeat(np.arange(nt), counts)).astype(np.int32)
    theta = np.arange(total) * (2 * math.pi / total)
    geo = cone_geometry(128, 128, 64, 2.0, 192, 256, 2.0, 2.0, theta, 1000.0, 1500.0)
    return geo, phase, counts


def config4(iters=30, tv_weight=5.0, tv_eps=0.05, wt=1.0, photons=1e5, mu_scale=0.005):
    from .phantoms import cone_projection_ellipsoids, transmission_noise
    from .projectors import Projector
    from .reconstruction import DataTerm, Objective, TotalVariation

    torch = _torch()
    nt = 10
    geo, phase, counts = config4_setup(nt)
    dev = torch.device("cuda")
    series, phase_ells = respiratory_series(nt, params=True)
    truth = torch.from_numpy(series).to(dev)
    P = _tuned(Projector(geo, model="cubic", precision="f32", nt=nt, t0=phase, t1=phase,
                         w0=np.ones(phase.size), w1=np.zeros(phase.size)))
    sino = torch.zeros(P.sino_shape, dtype=torch.float32, device=dev)
    fwd, bwd = _projection_times(P, truth, sino)
    del sino
    # analytic data of each view's phase (pixel-averaged line integrals in
    # the projector's units, s_scl u_scl v_scl = 8 times the integral in
    # voxel lengths, i.e. pitch^2 times the integral in mm), so 0.005 per
    # unit is 0.02 mm^-1 per object unit at the 2 mm pitch
    data = cone_projection_ellipsoids(geo, [phase_ells[p] for p in phase], device=dev)
    gen = torch.Generator(device=dev).manual_seed(4)
    y, w = transmission_noise(data, photons, mu_scale, generator=gen)
    del data
    D = DataTerm(P, y, w)
    shape = (nt, 64, 128, 128)
    tv = TotalVariation(shape, tv_weight, tv_eps, wt=wt, samples=True, dtype=torch.float32,
                        device=dev)
    D.capture(torch.zeros(shape, device=dev), torch.zeros(shape, device=dev))
    _, sol = _solve(Objective(D, [tv]), torch.zeros(shape, device=dev), iters, truth,
                    _sampler(shape, torch.float32, dev))
    P.close()
    return dict(config=4, workload="4D cone, 10 phases x 128x128x64 x 2 mm, 256x192 detector 2 mm, "
                "views per phase 30..60, analytic ellipsoid data", views=int(phase.size), views_per_phase=counts.tolist(),
                fwd_ms=1e3 * fwd, bwd_ms=1e3 * bwd, pairs_per_s=1.0 / (fwd + bwd), **sol)


def dynamic_series_2d(n, nt, seed=6):
    """nt frames of an n x n image: 12 fixed ellipses (seed) and an inner
    ellipse of value 0.8 whose centre moves by 0.08 n sin(2 pi p / nt)
    along y (a breathing motion)."""
    from .phantoms import ellipses_2d

    base, _ = ellipses_2d(n, 12, seed, centre_frac=0.5, axes_frac=(0.05, 0.3),
                          dtype=np.float32)
    yy, xx = np.meshgrid(np.arange(n) - (n - 1) / 2, np.arange(n) - (n - 1) / 2, indexing="ij")
    out = np.empty((nt, n, n), dtype=np.float32)
    for p in range(nt):
        cy = 0.08 * n * math.sin(2 * math.pi * p / nt)
        inner = ((xx / (0.1 * n)) ** 2 + ((yy - cy) / (0.07 * n)) ** 2) <= 1.0
        out[p] = base + np.where(inner, np.float32(0.8), np.float32(0))
    return out


def config6(iters=30, tv_weight=2.0, tv_eps=0.01, wt=1.0, photons=1e5, mu_scale=0.02):
    """2-D+t fan beam (Tomography_2D_dyn, Ytrtt2D_script_reconstructions_
    dynamic_new.i:5-59): views over one turn acquired during tacq = 127 s of
    a motion of period tcycl = 7 s, t_index = (t % tcycl) nt / tcycl
    (Ytrtt2D_system_dyn.i:500-509); cubic temporal B-spline (t_deg 3)."""
    from .geometry import fan_geometry_2d
    from .phantoms import transmission_noise
    from .projectors import DynamicProjector
    from .reconstruction import DataTerm, Objective, TotalVariation

    torch = _torch()
    n, nt, V, nv, Rsc, Rsd, tacq, tcycl = 256, 16, 720, 512, 1000.0, 1536.0, 127.0, 7.0
    theta = np.arange(V) * (2 * math.pi / V)
    t = np.arange(V) * (tacq / (V - 1))
    t_index = (t % tcycl) * (nt / tcycl)
    geo = fan_geometry_2d(n, n, 1.0, nv, 1.0, theta, Rsc, Rsd)
    dev = torch.device("cuda")
    truth = torch.from_numpy(dynamic_series_2d(n, nt)).to(dev)
    P = DynamicProjector(geo, nt, t_index, t_deg=3, model="cubic", precision="f32")
    sino = torch.zeros(P.sino_shape, dtype=torch.float32, device=dev)
    fwd, bwd = _projection_times(P, truth, sino)
    sino.zero_()
    P.forward(truth, sino)
    gen = torch.Generator(device=dev).manual_seed(6)
    y, w = transmission_noise(sino, photons, mu_scale, generator=gen)
    D = DataTerm(P, y, w)
    shape = (nt, n, n)
    # a series of nt single-slice frames: B-spline sampling and TV in x, y
    # within each frame, TV along t across frames
    series = (nt, 1, n, n)
    tv = TotalVariation(series, tv_weight, tv_eps, wt=wt, samples=True, dtype=torch.float32,
                        device=dev)
    D.capture(torch.zeros(shape, device=dev), torch.zeros(shape, device=dev))
    _, sol = _solve(Objective(D, [tv]), torch.zeros(shape, device=dev), iters, truth,
                    _sampler(series, torch.float32, dev))
    P.close()
    return dict(config=6, workload="2D+t fan 256^2 x 1 mm, 16 phases, 720 views / 360 deg over "
                "127 s of a 7 s motion, 512 bins 1 mm, Rsc 1000, Rsd 1536, cubic in t, FP32",
                fwd_ms=1e3 * fwd, bwd_ms=1e3 * bwd, **sol)


RUNNERS = {1: config1, 2: config2, 3: config3, 4: config4, 6: config6}


def main(argv=None):
    ap = argparse.ArgumentParser(description="run benchmark configurations 1-4 and 6 on the GPU")
    ap.add_argument("configs", nargs="*", type=int, default=[1, 2, 3, 4])
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--out", default=None, help="append JSON lines here")
    ap.add_argument("--tune-cache", default=None, help="autotuner table (default: shipped one)")
    args = ap.parse_args(argv)
    global TUNE_CACHE
    TUNE_CACHE = args.tune_cache
    for k in args.configs:
        kw = {} if args.iters is None else {"iters": args.iters}
        t0 = time.perf_counter()
        res = RUNNERS[k](**kw)
        res["wall_s"] = time.perf_counter() - t0
        line = json.dumps(res)
        print(line, flush=True)
        if args.out:
            with open(args.out, "a") as f:
                f.write(line + "\n")


if __name__ == "__main__":
    main()
