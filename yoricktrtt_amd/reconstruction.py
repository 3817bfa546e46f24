"""Objective of the TRTT reconstruction on the GPU.

* :class:`DataTerm` is the weighted least-squares data term of the per-view /
  per-phase operator loop (Tomography_4D/Ytrtt4D_optim.i:361-487):

      f(x) = 1/2 sum_k (C_k x - y_k)' W_k (C_k x - y_k)
      g(x) = sum_k C_k' W_k (C_k x - y_k)

  where C_k = A_k B_k (projector of view k after the time interpolation of
  its phase).  One call runs every view of every phase: forward launch,
  weighted residual with the FP64 sum, backward launch.  The sequence can be
  captured once in a HIP graph and replayed per evaluation.
* :class:`TotalVariation` is the isotropic TV with threshold applied in
  sample space (B-spline coefficients -> samples, Ytrtt4D_optim.i:448-482).
* :func:`noise_weights` is the reference noise model
  (Tomography_4D/Ytrtt4D_system.i:829-839).

Torch supplies device memory and streams; the computation is in
libtrtt_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class DataTerm:
    """Weighted least squares over every view of a :class:`Projector`."""

    def __init__(self, projector, y, w=None):
        import torch

        self.torch = torch
        self.P = projector
        shape = tuple(projector.sino_shape)
        # data and weights in the plan's type (the residual kernel reads them as such)
        self.y = y.to(dtype=projector.dtype).contiguous()
        self.w = None if w is None else w.to(dtype=projector.dtype).contiguous()
        for name, t in (("y", self.y), ("w", self.w)):
            if t is not None and tuple(t.shape) != shape:
                raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")
        self.sino = torch.zeros(shape, dtype=projector.dtype, device=y.device)
        self.fx = torch.zeros(1, dtype=torch.float64, device=y.device)
        self.graph = None
        self.x_static = None
        self.g_static = None

    def _run(self, x, g):
        self.sino.zero_()
        self.fx.zero_()
        self.P.forward(x, self.sino)
        self.P.weighted_residual(self.sino, self.y, self.w, self.fx)
        g.zero_()
        self.P.backward(self.sino, g)

    def check(self) -> None:
        """Raise when a launch met a voxel in the source plane, the case where
        the reference projectors return TRTT_FAILURE (trtt_3D_projectors.c:
        1264-1284); the device flag is sticky, so one check covers every
        evaluation since the plan was made.  Synchronous."""
        flag = getattr(self.P, "error_flag", None)
        if flag is not None and flag() != 0:
            raise RuntimeError("projector failure: a voxel lies in the source plane")

    def eval(self, x, g) -> float:
        """Return f(x) (FP64) and write g(x) into g.  Eager evaluations check
        the projector's error flag; graph replays leave it to check()."""
        if self.graph is not None:
            if x.data_ptr() != self.x_static.data_ptr():
                self.x_static.copy_(x)
            self.graph.replay()
            if g.data_ptr() != self.g_static.data_ptr():
                g.copy_(self.g_static)
            return float(self.fx.item())
        self._run(x, g)
        f = float(self.fx.item())
        self.check()
        return f

    def capture(self, x_static, g_static):
        """Capture the whole evaluation (all launches) in one HIP graph with
        x_static / g_static as its fixed input / output buffers."""
        torch = self.torch
        self.x_static, self.g_static = x_static, g_static
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._run(x_static, g_static)  # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.check()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self._run(x_static, g_static)
        self.graph = graph
        return graph


class TotalVariation:
    """w * sum (sqrt(eps^2 + |D S x|^2) - eps), S the B-spline sampling
    (or the identity when ``samples`` is False), for an image [ny][nx], a
    volume [nz][ny][nx] or a series [nt][nz][ny][nx]."""

    def __init__(self, shape, weight, eps, wt=1.0, samples=True, dtype=None, device=None):
        import torch

        self.torch = torch
        self.shape = tuple(shape)
        self.nt = shape[0] if len(shape) == 4 else 1
        self.nz, self.ny, self.nx = ((1,) + self.shape)[-3:] if len(shape) == 2 else shape[-3:]
        self.weight, self.eps, self.wt = float(weight), float(eps), float(wt)
        self.samples = samples
        dtype = dtype or torch.float32
        self.prec = _lib.F64 if dtype == torch.float64 else _lib.F32
        ncomp = 4 if self.nt > 1 else 3
        n = self.nt * self.nz * self.ny * self.nx
        self.work = torch.empty(ncomp * n, dtype=dtype, device=device)
        self.xs = torch.empty(self.shape, dtype=dtype, device=device) if samples else None
        self.gs = torch.empty(self.shape, dtype=dtype, device=device) if samples else None
        self.f = torch.zeros(1, dtype=torch.float64, device=device)
        self.lib = _lib.load()

    def _filter(self, src, dst, accumulate=False):
        r = self.lib.trtt_hip_bspline_filter(_vp(src), _vp(dst), self.nx, self.ny, self.nz, self.nt,
                                             int(accumulate), self.prec, _stream(self.torch))
        if r != 0:
            raise RuntimeError("bspline filter launch failed")

    def eval(self, x, g) -> float:
        """Return the TV value (FP64) and add its gradient to g."""
        self.f.zero_()
        if self.samples:
            self._filter(x, self.xs)
            self.gs.zero_()
            r = self.lib.trtt_hip_tv(_vp(self.xs), _vp(self.gs), _vp(self.work), _vp(self.f),
                                     self.nx, self.ny, self.nz, self.nt, self.weight, self.wt,
                                     self.eps, self.prec, _stream(self.torch))
            self._filter(self.gs, g, accumulate=True)
        else:
            r = self.lib.trtt_hip_tv(_vp(x), _vp(g), _vp(self.work), _vp(self.f), self.nx, self.ny,
                                     self.nz, self.nt, self.weight, self.wt, self.eps, self.prec,
                                     _stream(self.torch))
        if r != 0:
            raise RuntimeError("TV launch failed")
        return float(self.f.item())


class Objective:
    """Sum of a data term and optional regularisers: f, g = objective(x, g)."""

    def __init__(self, data: DataTerm, regs=()):
        self.data = data
        self.regs = list(regs)
        self.nevals = 0

    def __call__(self, x, g) -> float:
        f = self.data.eval(x, g)
        for r in self.regs:
            f += r.eval(x, g)
        self.nevals += 1
        return f


def noise_weights(y, photon_flux, read_noise=0.0):
    """Inverse variances of the transmission noise model
    var = (z + sigma_r^2) / z^2, z = I0 exp(-y) (Ytrtt4D_system.i:829-839);
    unit weights without photon flux."""
    y = np.asarray(y, dtype=np.float64)
    if photon_flux == 0.0:
        return np.ones_like(y)
    z = photon_flux * np.exp(-y)
    var = (z + read_noise ** 2) / (z * z)
    return 1.0 / var


def add_noise(y, photon_flux, read_noise=0.0, seed=0):
    """Noisy data following the same model (Gaussian approximation,
    Ytrtt4D_system.i:840-843) and its weights."""
    rng = np.random.default_rng(seed)
    w = noise_weights(y, photon_flux, read_noise)
    noisy = np.maximum(0.0, y + rng.standard_normal(y.shape) / np.sqrt(w))
    return noisy, w
