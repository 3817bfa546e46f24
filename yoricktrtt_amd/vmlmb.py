"""VMLM-B: bound-constrained variable-metric limited-memory minimiser.

The reference drives its reconstructions with OptimPack's ``op_mnb`` /
``opk_minimize`` (Tomography_4D/Ytrtt4D_optim.i:933-941), which is not
available here.  This is this build's own implementation of the published
method (Thiebaut 2002, "Optimization issues in blind deconvolution
algorithms", SPIE 4847): an L-BFGS recursion restricted to the free
variables, a projected search path x(a) = P(x + a d) and an Armijo
backtracking search on that path, the first step scaled to unit length.

Vectors are torch tensors on any device; every inner product and every
line-search scalar is FP64.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field


def dot(a, b) -> float:
    """FP64 inner product of two tensors of any float dtype."""
    import torch

    if a.dtype == torch.float64 and b.dtype == torch.float64:
        return float(torch.dot(a.reshape(-1), b.reshape(-1)).item())
    return float(torch.dot(a.reshape(-1).double(), b.reshape(-1).double()).item())


@dataclass
class Result:
    x: object
    f: float
    niter: int
    neval: int
    status: str
    history: list = field(default_factory=list)
    pairs: list = field(default_factory=list)  # L-BFGS memory (s, y, 1/s'y) at exit
    grad: object = None  # gradient at x
    g0norm: float = None  # projected gradient norm of the first iteration (grtol scale)


def save_checkpoint(path, x, pairs, niter, f, g0norm=None):
    """Solver state after an iteration: the iterate, the L-BFGS memory, the
    iteration count, f and the grtol scale, as an .npz archive (plain
    arrays, no pickled objects)."""
    import numpy as np

    arrays = {"x": x.detach().cpu().numpy(), "niter": np.array(niter), "f": np.array(f),
              "g0norm": np.array(np.nan if g0norm is None else g0norm),
              "npairs": np.array(len(pairs))}
    for k, (s, y, rho) in enumerate(pairs):
        arrays[f"s{k}"] = s.detach().cpu().numpy()
        arrays[f"y{k}"] = y.detach().cpu().numpy()
        arrays[f"rho{k}"] = np.array(rho)
    tmp = str(path) + ".tmp.npz"
    np.savez(tmp, **arrays)
    import os

    os.replace(tmp, str(path))


def load_checkpoint(path, device=None):
    """(x, pairs, niter, f, g0norm) of a checkpoint written by save_checkpoint,
    with tensors on ``device``."""
    import numpy as np
    import torch

    with np.load(str(path), allow_pickle=False) as z:
        x = torch.from_numpy(z["x"].copy()).to(device)
        pairs = [(torch.from_numpy(z[f"s{k}"].copy()).to(device),
                  torch.from_numpy(z[f"y{k}"].copy()).to(device), float(z[f"rho{k}"]))
                 for k in range(int(z["npairs"]))]
        g0 = float(z["g0norm"])
        return x, pairs, int(z["niter"]), float(z["f"]), (None if math.isnan(g0) else g0)


def _project(x, lower, upper):
    if lower is not None:
        x.clamp_(min=lower)
    if upper is not None:
        x.clamp_(max=upper)
    return x


def vmlmb(fg, x0, lower=None, upper=None, mem=5, maxiter=50, maxeval=None, gatol=0.0,
          grtol=0.0, frtol=0.0, ftol=1e-4, fstop=None, callback=None, pairs=None, g0norm=None,
          checkpoint=None, checkpoint_every=1):
    """Minimise f over lower <= x <= upper.

    ``fg(x, g)`` returns f(x) as a float and writes the gradient into g.
    Stops after ``maxiter`` iterations, when the projected gradient norm falls
    below max(gatol, grtol |pg_0|), when the relative decrease of f over
    an iteration is below ``frtol``, or once f <= ``fstop``.
    ``callback(it, x, f)`` is called after each iteration.  ``pairs`` seeds
    the L-BFGS memory (e.g. the ``pairs`` of an earlier run on the same
    objective, so a restart from a new point keeps its curvature model).
    ``checkpoint`` names a file that receives the solver state every
    ``checkpoint_every`` iterations (save_checkpoint); a run started from a
    checkpoint's x, pairs and g0norm (load_checkpoint) continues with the
    same iterates as the uninterrupted run.
    """
    import torch

    x = _project(x0.clone(), lower, upper)
    g = torch.empty_like(x)
    f = fg(x, g)
    neval = 1
    pairs = list(pairs[-mem:]) if pairs else []  # (s, y, rho)
    history = [f]
    status = "maxiter"
    xn = torch.empty_like(x)
    gn = torch.empty_like(x)
    it = 0
    if fstop is not None and f <= fstop:
        status = "fstop"
        maxiter = 0
    while it < maxiter:
        free = torch.ones_like(x, dtype=torch.bool)
        if lower is not None:
            free &= ~((x <= lower) & (g > 0))
        if upper is not None:
            free &= ~((x >= upper) & (g < 0))
        pg = torch.where(free, g, torch.zeros_like(g))
        gnorm = math.sqrt(dot(pg, pg))
        if g0norm is None:
            g0norm = gnorm
        if gnorm == 0.0 or gnorm <= max(gatol, grtol * g0norm):
            status = "gradient"
            break
        # two-loop recursion on the free variables
        q = pg.clone()
        alphas = []
        for s, y, rho in reversed(pairs):
            a = rho * dot(torch.where(free, s, 0.0), q)
            q.sub_(torch.where(free, y, 0.0), alpha=a)
            alphas.append(a)
        if pairs:
            s, y, _ = pairs[-1]
            q.mul_(dot(s, y) / dot(y, y))
        for (s, y, rho), a in zip(pairs, reversed(alphas)):
            b = rho * dot(torch.where(free, y, 0.0), q)
            q.add_(torch.where(free, s, 0.0), alpha=a - b)
        d = torch.where(free, -q, torch.zeros_like(q))
        gd = dot(g, d)
        if not (gd < 0.0):
            # not a descent direction: restart with steepest descent
            pairs.clear()
            d = -pg
            gd = -gnorm * gnorm
        step = 1.0 if pairs else 1.0 / gnorm
        # Armijo backtracking on the projected path
        accepted = False
        for _ in range(30):
            torch.add(x, d, alpha=step, out=xn)
            _project(xn, lower, upper)
            fnew = fg(xn, gn)
            neval += 1
            decrease = dot(g, xn - x)
            if fnew <= f + ftol * decrease:
                accepted = True
                break
            # safeguarded quadratic interpolation of the step
            denom = 2.0 * (fnew - f - gd * step)
            t = -gd * step * step / denom if denom > 0 else 0.5 * step
            step = min(max(t, 0.1 * step), 0.5 * step)
            if maxeval is not None and neval >= maxeval:
                break
        if not accepted:
            status = "linesearch"
            break
        s = xn - x
        y = gn - g
        sy = dot(s, y)
        if sy > 1e-12 * math.sqrt(dot(s, s) * dot(y, y)):
            pairs.append((s, y, 1.0 / sy))
            if len(pairs) > mem:
                pairs.pop(0)
        fprev = f
        x, xn = xn, x
        g, gn = gn, g
        f = fnew
        history.append(f)
        it += 1
        if callback is not None:
            callback(it, x, f)
        if checkpoint is not None and it % checkpoint_every == 0:
            save_checkpoint(checkpoint, x, pairs, it, f, g0norm)
        if frtol > 0.0 and abs(fprev - f) <= frtol * max(abs(f), abs(fprev)):
            status = "frtol"
            break
        if fstop is not None and f <= fstop:
            status = "fstop"
            break
        if maxeval is not None and neval >= maxeval:
            status = "maxeval"
            break
    return Result(x=x, f=f, niter=it, neval=neval, status=status, history=history, pairs=pairs,
                  grad=g, g0norm=g0norm)
