"""VMLM-B: bound-constrained variable-metric limited-memory minimiser.

The reference drives its reconstructions with OptimPack's ``op_mnb`` /
``opk_minimize`` (Tomography_4D/Ytrtt4D_optim.i:933-941), which is not
available here.  This is this build's own implementation of the published
method (Thiebaut 2002, "Optimization issues in blind deconvolution
algorithms", SPIE 4847): an L-BFGS recursion restricted to the free
variables, a projected search path x(a) = P(x + a d), and on that path the
More-Thuente line search (More & Thuente 1994, "Line search algorithms
with guaranteed sufficient decrease", ACM TOMS 20:286-307) that OptimPack
uses, or Armijo backtracking (``linesearch="armijo"``).  The first step of
a run is scaled to unit length.

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


# ---------------------------------------------------------------------------
# More-Thuente line search
# ---------------------------------------------------------------------------
def _cubic_min(a, fa, ga, b, fb, gb):
    """Minimiser of the cubic that interpolates f and f' at a and b (the
    cubic step of the method), written with a scaled discriminant."""
    theta = 3.0 * (fa - fb) / (b - a) + ga + gb
    s = max(abs(theta), abs(ga), abs(gb))
    disc = (theta / s) ** 2 - (ga / s) * (gb / s)
    gamma = s * math.sqrt(max(disc, 0.0))
    if b < a:
        gamma = -gamma
    p = (gamma - ga) + theta
    q = ((gamma - ga) + gamma) + gb
    return a + (p / q) * (b - a) if q != 0.0 else None


def mt_step(stx, fx, gx, sty, fy, gy, stp, fp, gp, brackt, stpmin, stpmax):
    """One safeguarded update of the trial step (the four cases of More &
    Thuente section 4).  (stx, fx, gx) is the best step so far, (sty, fy, gy)
    the other end of the interval, (stp, fp, gp) the trial just evaluated.
    Returns the updated interval ends, the next trial step and whether a
    minimiser is bracketed."""
    sgnd = gp * math.copysign(1.0, gx)
    if fp > fx:
        # case 1: higher value, the minimiser lies between stx and stp
        stpc = _cubic_min(stx, fx, gx, stp, fp, gp)
        stpq = stx + ((gx / ((fx - fp) / (stp - stx) + gx)) / 2.0) * (stp - stx)
        if stpc is None:
            stpc = stpq
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0.0:
        # case 2: lower value, derivatives of opposite signs
        stpc = _cubic_min(stp, fp, gp, stx, fx, gx)
        stpq = stp + (gp / (gp - gx)) * (stx - stp)
        if stpc is None:
            stpc = stpq
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(gp) < abs(gx):
        # case 3: lower value, same sign, the derivative decreases in size;
        # the cubic may have no minimiser in the direction of the step
        theta = 3.0 * (fx - fp) / (stp - stx) + gx + gp
        s = max(abs(theta), abs(gx), abs(gp))
        gamma = s * math.sqrt(max(0.0, (theta / s) ** 2 - (gx / s) * (gp / s)))
        if stp > stx:
            gamma = -gamma
        p = (gamma - gp) + theta
        q = (gamma + (gx - gp)) + gamma
        r = p / q if q != 0.0 else 0.0
        if r < 0.0 and gamma != 0.0:
            stpc = stp + r * (stx - stp)
        else:
            stpc = stpmax if stp > stx else stpmin
        stpq = stp + (gp / (gp - gx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stpc - stp) < abs(stpq - stp) else stpq
            lim = stp + 0.66 * (sty - stp)
            stpf = min(lim, stpf) if stp > stx else max(lim, stpf)
        else:
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            stpf = min(max(stpf, stpmin), stpmax)
    else:
        # case 4: lower value, same sign, the derivative does not decrease
        if brackt:
            stpc = _cubic_min(stp, fp, gp, sty, fy, gy)
            stpf = stpc if stpc is not None else 0.5 * (stp + sty)
        else:
            stpf = stpmax if stp > stx else stpmin
    # the new interval keeps the best step at stx
    if fp > fx:
        sty, fy, gy = stp, fp, gp
    else:
        if sgnd < 0.0:
            sty, fy, gy = stx, fx, gx
        stx, fx, gx = stp, fp, gp
    return stx, fx, gx, sty, fy, gy, stpf, brackt


def more_thuente(phi, stp, f0, g0, ftol=1e-4, gtol=0.9, xtol=0.1, stpmin=0.0, stpmax=1e20,
                 maxfev=30, feps=1e-12):
    """Step along a descent direction satisfying the strong Wolfe conditions
    f(stp) <= f0 + ftol stp g0 and |f'(stp)| <= gtol |g0| (More & Thuente
    1994, section 2).  ``phi(stp)`` returns (f, f'); g0 < 0.  Near a
    minimiser f stops changing in floating point and the decrease can no
    longer be seen: a step whose value is within feps |f0| of f0 and that
    meets the curvature condition is accepted too (the approximate Wolfe
    conditions of Hager & Zhang 2005).  Returns (stp, f, f', evaluations,
    status) with status "converged", "xtol", "stpmax", "stpmin", "rounding"
    or "maxfev"; the step is the best one found in every case."""
    if not (g0 < 0.0):
        raise ValueError("not a descent direction")
    brackt = False
    stage1 = True
    gtest = ftol * g0
    width = stpmax - stpmin
    width1 = 2.0 * width
    stx, fx, gx = 0.0, f0, g0
    sty, fy, gy = 0.0, f0, g0
    stmin, stmax = 0.0, stp + 4.0 * stp
    nfev = 0
    best = (0.0, f0, g0)
    while True:
        f, g = phi(stp)
        nfev += 1
        ftest = f0 + stp * gtest
        if f <= ftest and (f < best[1] or best[0] == 0.0):
            best = (stp, f, g)
        if stage1 and f <= ftest and g >= 0.0:
            stage1 = False
        flat = abs(f - f0) <= feps * abs(f0)
        if (f <= ftest or flat) and abs(g) <= -gtol * g0:
            return stp, f, g, nfev, "converged"
        # warnings, the later ones taking precedence
        if stp == stpmin and (f > ftest or g >= gtest):
            return (*best, nfev, "stpmin")
        if stp == stpmax and f <= ftest and g <= gtest:
            return stp, f, g, nfev, "stpmax"
        if brackt and stmax - stmin <= xtol * stmax:
            return (*best, nfev, "xtol")
        if brackt and (stp <= stmin or stp >= stmax):
            return (*best, nfev, "rounding")
        if nfev >= maxfev:
            return (*best, nfev, "maxfev")
        if stage1 and f <= fx and f > ftest:
            # first stage: the step update works on the function minus the
            # sufficient-decrease line (section 3)
            fm, fxm, fym = f - stp * gtest, fx - stx * gtest, fy - sty * gtest
            gm, gxm, gym = g - gtest, gx - gtest, gy - gtest
            stx, fxm, gxm, sty, fym, gym, stp, brackt = mt_step(
                stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax)
            fx, fy = fxm + stx * gtest, fym + sty * gtest
            gx, gy = gxm + gtest, gym + gtest
        else:
            stx, fx, gx, sty, fy, gy, stp, brackt = mt_step(
                stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax)
        if brackt:
            # bisect when the interval does not shrink fast enough
            if abs(sty - stx) >= 0.66 * width1:
                stp = stx + 0.5 * (sty - stx)
            width1 = width
            width = abs(sty - stx)
            stmin, stmax = min(stx, sty), max(stx, sty)
        else:
            stmin = stp + 1.1 * (stp - stx)
            stmax = stp + 4.0 * (stp - stx)
        stp = min(max(stp, stpmin), stpmax)
        if brackt and (stp <= stmin or stp >= stmax or stmax - stmin <= xtol * stmax):
            stp = stx


def _project(x, lower, upper):
    if lower is not None:
        x.clamp_(min=lower)
    if upper is not None:
        x.clamp_(max=upper)
    return x


def vmlmb(fg, x0, lower=None, upper=None, mem=5, maxiter=50, maxeval=None, gatol=0.0,
          grtol=0.0, frtol=0.0, ftol=1e-4, gtol=0.9, xtol=0.1, linesearch="more-thuente",
          fstop=None, callback=None, pairs=None, g0norm=None, checkpoint=None,
          checkpoint_every=1):
    """Minimise f over lower <= x <= upper.

    The line search on the projected path x(a) = P(x + a d) is More-Thuente
    (strong Wolfe conditions with ``ftol``, ``gtol``, ``xtol``; the path's
    derivative sums g' d over the components the projection leaves free) or,
    with ``linesearch="armijo"``, backtracking to sufficient decrease.

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
        accepted = False
        if linesearch == "more-thuente":
            last = [None]

            def phi(a):
                torch.add(x, d, alpha=a, out=xn)
                _project(xn, lower, upper)
                fa = fg(xn, gn)
                # derivative of f(P(x + a d)): the components the projection
                # clamps do not move
                moving = torch.ones_like(x, dtype=torch.bool)
                xa = x + a * d
                if lower is not None:
                    moving &= xa > lower
                if upper is not None:
                    moving &= xa < upper
                last[0] = a
                return fa, dot(gn, torch.where(moving, d, torch.zeros_like(d)))

            fev_cap = 30 if maxeval is None else max(1, min(30, maxeval - neval))
            step, fnew, _, nfev, ls = more_thuente(phi, step, f, gd, ftol, gtol, xtol,
                                                   maxfev=fev_cap)
            neval += nfev
            accepted = step > 0.0 and (ls == "converged" or fnew < f)
            if accepted and last[0] != step:
                # the search returned an earlier step: evaluate it again
                torch.add(x, d, alpha=step, out=xn)
                _project(xn, lower, upper)
                fnew = fg(xn, gn)
                neval += 1
        else:
            # Armijo backtracking on the projected path
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
