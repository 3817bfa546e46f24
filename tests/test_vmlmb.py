"""VMLM-B (yoricktrtt_amd.vmlmb) on problems with known minimisers.
OptimPack is not available, so the solver is pinned by the optimality
conditions and closed forms, not against op_mnb (parity unpinned)."""
import numpy as np
import pytest
import torch

from yoricktrtt_amd.vmlmb import dot, vmlmb


def _quadratic(n, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.geomspace(1.0, cond, n)
    A = torch.tensor((q * ev) @ q.T)
    b = torch.tensor(rng.normal(size=n))

    def fg(x, g):
        Ax = A @ x
        g.copy_(Ax - b)
        return float(0.5 * dot(x, Ax) - dot(b, x))

    return A, b, fg


def test_unconstrained_quadratic_closed_form():
    A, b, fg = _quadratic(40, 1)
    res = vmlmb(fg, torch.zeros(40, dtype=torch.float64), mem=5, maxiter=400, grtol=1e-12)
    xs = torch.linalg.solve(A, b)
    assert torch.max(torch.abs(res.x - xs)).item() < 1e-8 * torch.max(torch.abs(xs)).item()


def test_bound_constrained_quadratic_kkt():
    A, b, fg = _quadratic(60, 2)
    res = vmlmb(fg, torch.ones(60, dtype=torch.float64), lower=0.0, mem=5, maxiter=600, grtol=1e-12)
    x = res.x
    g = A @ x - b
    assert torch.all(x >= 0)
    active = x <= 1e-12
    assert active.any() and (~active).any()  # the bound matters
    assert torch.max(torch.abs(g[~active])).item() < 1e-7
    assert torch.min(g[active]).item() > -1e-7


def test_rosenbrock():
    def fg(x, g):
        a, b = x[0].item(), x[1].item()
        g[0] = -2 * (1 - a) - 400 * a * (b - a * a)
        g[1] = 200 * (b - a * a)
        return (1 - a) ** 2 + 100 * (b - a * a) ** 2

    res = vmlmb(fg, torch.tensor([-1.2, 1.0], dtype=torch.float64), mem=5, maxiter=500, grtol=1e-10)
    assert np.allclose(res.x.numpy(), [1.0, 1.0], atol=1e-6)


def test_monotone_and_history():
    A, b, fg = _quadratic(30, 3)
    res = vmlmb(fg, torch.zeros(30, dtype=torch.float64), lower=-0.1, upper=0.3, maxiter=25)
    h = np.array(res.history)
    assert np.all(np.diff(h) <= 1e-12)
    assert res.niter == len(h) - 1


def test_checkpoint_resume_continues_the_same_iterates(tmp_path):
    """A run stopped after 7 iterations and resumed from its checkpoint ends
    on the same iterate as an uninterrupted run of 18 (checkpointing has no
    counterpart in the reference; OptimPack keeps no state on disk)."""
    from yoricktrtt_amd.vmlmb import load_checkpoint

    A, b, fg = _quadratic(50, 4)
    x0 = torch.full((50,), 0.2, dtype=torch.float64)
    full = vmlmb(fg, x0, lower=0.0, mem=5, maxiter=18, grtol=1e-3)
    ck = tmp_path / "state.npz"
    part = vmlmb(fg, x0, lower=0.0, mem=5, maxiter=7, grtol=1e-3, checkpoint=ck,
                 checkpoint_every=1)
    assert part.niter == 7
    x, pairs, niter, f, g0 = load_checkpoint(ck)
    assert niter == 7 and f == part.f and len(pairs) == 5
    assert torch.equal(x, part.x)
    rest = vmlmb(fg, x, lower=0.0, mem=5, maxiter=18 - niter, grtol=1e-3, pairs=pairs, g0norm=g0)
    assert torch.equal(rest.x, full.x)
    assert rest.history[1:] == full.history[8:]
