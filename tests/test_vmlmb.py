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


# ---------------------------------------------------------------------------
# More-Thuente line search: the test functions of More & Thuente (1994),
# section 5; from start steps 1e-3 .. 1e3 the search must return a step that
# satisfies the strong Wolfe conditions it guarantees.
# ---------------------------------------------------------------------------
def _mt_funcs():
    import math as m

    def f1(a, b=2.0):
        return -a / (a * a + b), (a * a - b) / (a * a + b) ** 2

    def f2(a, b=0.004):
        return (a + b) ** 5 - 2.0 * (a + b) ** 4, 5.0 * (a + b) ** 4 - 8.0 * (a + b) ** 3

    def f3(a, b=0.01, l=39):
        if a <= 1.0 - b:
            p0, d0 = 1.0 - a, -1.0
        elif a >= 1.0 + b:
            p0, d0 = a - 1.0, 1.0
        else:
            p0, d0 = (a - 1.0) ** 2 / (2.0 * b) + 0.5 * b, (a - 1.0) / b
        w = l * m.pi / 2.0
        return (p0 + 2.0 * (1.0 - b) / (l * m.pi) * m.sin(w * a),
                d0 + (1.0 - b) * m.cos(w * a))

    return [(f1, 1e-3, 0.1), (f2, 0.1, 0.1), (f3, 0.1, 0.1)]


@pytest.mark.parametrize("case", [0, 1, 2])
@pytest.mark.parametrize("a0", [1e-3, 1e-1, 1e1, 1e3])
def test_more_thuente_strong_wolfe(case, a0):
    from yoricktrtt_amd.vmlmb import more_thuente

    phi, ftol, gtol = _mt_funcs()[case]
    f0, g0 = phi(0.0)
    # the paper's experiments: a small interval tolerance, no flat acceptance
    stp, f, g, nfev, status = more_thuente(phi, a0, f0, g0, ftol=ftol, gtol=gtol, xtol=1e-10,
                                           feps=0.0)
    assert status == "converged", (status, stp, nfev)
    assert f <= f0 + ftol * stp * g0
    assert abs(g) <= gtol * abs(g0)
    assert (f, g) == phi(stp)
    assert nfev <= 30


def test_vmlmb_line_searches_agree_on_rosenbrock():
    """Both line searches reach the minimiser of the 2-D Rosenbrock function
    (1, 1); More-Thuente with fewer evaluations."""
    from yoricktrtt_amd.vmlmb import vmlmb

    def fg(x, g):
        a, b = x[0].item(), x[1].item()
        g[0] = -2.0 * (1.0 - a) - 400.0 * a * (b - a * a)
        g[1] = 200.0 * (b - a * a)
        return (1.0 - a) ** 2 + 100.0 * (b - a * a) ** 2

    runs = {}
    for ls in ("more-thuente", "armijo"):
        r = vmlmb(fg, torch.tensor([-1.2, 1.0], dtype=torch.float64), maxiter=2000, gatol=1e-10,
                  linesearch=ls)
        assert abs(r.x[0].item() - 1.0) < 1e-8 and abs(r.x[1].item() - 1.0) < 1e-8, ls
        runs[ls] = r.neval
    assert runs["more-thuente"] < runs["armijo"]
