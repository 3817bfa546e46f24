"""Learned warm start of the VMLM-B reconstruction.

Problem class.  The acquisition of the 2-D fan-beam configuration
(source-to-isocentre 1000 mm, source-to-detector 1500 mm, Poisson noise at
1e5 incident photons, weighted least squares + TV on the B-spline samples,
positivity, VMLM-B) on a reduced 3-D cone-beam geometry, so that thousands
of 50-iteration solves fit on one GPU: 64^3 voxels of 1 mm on the cubic
B-spline basis, 120 views over 360 deg, a 96 x 96 flat detector at 1.5 mm.
Objects are :func:`phantoms.body_phantom` (seeds 10000 + i for training,
12000 + j held out).  Counts are I0 exp(-0.02 p) with p the projector output
(the footprint integrated over each 1.5 mm detector pixel, about 2.36 times
the line integral in mm here), i.e. about 0.047 mm^-1 per object unit.

Pairs.  For each training object the plain method runs 50 VMLM-B
iterations from zero; the iterate after 5 iterations is the network input
and the 50th the target.  Every solve, the network forward and backward, and
the optimiser run on the GPU (projectors, TV and U-Net kernels of
libtrtt_hip.so).

Use.  ``learned_solve`` runs 5 plain iterations, applies the network
(``out = x5 + U(x5, g5)``, projected on x >= 0, where g5 is the objective
gradient at x5 scaled by a global factor stored with the weights; the
image-only variant sees x5 alone), and keeps that start only if it
lowers the objective below f(x5) -- otherwise the plain iterate is the start
(the plain solve is the fallback); VMLM-B then continues until the objective
reaches the plain 50-iteration value within ``tol`` relative, or the
iteration budget ends.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import asdict, dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# default network: input channels (x5, objective gradient at x5)
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "warmstart_unet_cone64_grad.npz")
# image-only network: input x5
IMAGE_WEIGHTS = os.path.join(HERE, "weights", "warmstart_unet_cone64.npz")
TRAIN_SEED0, EVAL_SEED0 = 10000, 12000


@dataclass
class Setup:
    n: int = 64
    nviews: int = 120
    ndet: int = 96
    pitch: float = 1.5
    rsc: float = 1000.0
    rsd: float = 1500.0
    photons: float = 1e5
    mu_scale: float = 0.02
    tv_weight: float = 5.0
    tv_eps: float = 0.05
    n_first: int = 5
    n_plain: int = 50
    tol: float = 1e-3


def _sync():
    import torch

    torch.cuda.synchronize()
    return time.perf_counter()


class Reconstruction:
    """One projector, data term and TV shared by every object of the class;
    the data and weights of an object are copied into fixed device buffers so
    the data term's HIP graph stays valid."""

    def __init__(self, setup: Setup = Setup(), device="cuda", graph=True):
        import torch

        from .geometry import cone_geometry
        from .projectors import Projector
        from .reconstruction import DataTerm, Objective, TotalVariation

        self.torch = torch
        self.s = setup
        self.device = torch.device(device)
        theta = np.arange(setup.nviews) * (2 * math.pi / setup.nviews)
        n = setup.n
        geo = cone_geometry(n, n, n, 1.0, setup.ndet, setup.ndet, setup.pitch, setup.pitch, theta,
                            setup.rsc, setup.rsd)
        self.P = Projector(geo, model="cubic", precision="f32")
        f32 = dict(dtype=torch.float32, device=self.device)
        self.y = torch.zeros(self.P.sino_shape, **f32)
        self.w = torch.ones(self.P.sino_shape, **f32)
        self.data = DataTerm(self.P, self.y, self.w)
        self.tv = TotalVariation((n, n, n), setup.tv_weight, setup.tv_eps, samples=True,
                                 dtype=torch.float32, device=self.device)
        self.obj = Objective(self.data, [self.tv])
        if graph:
            self.data.capture(torch.zeros((n, n, n), **f32), torch.zeros((n, n, n), **f32))
        self.gen = torch.Generator(device=self.device)
        self._g = torch.empty((n, n, n), **f32)

    def simulate(self, truth: np.ndarray, seed: int):
        """Noisy data and weights of an object (seeded Poisson counts)."""
        from .phantoms import transmission_noise

        torch = self.torch
        xt = torch.from_numpy(np.ascontiguousarray(truth, dtype=np.float32)).to(self.device)
        p = self.P.forward(xt)
        self.gen.manual_seed(int(seed))
        y, w = transmission_noise(p, self.s.photons, self.s.mu_scale, generator=self.gen)
        self.y.copy_(y)
        self.w.copy_(w)

    def objective(self, x) -> float:
        return self.obj(x, self._g)

    def zeros(self):
        n = self.s.n
        return self.torch.zeros((n, n, n), dtype=self.torch.float32, device=self.device)

    def solve(self, x0, maxiter, fstop=None, callback=None, pairs=None):
        from .vmlmb import vmlmb

        return vmlmb(self.obj, x0, lower=0.0, maxiter=maxiter, fstop=fstop, callback=callback,
                     pairs=pairs)


def make_pairs(rec: Reconstruction, seeds, log=print, grad_input=False):
    """(x5, x50) of the plain method for every seed, on the device.  With
    ``grad_input`` the input also holds the objective gradient at x5 as a
    second channel (raw; scaled by :func:`grad_scale`).  The 50 iterations run
    as 5 + 45 with the L-BFGS memory carried over, which is the same iteration
    sequence."""
    from .phantoms import body_phantom

    torch, s = rec.torch, rec.s
    shape = (len(seeds), s.n, s.n, s.n) + ((2,) if grad_input else ())
    X5 = torch.empty(shape, dtype=torch.float32, device=rec.device)
    X50 = torch.empty((len(seeds), s.n, s.n, s.n), dtype=torch.float32, device=rec.device)
    t0 = time.perf_counter()
    for i, seed in enumerate(seeds):
        rec.simulate(body_phantom(s.n, seed), seed)
        r5 = rec.solve(rec.zeros(), s.n_first)
        if grad_input:
            X5[i, ..., 0].copy_(r5.x)
            X5[i, ..., 1].copy_(r5.grad)
        else:
            X5[i].copy_(r5.x)
        r = rec.solve(r5.x, s.n_plain - r5.niter, pairs=r5.pairs)
        X50[i].copy_(r.x)
        if (i + 1) % 50 == 0 or i + 1 == len(seeds):
            log(f"pairs {i + 1}/{len(seeds)}  {time.perf_counter() - t0:.1f} s  "
                f"(last: {r5.niter + r.niter} it, {r.status})")
    return X5, X50


def grad_scale(X5) -> float:
    """One global factor bringing the gradient channel to unit RMS over the
    training inputs (stored with the weights)."""
    g = X5[..., 1]
    return float(1.0 / max(g.pow(2).mean().sqrt().item(), 1e-30))


def net_input(net, x5, g5):
    """The network input for an iterate and its gradient."""
    if net.in_ch == 1:
        return x5
    return net.torch.stack([x5, g5 * float(net.meta["grad_scale"])], dim=-1)


def train(net, X5, X50, epochs, batch=4, lr=1e-3, seed=0, log=print):
    """Adam with a cosine learning-rate schedule on the MSE to the targets."""
    torch = net.torch
    rng = np.random.default_rng(seed)
    N = X5.shape[0]
    nb = N // batch
    total = epochs * nb
    k = 0
    t0 = time.perf_counter()
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(N)
        run = 0.0
        for b in range(nb):
            idx = torch.from_numpy(perm[b * batch:(b + 1) * batch]).to(X5.device)
            lr_k = 0.5 * lr * (1.0 + math.cos(math.pi * k / total))
            run += net.train_step(X5[idx], X50[idx], lr=lr_k)
            k += 1
        hist.append(run / nb)
        log(f"epoch {ep + 1}/{epochs}  loss {hist[-1]:.3e}  {time.perf_counter() - t0:.1f} s")
    return hist


def plain_solve(rec: Reconstruction):
    """n_plain iterations from zero with the objective and wall time of
    every iterate."""
    s = rec.s
    t0 = _sync()
    fh, th = [], []

    def cb(it, x, f):
        fh.append(f)
        th.append(time.perf_counter() - t0)

    r = rec.solve(rec.zeros(), s.n_plain, callback=cb)
    return r, fh, th, _sync() - t0


def learned_solve(rec: Reconstruction, net, fstop, budget=None, keep_memory=True):
    """5 plain iterations, the network, then VMLM-B to fstop.  With
    ``keep_memory`` the continued solve starts from the L-BFGS pairs of the
    first 5 iterations (curvature of the same objective) instead of an empty
    memory."""
    s = rec.s
    budget = s.n_plain if budget is None else budget
    t0 = _sync()
    r5 = rec.solve(rec.zeros(), s.n_first)
    xw = net(net_input(net, r5.x, r5.grad)).clamp_(min=0.0) if net is not None else None
    fw = rec.objective(xw) if xw is not None else math.inf
    use_net = fw < r5.f
    start = xw if use_net else r5.x
    r = rec.solve(start, budget - r5.niter, fstop=fstop,
                  pairs=r5.pairs if keep_memory else None)
    return r, dict(f5=r5.f, f_net=fw, used_net=bool(use_net), niter=r5.niter + r.niter,
                   time=_sync() - t0)


def evaluate(rec: Reconstruction, net, seeds, log=print, keep_memory=True):
    """Plain 50-iteration solve vs learned warm start on each held-out object."""
    from .phantoms import body_phantom

    s = rec.s
    torch = rec.torch
    rows = []
    # warm-up of every launch path outside the timed solves
    rec.simulate(body_phantom(s.n, seeds[0]), seeds[0])
    rw = rec.solve(rec.zeros(), 3)
    if net is not None:
        net(net_input(net, rw.x, rw.grad))
    for seed in seeds:
        truth = body_phantom(s.n, seed)
        tt = torch.from_numpy(truth).to(rec.device)
        rec.simulate(truth, seed)
        rp, fh, th, t_plain = plain_solve(rec)
        f50 = rp.f
        fstop = f50 + s.tol * abs(f50)
        k_tol = next((k for k, f in enumerate(fh) if f <= fstop), len(fh) - 1)
        rl, info = learned_solve(rec, net, fstop, keep_memory=keep_memory)
        rows.append(dict(
            seed=int(seed), f50=f50, plain_iters=rp.niter, plain_time=t_plain,
            plain_iters_to_tol=k_tol + 1, plain_time_to_tol=th[k_tol],
            learned_f=rl.f, learned_reached=bool(rl.f <= fstop), learned_iters=info["niter"],
            learned_time=info["time"], used_net=info["used_net"], f5=info["f5"],
            f_net=info["f_net"],
            rmse_plain=float(torch.sqrt(torch.mean((rp.x - tt) ** 2)).item()),
            rmse_learned=float(torch.sqrt(torch.mean((rl.x - tt) ** 2)).item())))
        r = rows[-1]
        log(f"seed {seed}: plain {r['plain_time']:.3f} s ({r['plain_iters_to_tol']} it to tol), "
            f"learned {r['learned_time']:.3f} s ({r['learned_iters']} it, reached "
            f"{r['learned_reached']}, net {r['used_net']}), rmse {r['rmse_plain']:.4f} / "
            f"{r['rmse_learned']:.4f}")
    return rows


def summarize(rows):
    a = lambda k: np.array([r[k] for r in rows], dtype=np.float64)  # noqa: E731
    return dict(
        n=len(rows),
        reached=int(a("learned_reached").sum()),
        used_net=int(a("used_net").sum()),
        plain_time_50=float(a("plain_time").mean()),
        plain_time_to_tol=float(a("plain_time_to_tol").mean()),
        learned_time=float(a("learned_time").mean()),
        time_ratio=float(a("learned_time").sum() / a("plain_time").sum()),
        plain_iters_to_tol=float(a("plain_iters_to_tol").mean()),
        learned_iters=float(a("learned_iters").mean()),
        rmse_plain=float(a("rmse_plain").mean()),
        rmse_learned=float(a("rmse_learned").mean()),
        rel_objective_gap=float(np.mean((a("learned_f") - a("f50")) / np.abs(a("f50")))),
    )


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("mode", choices=["train", "eval"])
    ap.add_argument("--n-train", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--n-eval", type=int, default=100)
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--out", default=None, help="write the trained weights here (train)")
    ap.add_argument("--report", default=None, help="JSON report of the evaluation")
    ap.add_argument("--fresh-memory", action="store_true",
                    help="restart VMLM-B after the network with an empty L-BFGS memory")
    ap.add_argument("--image-only", action="store_true",
                    help="network input x5 only (default: x5 and the objective gradient at x5)")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    import torch

    from .unet import UNet3D

    if not torch.cuda.is_available():
        sys.exit("the warm start runs on the GPU")
    setup = Setup()
    rec = Reconstruction(setup)
    report = dict(setup=asdict(setup))
    grad_input = not args.image_only
    if args.image_only and args.weights == DEFAULT_WEIGHTS:
        args.weights = IMAGE_WEIGHTS
    if args.mode == "train":
        t0 = time.perf_counter()
        X5, X50 = make_pairs(rec, list(range(TRAIN_SEED0, TRAIN_SEED0 + args.n_train)), log,
                             grad_input=grad_input)
        report["pairs_time"] = time.perf_counter() - t0
        net = UNet3D(seed=0, in_ch=2 if grad_input else 1)
        if grad_input:
            net.meta["grad_scale"] = grad_scale(X5)
            X5[..., 1] *= net.meta["grad_scale"]
        t0 = time.perf_counter()
        report["train_loss"] = train(net, X5, X50, args.epochs, args.batch, args.lr, log=log)
        report["train_time"] = time.perf_counter() - t0
        report.update(n_train=args.n_train, epochs=args.epochs, batch=args.batch, lr=args.lr)
        out = args.out or args.weights
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        net.save(out)
        log(f"weights -> {out}")
        del X5, X50
    else:
        net = UNet3D.load(args.weights)
    if args.n_eval > 0:
        rows = evaluate(rec, net, list(range(EVAL_SEED0, EVAL_SEED0 + args.n_eval)), log,
                        keep_memory=not args.fresh_memory)
        report["eval"] = summarize(rows)
        report["eval_rows"] = rows
        log(json.dumps(report["eval"]))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
