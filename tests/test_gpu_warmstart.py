"""Learned warm start against the plain VMLM-B solve on 100 held-out objects
(seeds 12000 + j, j < 100, never seen in training), with the committed
networks.

Measured here, per object: the plain 50-iteration objective f50, the
iterations and wall time the plain solve needs to come within 1e-3 of f50,
and the same for the warm start (5 plain iterations, the U-Net, VMLM-B to
the same objective).  The default network (input: x5 and the objective
gradient at x5) must reach the target on every object in fewer iterations,
at most half the plain 50-iteration GPU wall time, with a lower RMSE against
the true object; the image-only network must still gain over the plain
solve; a network that makes the objective worse must be rejected in favour
of the plain iterate (the fallback)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def rec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from yoricktrtt_amd.warmstart import Reconstruction, Setup

    return Reconstruction(Setup())


@pytest.mark.parametrize("which", ["default", "image_only"])
def test_trained_network_gains_over_plain(rec, which):
    from yoricktrtt_amd.unet import UNet3D
    from yoricktrtt_amd.warmstart import (DEFAULT_WEIGHTS, EVAL_SEED0, IMAGE_WEIGHTS, evaluate,
                                          summarize)

    _say("warm start vs plain on held-out objects:", which)
    net = UNet3D.load(DEFAULT_WEIGHTS if which == "default" else IMAGE_WEIGHTS)
    assert net.in_ch == (2 if which == "default" else 1)
    rows = evaluate(rec, net, [EVAL_SEED0 + j for j in range(100)], log=lambda m: _say(" ", m))
    s = summarize(rows)
    _say(" ", s)
    for r in rows:
        assert r["used_net"], r["seed"]
        assert r["learned_reached"], r["seed"]
        assert r["learned_iters"] < r["plain_iters_to_tol"], r["seed"]
        assert r["rmse_learned"] < r["rmse_plain"], r["seed"]
    # target: the plain 50-iteration objective within 1e-3 in at most half
    # the plain GPU wall time (default network); the image-only one gains less
    assert s["time_ratio"] <= (0.5 if which == "default" else 0.85)


def test_bad_network_falls_back_to_plain(rec):
    """A network with a large random head worsens the objective: the warm
    start keeps the plain 5-iteration iterate and the solve proceeds as the
    plain method would."""
    from yoricktrtt_amd.phantoms import body_phantom
    from yoricktrtt_amd.unet import UNet3D
    from yoricktrtt_amd.warmstart import EVAL_SEED0, learned_solve

    _say("fallback")
    net = UNet3D(seed=3)
    flat = net.flat()
    lay = net.L["head"]
    flat[lay.w_off:lay.w_off + 16] = np.random.default_rng(0).normal(0, 5.0, size=16)
    net.set_flat(flat)
    seed = EVAL_SEED0 + 50
    rec.simulate(body_phantom(rec.s.n, seed), seed)
    plain = rec.solve(rec.zeros(), 12)
    r, info = learned_solve(rec, net, fstop=None, budget=12)
    assert not info["used_net"] and info["f_net"] > info["f5"]
    # the fallback continues from the plain iterate with the plain memory
    assert info["niter"] == 12
    assert abs(r.f - plain.f) <= 1e-3 * abs(plain.f)
