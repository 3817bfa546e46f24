"""Pins the numpy U-Net restatement (oracle/unet_ref.py) that the GPU U-Net
tests compare against: brute-force convolution, adjoint identities of every
backward operator, finite differences of the whole network, and BF16
rounding against torch's own conversion.  CPU only."""
import numpy as np
import pytest

from oracle import unet_ref as U


def _brute_conv(x, w, b):
    n, d, h, wd, cin = x.shape
    cout = w.shape[0]
    out = np.zeros((n, d, h, wd, cout))
    for i in range(n):
        for z in range(d):
            for y in range(h):
                for xx in range(wd):
                    for t, (dz, dy, dx) in enumerate(U.TAPS3):
                        zz, yy, x2 = z + dz, y + dy, xx + dx
                        if 0 <= zz < d and 0 <= yy < h and 0 <= x2 < wd:
                            out[i, z, y, xx] += w[:, t, :] @ x[i, zz, yy, x2]
    return out + b


def test_conv_brute_force():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 3, 4, 5, 3))
    w = rng.normal(size=(4, 27, 3))
    b = rng.normal(size=4)
    assert np.allclose(U.conv(x, w, b, 3), _brute_conv(x, w, b), rtol=1e-12, atol=1e-12)
    w1 = rng.normal(size=(2, 1, 3))
    assert np.allclose(U.conv(x, w1, 0.0, 1), x @ w1[:, 0, :].T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ks", [3, 1])
def test_conv_adjoints(ks):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 4, 3, 6, 5))
    taps = ks ** 3
    w = rng.normal(size=(7, taps, 5))
    y = rng.normal(size=(2, 4, 3, 6, 7))
    lhs = np.sum(U.conv(x, w, 0.0, ks) * y)
    assert abs(lhs - np.sum(x * U.conv_dgrad(y, w, ks))) < 1e-10 * abs(lhs)
    dw, db = U.conv_wgrad(x, y, ks)
    assert abs(lhs - np.sum(w * dw)) < 1e-10 * abs(lhs)
    assert np.allclose(db, y.reshape(-1, 7).sum(0))


def test_pool_upsample_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 4, 6, 8, 3))
    p, arg = U.pool(x)
    # the max and its first index, cell by cell
    assert p[1, 1, 2, 3, 2] == x[1, 2:4, 4:6, 6:8, 2].max()
    o = arg[1, 1, 2, 3, 2]
    assert x[1, 2 + (o >> 2), 4 + ((o >> 1) & 1), 6 + (o & 1), 2] == p[1, 1, 2, 3, 2]
    y = rng.normal(size=p.shape)
    # selection S: <S x, y> = <x, S' y>
    sel = np.sum(np.take_along_axis(U._children(x), arg[:, :, :, :, None, :], 4)[..., 0, :] * y)
    assert abs(sel - np.sum(x * U.pool_bwd(y, arg, x.shape))) < 1e-10 * abs(sel)
    c = rng.normal(size=(2, 2, 3, 4, 3))
    f = rng.normal(size=(2, 4, 6, 8, 3))
    assert abs(np.sum(U.upsample(c) * f) - np.sum(c * U.upsample_bwd(f))) < 1e-10
    assert U.upsample(c)[1, 3, 5, 7, 2] == c[1, 1, 2, 3, 2]


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-6, 6, size=4000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(U.bf16(v), ref)


def _random_params(specs, rng):
    return {name: (rng.normal(0.0, np.sqrt(2.0 / (ks ** 3 * cin)), size=(cout, ks ** 3, cin)),
                   rng.normal(0.0, 0.1, size=cout))
            for name, cin, cout, ks, _ in specs}


def test_network_gradient_finite_differences():
    """FP64 U-Net (no rounding): analytic gradients of the MSE against central
    differences on random entries of every layer."""
    rng = np.random.default_rng(4)
    ch = (2, 3, 4)
    params = _random_params(U.layer_specs(ch), rng)
    x = rng.normal(size=(1, 4, 4, 8))
    t = rng.normal(size=x.shape)
    net = U.UNetRef(params, ch)
    loss, g = net.loss_and_grads(x, t)
    h = 1e-6
    for name, (w, b) in params.items():
        gw, gb = g[name]
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            w0 = w[idx]
            w[idx] = w0 + h
            fp = net.loss_and_grads(x, t)[0]
            w[idx] = w0 - h
            fm = net.loss_and_grads(x, t)[0]
            w[idx] = w0
            fd = (fp - fm) / (2 * h)
            assert abs(fd - gw[idx]) <= 1e-6 * max(1e-3, abs(gw[idx])), (name, idx)
        j = int(rng.integers(0, b.size))
        b0 = b[j]
        b[j] = b0 + h
        fp = net.loss_and_grads(x, t)[0]
        b[j] = b0 - h
        fm = net.loss_and_grads(x, t)[0]
        b[j] = b0
        assert abs((fp - fm) / (2 * h) - gb[j]) <= 1e-6 * max(1e-3, abs(gb[j])), name
