"""Numpy restatement of the warm-start U-Net (test infrastructure only).

FP64 forward and backward of the 3-level 3-D U-Net used to warm-start the
reconstruction, written directly from the definitions (convolution as a sum
over taps of shifted inputs, its adjoint as the sum of shifted outputs), not
from the GPU formulation (implicit GEMM with mirrored transposed weights).
With ``bf16=True`` every tensor the GPU stores in BF16 is rounded to BF16 at
the same point, so the GPU gradients can be compared at BF16 accuracy.  Only
the tests import this module; the numpy backward itself is pinned by finite
differences in tests/test_unet_ref.py.

Layout: NDHWC arrays; weights w[cout][taps][cin] with tap t the offset
(t//9 - 1, (t//3)%3 - 1, t%3 - 1) along (z, y, x); bias b[cout].
"""
from __future__ import annotations

import numpy as np

TAPS3 = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_specs(channels=(16, 32, 64), in_ch=1):
    """(name, cin, cout, ks, relu) of every convolution, in forward order."""
    c1, c2, c3 = channels
    return [("enc1a", in_ch, c1, 3, True), ("enc1b", c1, c1, 3, True),
            ("enc2a", c1, c2, 3, True), ("enc2b", c2, c2, 3, True),
            ("bota", c2, c3, 3, True), ("botb", c3, c3, 3, True),
            ("dec2a", c2 + c3, c2, 3, True), ("dec2b", c2, c2, 3, True),
            ("dec1a", c1 + c2, c1, 3, True), ("dec1b", c1, c1, 3, True),
            ("head", c1, 1, 1, False)]


def _taps(ks):
    return TAPS3 if ks == 3 else [(0, 0, 0)]


def _shift(x, dz, dy, dx):
    """y[p] = x[p + (dz, dy, dx)] with zeros outside."""
    n, d, h, w, c = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    return xp[:, 1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :]


def conv(x, w, b, ks):
    out = np.zeros(x.shape[:4] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(_taps(ks)):
        out += _shift(x, dz, dy, dx) @ w[:, t, :].T
    return out + b


def conv_wgrad(x, dz, ks):
    cin, cout = x.shape[-1], dz.shape[-1]
    taps = _taps(ks)
    dw = np.zeros((cout, len(taps), cin))
    g = dz.reshape(-1, cout)
    for t, (a, b_, c) in enumerate(taps):
        dw[:, t, :] = g.T @ _shift(x, a, b_, c).reshape(-1, cin)
    return dw, g.sum(axis=0)


def conv_dgrad(dz, w, ks):
    """Adjoint of x -> conv(x, w, 0): dx[p] = sum_t dz[p - off_t] w_t."""
    n, d, h, wd, _ = dz.shape
    dxp = np.zeros((n, d + 2, h + 2, wd + 2, w.shape[2]))
    for t, (a, b_, c) in enumerate(_taps(ks)):
        dxp[:, 1 + a:1 + a + d, 1 + b_:1 + b_ + h, 1 + c:1 + c + wd, :] += dz @ w[:, t, :]
    return dxp[:, 1:-1, 1:-1, 1:-1, :]


def _children(x):
    """[..., 8, c] children of every 2x2x2 cell, child o = 4 bz + 2 by + bx."""
    n, d, h, w, c = x.shape
    r = x.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c)
    r = r.transpose(0, 1, 3, 5, 2, 4, 6, 7)
    return r.reshape(n, d // 2, h // 2, w // 2, 8, c)


def _from_children(r, shape):
    n, d, h, w, c = shape
    r = r.reshape(n, d // 2, h // 2, w // 2, 2, 2, 2, c).transpose(0, 1, 4, 2, 5, 3, 6, 7)
    return r.reshape(shape)


def pool(x):
    ch = _children(x)
    arg = np.argmax(ch, axis=4)  # first maximum
    return ch.max(axis=4), arg


def pool_bwd(dy, arg, shape):
    r = np.zeros(arg.shape[:4] + (8,) + arg.shape[4:])
    np.put_along_axis(r, arg[:, :, :, :, None, :], dy[:, :, :, :, None, :], axis=4)
    return _from_children(r, shape)


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def upsample_bwd(dy):
    return _children(dy).sum(axis=4)


class UNetRef:
    """Forward/backward of the U-Net with parameters {name: (w, b)}."""

    def __init__(self, params, channels=(16, 32, 64), bf16_emulation=False, in_ch=1):
        self.specs = layer_specs(channels, in_ch)
        self.p = params
        self.r = bf16 if bf16_emulation else (lambda a: np.asarray(a, dtype=np.float64))
        self.emulate = bf16_emulation

    def _w(self, name):
        w, b = self.p[name]
        return (bf16(w) if self.emulate else w), b

    def _conv(self, name, x, relu=True):
        spec = next(s for s in self.specs if s[0] == name)
        w, b = self._w(name)
        z = conv(x, w, b, spec[3])
        return self.r(np.maximum(z, 0.0) if relu else z)

    def forward(self, x):
        """x (n, d, h, w), or (n, d, h, w, c) with c input channels ->
        (out, cache); out = x[..., 0] + head(x)."""
        r = self.r
        xin = x[..., None] if x.ndim == 4 else x
        a = {"x": x, "xin": r(xin)}
        a["e1a"] = self._conv("enc1a", a["xin"])
        a["e1b"] = self._conv("enc1b", a["e1a"])
        a["p1"], a["arg1"] = pool(a["e1b"])
        a["e2a"] = self._conv("enc2a", a["p1"])
        a["e2b"] = self._conv("enc2b", a["e2a"])
        a["p2"], a["arg2"] = pool(a["e2b"])
        a["ba"] = self._conv("bota", a["p2"])
        a["bb"] = self._conv("botb", a["ba"])
        a["cat2"] = np.concatenate([a["e2b"], upsample(a["bb"])], axis=-1)
        a["d2a"] = self._conv("dec2a", a["cat2"])
        a["d2b"] = self._conv("dec2b", a["d2a"])
        a["cat1"] = np.concatenate([a["e1b"], upsample(a["d2b"])], axis=-1)
        a["d1a"] = self._conv("dec1a", a["cat1"])
        a["d1b"] = self._conv("dec1b", a["d1a"])
        w, b = self._w("head")
        head = conv(a["d1b"], w, b, 1)[..., 0]  # FP32 on the GPU, not rounded
        a["out"] = (x if x.ndim == 4 else x[..., 0]) + head
        return a["out"], a

    def loss_and_grads(self, x, target):
        """Mean squared error of forward(x) against target and the gradient of
        every parameter (FP64)."""
        r = self.r
        out, a = self.forward(x)
        diff = out - target
        loss = float(np.mean(diff * diff))
        g = {}
        dh = r(2.0 * diff / diff.size)[..., None]

        def back(name, inp, dz, need_dx=True):
            spec = next(s for s in self.specs if s[0] == name)
            w, _ = self._w(name)
            g[name] = conv_wgrad(inp, dz, spec[3])
            return r(conv_dgrad(dz, w, spec[3])) if need_dx else None

        def relu_b(gr, act):
            return r(np.where(act > 0, gr, 0.0))

        gd1b = back("head", a["d1b"], dh)
        gd1a = back("dec1b", a["d1a"], relu_b(gd1b, a["d1b"]))
        gcat1 = back("dec1a", a["cat1"], relu_b(gd1a, a["d1a"]))
        c1 = a["e1b"].shape[-1]
        ge1_skip, gd2b = gcat1[..., :c1], r(upsample_bwd(gcat1[..., c1:]))
        gd2a = back("dec2b", a["d2a"], relu_b(gd2b, a["d2b"]))
        gcat2 = back("dec2a", a["cat2"], relu_b(gd2a, a["d2a"]))
        c2 = a["e2b"].shape[-1]
        ge2_skip, gbb = gcat2[..., :c2], r(upsample_bwd(gcat2[..., c2:]))
        gba = back("botb", a["ba"], relu_b(gbb, a["bb"]))
        gp2 = back("bota", a["p2"], relu_b(gba, a["ba"]))
        ge2b = pool_bwd(gp2, a["arg2"], a["e2b"].shape) + ge2_skip
        ge2a = back("enc2b", a["e2a"], relu_b(ge2b, a["e2b"]))
        gp1 = back("enc2a", a["p1"], relu_b(ge2a, a["e2a"]))
        ge1b = pool_bwd(gp1, a["arg1"], a["e1b"].shape) + ge1_skip
        ge1a = back("enc1b", a["e1a"], relu_b(ge1b, a["e1b"]))
        back("enc1a", a["xin"], relu_b(ge1a, a["e1a"]), need_dx=False)
        return loss, g
