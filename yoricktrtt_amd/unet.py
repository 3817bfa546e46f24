"""3-D U-Net of the learned warm start, run by the BF16 MFMA kernels of
libtrtt_hip.so (trtt_unet_* entry points).

Network (3 levels, 16/32/64 channels by default, 3x3x3 convolutions, ReLU,
2x2x2 max pooling, nearest upsampling, skip connections by concatenation,
1x1x1 linear head), predicting a correction of its input (of input channel
0 when it has several):

    out = x + head(dec1(cat(enc1, up(dec2(cat(enc2, up(bottom)))))))

Activations and their gradients are BF16 (NDHWC), the products accumulate in
FP32 on the matrix cores, and the parameters, their gradients and the Adam
state are FP32.  Torch provides the device memory and the stream; every
operation of the forward, the backward and the optimiser step is a launch of
this project's kernels.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _pad16(k):
    return (k + 15) // 16 * 16


def layer_specs(channels=(16, 32, 64), in_ch=1):
    """(name, cin, cout, ks, relu) of every convolution, in forward order."""
    c1, c2, c3 = channels
    return [("enc1a", in_ch, c1, 3, True), ("enc1b", c1, c1, 3, True),
            ("enc2a", c1, c2, 3, True), ("enc2b", c2, c2, 3, True),
            ("bota", c2, c3, 3, True), ("botb", c3, c3, 3, True),
            ("dec2a", c2 + c3, c2, 3, True), ("dec2b", c2, c2, 3, True),
            ("dec1a", c1 + c2, c1, 3, True), ("dec1b", c1, c1, 3, True),
            ("head", c1, 1, 1, False)]


class _Layer:
    def __init__(self, name, cin, cout, ks, relu, offset):
        self.name, self.cin, self.cout, self.ks, self.relu = name, cin, cout, ks, relu
        self.taps = ks ** 3
        self.kp = _pad16(self.taps * cin)
        self.kpt = _pad16(self.taps * cout)
        self.w_off = offset
        self.b_off = offset + cout * self.kp
        self.size = cout * self.kp + cout


class UNet3D:
    """The warm-start network on one GPU."""

    def __init__(self, channels=(16, 32, 64), seed=0, device="cuda", in_ch=1):
        import torch

        self.torch = torch
        self.lib = _lib.load()
        self.channels = tuple(int(c) for c in channels)
        self.in_ch = int(in_ch)
        self.meta = {}  # extra scalars saved with the weights (e.g. input scales)
        self.device = torch.device(device)
        self.layers = []
        off = 0
        for spec in layer_specs(self.channels, self.in_ch):
            lay = _Layer(*spec, off)
            self.layers.append(lay)
            off += lay.size
        self.L = {lay.name: lay for lay in self.layers}
        self.nparams = off
        self.params = torch.zeros(off, dtype=torch.float32, device=self.device)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.nstep = 0
        self.wk = {lay.name: torch.empty(lay.cout * lay.kp, dtype=torch.bfloat16, device=self.device)
                   for lay in self.layers}
        self.wt = {lay.name: torch.empty(lay.cin * lay.kpt, dtype=torch.bfloat16, device=self.device)
                   for lay in self.layers if lay.name != "enc1a"}
        self.loss_dev = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._bufs = {}
        self.init_params(seed)

    # ---------------------------------------------------------------- params
    def init_params(self, seed):
        """He-normal weights for the ReLU layers, a small head (the network
        starts close to the identity map), zero biases."""
        rng = np.random.default_rng(seed)
        host = np.zeros(self.nparams, dtype=np.float32)
        for lay in self.layers:
            fan_in = lay.taps * lay.cin
            std = math.sqrt(2.0 / fan_in) if lay.relu else 1e-3
            w = np.zeros((lay.cout, lay.kp), dtype=np.float32)
            w[:, :fan_in] = rng.normal(0.0, std, size=(lay.cout, fan_in))
            host[lay.w_off:lay.b_off] = w.reshape(-1)
        self.set_flat(host)

    def set_flat(self, host):
        self.params.copy_(self.torch.from_numpy(np.ascontiguousarray(host, dtype=np.float32)))
        self.m.zero_()
        self.v.zero_()
        self.nstep = 0
        self.prepare()

    def flat(self):
        return self.params.cpu().numpy().copy()

    def layer_params(self, name, source=None):
        """(w[cout][taps][cin], b[cout]) of a layer as numpy FP64, from the
        parameters or from another flat buffer of the same layout (e.g. the
        gradient)."""
        lay = self.L[name]
        buf = (self.params if source is None else source).cpu().numpy().astype(np.float64)
        w = buf[lay.w_off:lay.b_off].reshape(lay.cout, lay.kp)[:, :lay.taps * lay.cin]
        return w.reshape(lay.cout, lay.taps, lay.cin), buf[lay.b_off:lay.b_off + lay.cout]

    def prepare(self):
        """BF16 operands of every layer from the FP32 master weights."""
        st = self._stream()
        for lay in self.layers:
            wt = self.wt.get(lay.name)
            r = self.lib.trtt_unet_prep_weights(_vp(self.params[lay.w_off:]), _vp(self.wk[lay.name]),
                                                _vp(wt), lay.taps, lay.cin, lay.cout, lay.kp,
                                                lay.kpt, st)
            self._check(r, "prep_weights")

    def save(self, path):
        np.savez(path, params=self.flat(), channels=np.array(self.channels),
                 in_ch=np.array(self.in_ch),
                 **{"meta_" + k: np.array(v, dtype=np.float64) for k, v in self.meta.items()})

    @classmethod
    def load(cls, path, device="cuda"):
        with np.load(path, allow_pickle=False) as z:
            in_ch = int(z["in_ch"]) if "in_ch" in z.files else 1
            net = cls(tuple(int(c) for c in z["channels"]), device=device, in_ch=in_ch)
            net.set_flat(z["params"])
            net.meta = {k[5:]: float(z[k]) for k in z.files if k.startswith("meta_")}
        return net

    # --------------------------------------------------------------- helpers
    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    @staticmethod
    def _check(r, what):
        if r != 0:
            raise RuntimeError(f"U-Net kernel launch failed: {what}")

    def _buffers(self, n, d, h, w):
        key = (n, d, h, w)
        if key in self._bufs:
            return self._bufs[key]
        if d % 4 or h % 4 or w % 4:
            raise ValueError("U-Net volume sizes must be multiples of 4")
        torch = self.torch
        c1, c2, c3 = self.channels
        P1 = n * d * h * w
        P2, P3 = P1 // 8, P1 // 64
        bf = dict(dtype=torch.bfloat16, device=self.device)
        B = {}
        sizes16 = {
            "xin": P1 * self.in_ch, "e1a": P1 * c1, "e1b": P1 * c1, "p1": P2 * c1, "e2a": P2 * c2,
            "e2b": P2 * c2, "p2": P3 * c2, "ba": P3 * c3, "bb": P3 * c3, "cat2": P2 * (c2 + c3),
            "d2a": P2 * c2, "d2b": P2 * c2, "cat1": P1 * (c1 + c2), "d1a": P1 * c1, "d1b": P1 * c1,
            "head16": P1, "dhead": P1,
            # gradients (reused where lifetimes do not overlap)
            "g1": P1 * c1, "z1": P1 * c1, "gcat1": P1 * (c1 + c2), "skip1": P1 * c1,
            "g2": P2 * c2, "z2": P2 * c2, "gcat2": P2 * (c2 + c3), "skip2": P2 * c2,
            "g3": P3 * c3, "z3": P3 * c3, "gp2": P3 * c2, "gp1": P2 * c1,
        }
        for k, s in sizes16.items():
            B[k] = torch.empty(s, **bf)
        B["arg1"] = torch.empty(P2 * c1, dtype=torch.uint8, device=self.device)
        B["arg2"] = torch.empty(P3 * c2, dtype=torch.uint8, device=self.device)
        f32 = dict(dtype=torch.float32, device=self.device)
        B["head32"] = torch.empty(P1, **f32)
        B["out"] = torch.empty(P1, **f32)
        B["pool1"] = torch.empty(P1 * c1, **f32)
        B["pool2"] = torch.empty(P2 * c2, **f32)
        self._bufs = {key: B}  # one shape at a time
        return B

    def _conv(self, name, x, y, dims, y32=None):
        lay = self.L[name]
        n, d, h, w = dims
        r = self.lib.trtt_unet_conv_fwd(_vp(x), _vp(self.wk[name]), _vp(self.params[lay.b_off:]),
                                        _vp(y), _vp(y32), n, d, h, w, lay.cin, lay.cout, lay.ks,
                                        lay.kp, int(lay.relu), self._stream())
        self._check(r, name)

    def _dgrad(self, name, dz, dx, dims):
        lay = self.L[name]
        n, d, h, w = dims
        r = self.lib.trtt_unet_conv_fwd(_vp(dz), _vp(self.wt[name]), None, _vp(dx), None, n, d, h,
                                        w, lay.cout, lay.cin, lay.ks, lay.kpt, 0, self._stream())
        self._check(r, name + " data gradient")

    def _wgrad(self, name, x, dz, dims):
        lay = self.L[name]
        n, d, h, w = dims
        r = self.lib.trtt_unet_conv_wgrad(_vp(x), _vp(dz), _vp(self.grad[lay.w_off:]),
                                          _vp(self.grad[lay.b_off:]), n, d, h, w, lay.cin, lay.cout,
                                          lay.ks, lay.kp, self._stream())
        self._check(r, name + " weight gradient")

    def _relu_bwd(self, g32, g16, act, out):
        r = self.lib.trtt_unet_relu_bwd(_vp(g32), _vp(g16), _vp(act), _vp(out), act.numel(),
                                        self._stream())
        self._check(r, "relu backward")

    # --------------------------------------------------------------- forward
    def _forward(self, x, target=None, want_grad=False):
        lib, st = self.lib, self._stream()
        if self.in_ch == 1 and x.dim() == 3:
            x = x.unsqueeze(0)
        if self.in_ch > 1 and x.dim() == 4:
            x = x.unsqueeze(0)
        x = x.contiguous()
        if x.dtype != self.torch.float32:
            raise TypeError("U-Net input must be float32")
        if (self.in_ch == 1 and x.dim() != 4) or (self.in_ch > 1 and (
                x.dim() != 5 or x.shape[-1] != self.in_ch)):
            raise ValueError(f"expected [n][d][h][w] input with {self.in_ch} channel(s) last")
        n, d, h, w = x.shape[:4]
        c1, c2, c3 = self.channels
        B = self._buffers(n, d, h, w)
        L1, L2, L3 = (n, d, h, w), (n, d // 2, h // 2, w // 2), (n, d // 4, h // 4, w // 4)
        P1 = n * d * h * w
        self._check(lib.trtt_unet_cast_bf16(_vp(x), _vp(B["xin"]), P1 * self.in_ch, st), "cast")
        # the network corrects channel 0
        base = x if self.in_ch == 1 else x[..., 0].contiguous()
        self._conv("enc1a", B["xin"], B["e1a"], L1)
        self._conv("enc1b", B["e1a"], B["e1b"], L1)
        self._check(lib.trtt_unet_pool(_vp(B["e1b"]), _vp(B["p1"]), _vp(B["arg1"]), None, None, *L1,
                                       c1, 0, st), "pool1")
        self._conv("enc2a", B["p1"], B["e2a"], L2)
        self._conv("enc2b", B["e2a"], B["e2b"], L2)
        self._check(lib.trtt_unet_pool(_vp(B["e2b"]), _vp(B["p2"]), _vp(B["arg2"]), None, None, *L2,
                                       c2, 0, st), "pool2")
        self._conv("bota", B["p2"], B["ba"], L3)
        self._conv("botb", B["ba"], B["bb"], L3)
        self._check(lib.trtt_unet_channel_copy(_vp(B["e2b"]), _vp(B["cat2"]), P1 // 8, c2, c2 + c3,
                                               0, 0, st), "cat2")
        self._check(lib.trtt_unet_upsample(_vp(B["bb"]), _vp(B["cat2"]), *L2, c3, c2 + c3, c2, 0,
                                           st), "up2")
        self._conv("dec2a", B["cat2"], B["d2a"], L2)
        self._conv("dec2b", B["d2a"], B["d2b"], L2)
        self._check(lib.trtt_unet_channel_copy(_vp(B["e1b"]), _vp(B["cat1"]), P1, c1, c1 + c2, 0, 0,
                                               st), "cat1")
        self._check(lib.trtt_unet_upsample(_vp(B["d2b"]), _vp(B["cat1"]), *L1, c2, c1 + c2, c1, 0,
                                           st), "up1")
        self._conv("dec1a", B["cat1"], B["d1a"], L1)
        self._conv("dec1b", B["d1a"], B["d1b"], L1)
        self._conv("head", B["d1b"], B["head16"], L1, y32=B["head32"])
        if target is not None:
            target = target.reshape(base.shape).contiguous()
            self.loss_dev.zero_()
        r = lib.trtt_unet_residual_mse(_vp(base), _vp(B["head32"]), _vp(target), _vp(B["out"]),
                                       _vp(B["dhead"]) if want_grad else None, P1,
                                       _vp(self.loss_dev) if target is not None else None, st)
        self._check(r, "residual")
        return B, (L1, L2, L3)

    def __call__(self, x):
        """Warm start of a volume (or a batch [n][d][h][w]) in FP32; with
        several input channels, x is [n][d][h][w][c] (or [d][h][w][c]) and the
        result has the shape of channel 0."""
        shape = x.shape if self.in_ch == 1 else x.shape[:-1]
        B, _ = self._forward(x)
        return B["out"].reshape(shape).clone()

    # -------------------------------------------------------------- backward
    def _backward(self, B, dims):
        lib, st = self.lib, self._stream()
        L1, L2, L3 = dims
        c1, c2, c3 = self.channels
        P1 = L1[0] * L1[1] * L1[2] * L1[3]
        self.grad.zero_()
        self._wgrad("head", B["d1b"], B["dhead"], L1)
        self._dgrad("head", B["dhead"], B["g1"], L1)
        self._relu_bwd(None, B["g1"], B["d1b"], B["z1"])
        self._wgrad("dec1b", B["d1a"], B["z1"], L1)
        self._dgrad("dec1b", B["z1"], B["g1"], L1)
        self._relu_bwd(None, B["g1"], B["d1a"], B["z1"])
        self._wgrad("dec1a", B["cat1"], B["z1"], L1)
        self._dgrad("dec1a", B["z1"], B["gcat1"], L1)
        self._check(lib.trtt_unet_channel_copy(_vp(B["gcat1"]), _vp(B["skip1"]), P1, c1, c1 + c2, 0,
                                               1, st), "cat1 backward")
        self._check(lib.trtt_unet_upsample(_vp(B["gcat1"]), _vp(B["g2"]), *L1, c2, c1 + c2, c1, 1,
                                           st), "up1 backward")
        self._relu_bwd(None, B["g2"], B["d2b"], B["z2"])
        self._wgrad("dec2b", B["d2a"], B["z2"], L2)
        self._dgrad("dec2b", B["z2"], B["g2"], L2)
        self._relu_bwd(None, B["g2"], B["d2a"], B["z2"])
        self._wgrad("dec2a", B["cat2"], B["z2"], L2)
        self._dgrad("dec2a", B["z2"], B["gcat2"], L2)
        self._check(lib.trtt_unet_channel_copy(_vp(B["gcat2"]), _vp(B["skip2"]), P1 // 8, c2,
                                               c2 + c3, 0, 1, st), "cat2 backward")
        self._check(lib.trtt_unet_upsample(_vp(B["gcat2"]), _vp(B["g3"]), *L2, c3, c2 + c3, c2, 1,
                                           st), "up2 backward")
        self._relu_bwd(None, B["g3"], B["bb"], B["z3"])
        self._wgrad("botb", B["ba"], B["z3"], L3)
        self._dgrad("botb", B["z3"], B["g3"], L3)
        self._relu_bwd(None, B["g3"], B["ba"], B["z3"])
        self._wgrad("bota", B["p2"], B["z3"], L3)
        self._dgrad("bota", B["z3"], B["gp2"], L3)
        B["pool2"].zero_()
        self._check(lib.trtt_unet_pool(None, None, _vp(B["arg2"]), _vp(B["pool2"]), _vp(B["gp2"]),
                                       *L2, c2, 1, st), "pool2 backward")
        self._relu_bwd(B["pool2"], B["skip2"], B["e2b"], B["z2"])
        self._wgrad("enc2b", B["e2a"], B["z2"], L2)
        self._dgrad("enc2b", B["z2"], B["g2"], L2)
        self._relu_bwd(None, B["g2"], B["e2a"], B["z2"])
        self._wgrad("enc2a", B["p1"], B["z2"], L2)
        self._dgrad("enc2a", B["z2"], B["gp1"], L2)
        B["pool1"].zero_()
        self._check(lib.trtt_unet_pool(None, None, _vp(B["arg1"]), _vp(B["pool1"]), _vp(B["gp1"]),
                                       *L1, c1, 1, st), "pool1 backward")
        self._relu_bwd(B["pool1"], B["skip1"], B["e1b"], B["z1"])
        self._wgrad("enc1b", B["e1a"], B["z1"], L1)
        self._dgrad("enc1b", B["z1"], B["g1"], L1)
        self._relu_bwd(None, B["g1"], B["e1a"], B["z1"])
        self._wgrad("enc1a", B["xin"], B["z1"], L1)

    def loss_and_grad(self, x, target):
        """Mean squared error of the network output against target; the
        parameter gradient is left in ``self.grad``.  Returns the loss."""
        B, dims = self._forward(x, target, want_grad=True)
        self._backward(B, dims)
        return float(self.loss_dev.item())

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.nstep += 1
        lr_t = lr * math.sqrt(1.0 - b2 ** self.nstep) / (1.0 - b1 ** self.nstep)
        r = self.lib.trtt_unet_adam(_vp(self.params), _vp(self.grad), _vp(self.m), _vp(self.v),
                                    self.nparams, lr_t, b1, b2, eps, self._stream())
        self._check(r, "adam")
        self.prepare()

    def train_step(self, x, target, lr=1e-3):
        loss = self.loss_and_grad(x, target)
        self.adam_step(lr)
        return loss
