"""GPU U-Net of the warm start (BF16 MFMA kernels) against the numpy
restatement in oracle/unet_ref.py.

Tolerances (stated per check):
* single layers: FP32 outputs of the matrix-core kernels (forward, data
  gradient, weight gradient) from BF16 operands vs FP64 numpy on the same BF16
  values: 1e-5 relative (max norm) -- FP32 accumulation order only;
* whole network: each layer recomputed from the GPU's own input differs only
  by 1-ulp BF16 rounding flips on < 0.1% of outputs; end to end, the output,
  the loss and every parameter gradient agree to 1e-2 relative (L2) with
  the numpy network that emulates BF16 rounding at the same points (the
  error against the unrounded FP64 network is printed);
* training: the MSE of a learnable map falls below half its start value."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _st():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _bf(a):
    return torch.tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).cuda()


@pytest.mark.parametrize("ks,cin,cout", [(3, 1, 16), (3, 16, 32), (3, 48, 16), (3, 96, 32),
                                         (3, 64, 64), (3, 32, 128), (1, 16, 1), (3, 3, 20)])
def test_conv_layer_vs_numpy(ks, cin, cout):
    from oracle import unet_ref as U
    from yoricktrtt_amd import _lib

    _say("conv layer", ks, cin, cout)
    lib = _lib.load()
    rng = np.random.default_rng(cin * 1000 + cout)
    n, d, h, w = 2, 5, 6, 21
    taps = ks ** 3
    kp = (taps * cin + 15) // 16 * 16
    kpt = (taps * cout + 15) // 16 * 16
    x = U.bf16(rng.normal(size=(n, d, h, w, cin)))
    wm = np.zeros((cout, kp), dtype=np.float32)
    wm[:, :taps * cin] = rng.normal(0, 0.2, size=(cout, taps * cin))
    b = rng.normal(size=cout).astype(np.float32)
    wd = torch.tensor(wm).cuda()
    wk = torch.empty(cout * kp, dtype=torch.bfloat16, device="cuda")
    wt = torch.empty(cin * kpt, dtype=torch.bfloat16, device="cuda")
    assert lib.trtt_unet_prep_weights(_vp(wd), _vp(wk), _vp(wt), taps, cin, cout, kp, kpt,
                                      _st()) == 0
    w3 = U.bf16(wm[:, :taps * cin]).reshape(cout, taps, cin)
    xd = _bf(x)
    y = torch.empty(n * d * h * w * cout, dtype=torch.bfloat16, device="cuda")
    y32 = torch.empty(n * d * h * w * cout, dtype=torch.float32, device="cuda")
    bd = torch.tensor(b).cuda()
    assert lib.trtt_unet_conv_fwd(_vp(xd), _vp(wk), _vp(bd), _vp(y), _vp(y32), n, d, h, w, cin,
                                  cout, ks, kp, 1, _st()) == 0
    ref = np.maximum(U.conv(x, w3, b.astype(np.float64), ks), 0.0)
    got = y32.cpu().numpy().reshape(ref.shape)
    assert rel_err(got, ref) < 1e-5
    assert np.array_equal(y.float().cpu().numpy().reshape(ref.shape), U.bf16(got))
    # data gradient: the same kernel on dY with the mirrored transposed weights
    dz = U.bf16(rng.normal(size=(n, d, h, w, cout)))
    dzd = _bf(dz)
    dx = torch.empty(n * d * h * w * cin, dtype=torch.bfloat16, device="cuda")
    dx32 = torch.empty(n * d * h * w * cin, dtype=torch.float32, device="cuda")
    assert lib.trtt_unet_conv_fwd(_vp(dzd), _vp(wt), None, _vp(dx), _vp(dx32), n, d, h, w, cout,
                                  cin, ks, kpt, 0, _st()) == 0
    assert rel_err(dx32.cpu().numpy().reshape(x.shape), U.conv_dgrad(dz, w3, ks)) < 1e-5
    # weight and bias gradients
    dw = torch.zeros(cout * kp, dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    assert lib.trtt_unet_conv_wgrad(_vp(xd), _vp(dzd), _vp(dw), _vp(db), n, d, h, w, cin, cout, ks,
                                    kp, _st()) == 0
    rdw, rdb = U.conv_wgrad(x, dz, ks)
    gdw = dw.cpu().numpy().reshape(cout, kp)
    assert np.all(gdw[:, taps * cin:] == 0.0)
    assert rel_err(gdw[:, :taps * cin].reshape(rdw.shape), rdw) < 1e-5
    assert rel_err(db.cpu().numpy(), rdb) < 1e-5


def test_pool_upsample_copy_exact():
    from oracle import unet_ref as U
    from yoricktrtt_amd import _lib

    _say("pool/upsample")
    lib = _lib.load()
    rng = np.random.default_rng(7)
    n, d, h, w, c = 2, 4, 6, 8, 5
    x = U.bf16(rng.normal(size=(n, d, h, w, c)))
    x[0, :2, :2, :2, 0] = 0.5  # ties: the first child wins
    xd = _bf(x)
    m = x.size // 8
    y = torch.empty(m, dtype=torch.bfloat16, device="cuda")
    arg = torch.empty(m, dtype=torch.uint8, device="cuda")
    assert lib.trtt_unet_pool(_vp(xd), _vp(y), _vp(arg), None, None, n, d, h, w, c, 0, _st()) == 0
    rp, rarg = U.pool(x)
    assert np.array_equal(y.float().cpu().numpy().reshape(rp.shape), rp)
    assert np.array_equal(arg.cpu().numpy().reshape(rarg.shape), rarg)
    dy = U.bf16(rng.normal(size=rp.shape))
    dx = torch.zeros(x.size, dtype=torch.float32, device="cuda")
    assert lib.trtt_unet_pool(None, None, _vp(arg), _vp(dx), _vp(_bf(dy)), n, d, h, w, c, 1,
                              _st()) == 0
    assert np.array_equal(dx.cpu().numpy().reshape(x.shape), U.pool_bwd(dy, rarg, x.shape))
    # upsample into a channel slice of a wider tensor, and back
    ctot, off = 9, 3
    coarse = U.bf16(rng.normal(size=(n, d // 2, h // 2, w // 2, c)))
    fine = torch.zeros(n * d * h * w * ctot, dtype=torch.bfloat16, device="cuda")
    assert lib.trtt_unet_upsample(_vp(_bf(coarse)), _vp(fine), n, d, h, w, c, ctot, off, 0,
                                  _st()) == 0
    f = fine.float().cpu().numpy().reshape(n, d, h, w, ctot)
    assert np.array_equal(f[..., off:off + c], U.upsample(coarse))
    assert np.all(f[..., :off] == 0) and np.all(f[..., off + c:] == 0)
    g = U.bf16(rng.normal(size=(n, d, h, w, ctot)) * 0.25)
    gc = torch.empty(coarse.size, dtype=torch.bfloat16, device="cuda")
    assert lib.trtt_unet_upsample(_vp(_bf(g)), _vp(gc), n, d, h, w, c, ctot, off, 1, _st()) == 0
    assert rel_err(gc.float().cpu().numpy().reshape(coarse.shape),
                   U.upsample_bwd(g[..., off:off + c])) < 4e-3  # one BF16 rounding
    sl = torch.empty(n * d * h * w * c, dtype=torch.bfloat16, device="cuda")
    assert lib.trtt_unet_channel_copy(_vp(_bf(g)), _vp(sl), n * d * h * w, c, ctot, off, 1,
                                      _st()) == 0
    assert np.array_equal(sl.float().cpu().numpy().reshape(n, d, h, w, c), g[..., off:off + c])


def _teacher_forced_layers(net, B, acts, U):
    """Each layer recomputed in numpy from the GPU's own BF16 input: the GPU
    output may differ only by 1-ulp rounding flips (FP32 accumulation order)
    on a tiny fraction of elements; pooling, upsampling and concatenation
    are exact."""
    g = {k: B[k].float().cpu().numpy().reshape(acts[k].shape) for k in acts
         if k in B and k not in ("arg1", "arg2")}
    g["xin"] = acts["xin"]
    chain = [("enc1a", "xin", "e1a"), ("enc1b", "e1a", "e1b"), ("enc2a", "p1", "e2a"),
             ("enc2b", "e2a", "e2b"), ("bota", "p2", "ba"), ("botb", "ba", "bb"),
             ("dec2a", "cat2", "d2a"), ("dec2b", "d2a", "d2b"), ("dec1a", "cat1", "d1a"),
             ("dec1b", "d1a", "d1b")]
    for name, src, dst in chain:
        w, b = net.layer_params(name)
        ref = U.bf16(np.maximum(U.conv(g[src], U.bf16(w), b, 3), 0.0))
        diff = np.abs(g[dst] - ref)
        nflip = int(np.sum(diff > 0))
        _say(f"  {name}: {nflip} of {ref.size} outputs differ (1-ulp rounding flips)")
        assert nflip <= max(4, ref.size // 1000), name
        assert np.all(diff <= 2.0 ** -7 * np.abs(ref) + 1e-30), name
    assert np.array_equal(g["p1"], U.pool(g["e1b"])[0])
    assert np.array_equal(g["p2"], U.pool(g["e2b"])[0])
    assert np.array_equal(g["cat2"], np.concatenate([g["e2b"], U.upsample(g["bb"])], -1))
    assert np.array_equal(g["cat1"], np.concatenate([g["e1b"], U.upsample(g["d2b"])], -1))


@pytest.mark.parametrize("in_ch", [1, 2])
def test_network_vs_numpy(in_ch):
    """Forward layer by layer (teacher forced) and end to end; loss and every
    parameter gradient against the BF16-emulating numpy network."""
    from oracle import unet_ref as U
    from yoricktrtt_amd.unet import UNet3D

    _say("network forward and gradients, input channels", in_ch)
    net = UNet3D(seed=5, in_ch=in_ch)
    rng = np.random.default_rng(6)
    # a non-trivial head so the whole network carries signal
    flat = net.flat()
    lay = net.L["head"]
    flat[lay.w_off:lay.w_off + 16] = rng.normal(0, 0.3, size=16)
    net.set_flat(flat)
    shape = (2, 8, 8, 20) if in_ch == 1 else (2, 8, 8, 20, in_ch)
    x = rng.uniform(0, 1, size=shape).astype(np.float32)
    x0 = x if in_ch == 1 else x[..., 0]
    t = (x0 + 0.1 * rng.normal(size=x0.shape)).astype(np.float32)
    xd = torch.tensor(x).cuda()
    params = {name: net.layer_params(name) for name in net.L}
    ref = U.UNetRef(params, net.channels, bf16_emulation=True, in_ch=in_ch)
    out = net(xd).cpu().numpy()
    rout, acts = ref.forward(x.astype(np.float64))
    _teacher_forced_layers(net, net._bufs[x.shape[:4]], acts, U)
    _say(f"  output rel L2 end to end {_l2(out, rout):.1e}")
    assert _l2(out, rout) < 1e-2
    loss = net.loss_and_grad(xd, torch.tensor(t).cuda())
    rloss, rg = ref.loss_and_grads(x.astype(np.float64), t.astype(np.float64))
    _, g64 = U.UNetRef(params, net.channels, in_ch=in_ch).loss_and_grads(
        x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - rloss) <= 1e-2 * rloss
    worst = 0.0
    for name in net.L:
        gw, gb = net.layer_params(name, source=net.grad)
        ew, eb = _l2(gw, rg[name][0]), _l2(gb, rg[name][1])
        worst = max(worst, ew, eb)
        _say(f"  {name}: grad rel L2 vs BF16-emulated {ew:.2e} / {eb:.2e}; vs FP64 "
             f"{_l2(gw, g64[name][0]):.2e}")
        assert ew < 1e-2 and eb < 1e-2, name
    _say("  worst", worst)


def test_training_reduces_loss():
    """Adam on the GPU learns a fixed smoothing map on random volumes."""
    from scipy import ndimage

    from yoricktrtt_amd.unet import UNet3D

    _say("training")
    rng = np.random.default_rng(8)
    net = UNet3D(seed=9)
    xs = rng.uniform(0, 1, size=(8, 16, 16, 16)).astype(np.float32)
    ts = np.stack([ndimage.uniform_filter(v, 3, mode="constant") for v in xs]).astype(np.float32)
    xd, td = torch.tensor(xs).cuda(), torch.tensor(ts).cuda()
    losses = [net.train_step(xd[i % 8:i % 8 + 4], td[i % 8:i % 8 + 4], lr=2e-3)
              for i in range(0, 120, 4)]
    _say("  loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.5 * losses[0]
    assert all(np.isfinite(losses))


def test_save_load_roundtrip(tmp_path):
    from yoricktrtt_amd.unet import UNet3D

    net = UNet3D(seed=11)
    x = torch.rand(1, 8, 8, 16, device="cuda")
    y0 = net(x)
    p = tmp_path / "w.npz"
    net.save(str(p))
    y1 = UNet3D.load(str(p))(x)
    assert torch.equal(y0, y1)
