"""GPU reconstruction layer: weighted LS data term of all views and phases
against the FP64 oracle, its HIP-graph replay, TV and the B-spline sampling
filter against numpy, and VMLM-B on the GPU against the same solver run on
the CPU with the oracle data term.

Tolerances: FP64 plans 1e-12 relative (max |diff| / max |ref|); FP32 1e-5;
TV finite differences 1e-6 relative; solver iterates 1e-10 relative."""
import math

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


def _problem(nt=1, V=10, seed=0):
    from yoricktrtt_amd.geometry import cone_geometry

    rng = np.random.default_rng(seed)
    nx, ny, nz, nu, nv = 20, 18, 14, 24, 28
    th = np.sort(rng.uniform(0, 2 * math.pi, size=V))
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 80.0, 120.0)
    t0 = rng.integers(0, nt, size=V).astype(np.int32)
    return geo, t0, rng


def _oracle_fg(oracle, geo, t0, nt, y, w, x):
    """CPU FP64 data term: f = 1/2 sum w (Ax - y)^2, g = A' w (Ax - y)."""
    nz, ny, nx = geo.nz, geo.ny, geo.nx
    f = 0.0
    g = np.zeros((nt, nz, ny, nx))
    for v in range(geo.nviews):
        ax = oracle.sb3d(np.zeros((geo.nv, geo.nu)), x[t0[v]], nx, ny, nz, 1.0, geo.nu, geo.nv, 1.0,
                         1.0, geo.xos[v], geo.xsd[v], True, 0)
        d = ax - y[v]
        r = w[v] * d
        f += 0.5 * float(np.sum(r * d))
        oracle.sb3d(g[t0[v]], r, nx, ny, nz, 1.0, geo.nu, geo.nv, 1.0, 1.0, geo.xos[v], geo.xsd[v],
                    True, 1)
    return f, g


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_data_term_vs_oracle(oracle, prec):
    from yoricktrtt_amd.projectors import Projector
    from yoricktrtt_amd.reconstruction import DataTerm

    _say("data term", prec)
    nt = 3
    geo, t0, rng = _problem(nt=nt, V=12, seed=1)
    x = rng.uniform(size=(nt, geo.nz, geo.ny, geo.nx))
    y = rng.uniform(size=(geo.nviews, geo.nv, geo.nu))
    w = rng.uniform(0.5, 2.0, size=y.shape)
    dt = torch.float64 if prec == "f64" else torch.float32
    P = Projector(geo, model="cubic", precision=prec, nt=nt, t0=t0)
    D = DataTerm(P, torch.tensor(y, dtype=dt, device="cuda"), torch.tensor(w, dtype=dt, device="cuda"))
    xd = torch.tensor(x, dtype=dt, device="cuda")
    gd = torch.empty_like(xd)
    f = D.eval(xd, gd)
    rf, rg = _oracle_fg(oracle, geo, t0, nt, y, w, x)
    tol = 1e-12 if prec == "f64" else 1e-5
    assert abs(f - rf) <= tol * abs(rf)
    assert rel_err(gd.double().cpu().numpy(), rg) < tol
    P.close()


def test_data_term_graph_replay():
    """One HIP graph per evaluation: every replay, with eager work in
    between, equals the eager evaluation."""
    from yoricktrtt_amd.projectors import Projector
    from yoricktrtt_amd.reconstruction import DataTerm

    _say("graph replay")
    nt = 2
    geo, t0, rng = _problem(nt=nt, V=9, seed=2)
    y = torch.tensor(rng.uniform(size=(geo.nviews, geo.nv, geo.nu)), dtype=torch.float32,
                     device="cuda")
    P = Projector(geo, model="cubic", precision="f32", nt=nt, t0=t0)
    D = DataTerm(P, y)
    xs = [torch.tensor(rng.uniform(size=(nt, geo.nz, geo.ny, geo.nx)), dtype=torch.float32,
                       device="cuda") for _ in range(3)]
    eager = []
    for x in xs:
        g = torch.empty_like(x)
        eager.append((D.eval(x, g), g.clone()))
    x_static = torch.zeros_like(xs[0])
    g_static = torch.zeros_like(xs[0])
    D.capture(x_static, g_static)
    for rep in range(2):
        for x, (fe, ge) in zip(xs, eager):
            x_static.copy_(x)
            f = D.eval(x_static, g_static)
            torch.cuda.synchronize()
            _ = (x * 2.0).sum().item()  # eager work between replays
            assert abs(f - fe) <= 1e-6 * abs(fe)
            assert rel_err(g_static.double().cpu().numpy(), ge.double().cpu().numpy()) < 1e-6
    P.close()


def _np_filter(x):
    from scipy import ndimage

    k1 = np.array([1 / 6, 2 / 3, 1 / 6])
    out = x.copy()
    for ax in (-1, -2, -3):
        out = ndimage.convolve1d(out, k1, axis=ax, mode="constant", cval=0.0)
    return out


def _np_tv(x, weight, eps, wt):
    nt = x.shape[0]
    d = []
    for ax in (-1, -2, -3):
        dd = np.diff(x, axis=ax, append=np.take(x, [-1], axis=ax))
        d.append(dd)
    if nt > 1:
        d.append(wt * (np.roll(x, -1, axis=0) - x))
    s = np.sqrt(eps * eps + sum(dd * dd for dd in d))
    return weight * float(np.sum(s - eps))


@pytest.mark.parametrize("nt", [1, 3])
def test_tv_and_filter(nt):
    from yoricktrtt_amd.reconstruction import TotalVariation

    _say("tv", nt)
    rng = np.random.default_rng(3)
    shape = (nt, 7, 9, 11)
    x = rng.normal(size=shape)
    xd = torch.tensor(x, device="cuda")
    # filter against numpy
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.reconstruction import _stream, _vp

    lib = _lib.load()
    out = torch.empty_like(xd)
    assert lib.trtt_hip_bspline_filter(_vp(xd), _vp(out), 11, 9, 7, nt, 0, _lib.F64,
                                       _stream(torch)) == 0
    ref = np.stack([_np_filter(x[t]) for t in range(nt)])
    assert rel_err(out.cpu().numpy(), ref) < 1e-13
    # TV value (samples) against numpy, gradient against finite differences
    tv = TotalVariation(shape if nt > 1 else shape[1:], weight=0.7, eps=0.05, wt=0.5,
                        samples=True, dtype=torch.float64, device="cuda")
    xin = xd if nt > 1 else xd[0]
    g = torch.zeros_like(xin)
    f = tv.eval(xin, g)
    xs = ref if nt > 1 else ref
    assert abs(f - _np_tv(xs, 0.7, 0.05, 0.5)) < 1e-12 * abs(f)
    gnp = g.cpu().numpy().reshape(-1)
    flat = xin.cpu().numpy().reshape(-1)
    idx = rng.choice(flat.size, size=12, replace=False)
    h = 1e-6
    for i in idx:
        xp, xm = flat.copy(), flat.copy()
        xp[i] += h
        xm[i] -= h
        gp = torch.zeros_like(xin)
        fp = tv.eval(torch.tensor(xp.reshape(xin.shape), device="cuda"), gp)
        fm = tv.eval(torch.tensor(xm.reshape(xin.shape), device="cuda"), gp)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - gnp[i]) <= 1e-6 * max(1.0, abs(gnp[i]))


def test_filter_single_slice_is_2d():
    """nz == 1: the sampling filter acts along x and y only."""
    from scipy import ndimage

    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.reconstruction import _stream, _vp

    _say("2-D filter")
    lib = _lib.load()
    x = np.random.default_rng(9).normal(size=(13, 17))
    xd = torch.tensor(x, device="cuda")
    out = torch.empty_like(xd)
    assert lib.trtt_hip_bspline_filter(_vp(xd), _vp(out), 17, 13, 1, 1, 0, _lib.F64,
                                       _stream(torch)) == 0
    k1 = np.array([1 / 6, 2 / 3, 1 / 6])
    ref = ndimage.convolve1d(ndimage.convolve1d(x, k1, axis=0, mode="constant"), k1, axis=1,
                             mode="constant")
    assert rel_err(out.cpu().numpy(), ref) < 1e-13


def test_vmlmb_gpu_matches_cpu_fp64(oracle):
    """Five VMLM-B iterations on the GPU (FP64 plan) and on the CPU (FP64
    oracle data term) from the same start: same iterates."""
    from yoricktrtt_amd.projectors import Projector
    from yoricktrtt_amd.reconstruction import DataTerm
    from yoricktrtt_amd.vmlmb import vmlmb

    _say("vmlmb gpu vs cpu")
    geo, t0, rng = _problem(nt=1, V=8, seed=4)
    t0[:] = 0
    xt = rng.uniform(size=(geo.nz, geo.ny, geo.nx))
    y = np.stack([oracle.sb3d(np.zeros((geo.nv, geo.nu)), xt, geo.nx, geo.ny, geo.nz, 1.0, geo.nu,
                              geo.nv, 1.0, 1.0, geo.xos[v], geo.xsd[v], True, 0)
                  for v in range(geo.nviews)])
    y += 0.01 * rng.normal(size=y.shape)
    w = np.ones_like(y)
    P = Projector(geo, model="cubic", precision="f64")
    D = DataTerm(P, torch.tensor(y, device="cuda"))

    def fg_cpu(x, g):
        f, gg = _oracle_fg(oracle, geo, t0, 1, y, w, x.numpy()[None])
        g.copy_(torch.tensor(gg[0]))
        return f

    x0 = np.zeros((geo.nz, geo.ny, geo.nx))
    rg = vmlmb(D.eval, torch.tensor(x0, device="cuda"), lower=0.0, maxiter=5)
    rc = vmlmb(fg_cpu, torch.tensor(x0), lower=0.0, maxiter=5)
    assert rg.niter == rc.niter == 5
    assert rel_err(rg.x.cpu().numpy(), rc.x.numpy()) < 1e-10
    assert abs(rg.f - rc.f) <= 1e-10 * abs(rc.f)
    assert rg.history[-1] < 0.2 * rg.history[0]
    P.close()


def test_autotune_candidates_and_cache(tmp_path):
    """Every launch shape reproduces the default projections (checked inside
    the tuner), the chosen shape is applied and cached, and a second call is
    a cache hit; out-of-range shapes are refused."""
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("autotune")
    th = np.arange(16) * (2 * math.pi / 16)
    geo = cone_geometry(40, 36, 70, 1.0, 80, 72, 1.0, 1.0, th, 90.0, 140.0)
    P = Projector(geo, model="cubic", precision="f32")
    assert P.path > 0
    cache = str(tmp_path / "tuning.json")
    r = AT.autotune(P, cache=cache, reps=1, windows=(8192, 16384), vpbs=(0, 2), chunks=(0, 3))
    assert not r["cached"] and AT.get_tuning(P) == r["tuning"]
    assert len(r["rows"]) == len(AT.FWD_TILES) * 2 * 2 * (1 + 2) + sum(len(w) for w in AT.BWD_WINDOWS.values())
    r2 = AT.autotune(P, cache=cache)
    assert r2["cached"] and r2["tuning"] == r["tuning"]
    with pytest.raises(RuntimeError):
        AT.set_tuning(P, fwd_tile=7)
    with pytest.raises(RuntimeError):
        AT.set_tuning(P, fwd_window_bytes=1000)
    # a tuned plan still matches the oracle-checked default on fresh inputs
    x = torch.rand(P.vol_shape, device="cuda")
    a = P.forward(x)
    AT.set_tuning(P, **r["default"])
    b = P.forward(x)
    assert rel_err(a.cpu().numpy(), b.cpu().numpy()) < 1e-5
    P.close()


def test_overlapped_reduction_eager_and_graph():
    """Views in several forward chunks: the reduction of one chunk runs on a
    second stream while the next is projected.  Same projections as one
    chunk, also when the evaluation is captured in a HIP graph and replayed."""
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.projectors import Projector
    from yoricktrtt_amd.reconstruction import DataTerm

    _say("overlapped reduction")
    nt = 2
    geo, t0, rng = _problem(nt=nt, V=11, seed=5)
    P = Projector(geo, model="cubic", precision="f32", nt=nt, t0=t0)
    assert P.path > 0
    x = torch.tensor(rng.uniform(size=(nt, geo.nz, geo.ny, geo.nx)), dtype=torch.float32,
                     device="cuda")
    one = P.forward(x).clone()
    # scratch windows and the reduction kernel: 6 chunks, two buffers, two streams
    AT.set_tuning(P, fwd_chunk_views=2, fwd_flush=0)
    for _ in range(3):
        many = P.forward(x)
        assert rel_err(many.cpu().numpy(), one.cpu().numpy()) < 1e-6
    y = torch.tensor(rng.uniform(size=(geo.nviews, geo.nv, geo.nu)), dtype=torch.float32,
                     device="cuda")
    D = DataTerm(P, y)
    g = torch.empty_like(x)
    fe = D.eval(x, g)
    ge = g.clone()
    D.capture(torch.zeros_like(x), torch.zeros_like(x))
    for rep in range(3):
        g2 = torch.empty_like(x)
        f = D.eval(x, g2)
        torch.cuda.synchronize()
        _ = (x * 3.0).sum().item()  # eager work between replays
        assert abs(f - fe) <= 1e-6 * abs(fe)
        assert rel_err(g2.cpu().numpy(), ge.cpu().numpy()) < 1e-6
    P.close()
