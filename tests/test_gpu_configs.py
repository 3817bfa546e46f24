"""Parity at the full sizes of the five benchmark configurations.

Every case runs the HIP path at the configuration's own geometry and compares
it with the FP64 oracle (OpenMP over views), or with the reference C plugin
compiled from its sources when ``oracle/_ref`` is present:

  1  2-D parallel 256^2, 360 views / 180 deg, 367 bins, FP64: cubic and table,
     batched plan and per-view drop-in symbols, every view, 1e-12;
  2  2-D fan 512^2 x 0.5 mm, 720 views, 736 flat bins x 0.75 mm, Rsc 1000,
     Rsd 1500, FP32: cubic, table and Long & Fessler, every view, 1e-5;
  3  3-D cone 256^3, 360 views, 512 (v) x 384 (u) at 1 mm, FP32: the shipped
     launch shape and the default one; forward on 12 evenly spaced views of
     the 360-view plan, backward of all 360 views on four 4-slice slabs
     (the FP32 register accumulators run over all 360 views), the full
     360-view adjoint identity in FP64 sums, 1e-5 / 1e-6; the tabulated
     footprint on the same views and slabs, 1e-5;
  4  4-D cone, 10 phases x 128x128x64 x 2 mm, 516 views (30-60 per phase),
     256 x 192 at 2 mm: cost and gradient of the weighted data term over all
     phases against the oracle loop, eager and graph-replayed, 1e-5.

Parity measure: max |gpu - ref| / max |ref| per output.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# launch shape shipped in yoricktrtt_amd/tuning/tuning.json for configuration 3
_TUNING = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..",
                                      "yoricktrtt_amd", "tuning", "tuning.json")))
SHIPPED3 = next(v["tuning"] for k, v in _TUNING.items() if "cone:256x256x256x1:384x512:360" in k)
SLABS3 = [(0, 4), (84, 88), (170, 174), (252, 256)]


def _say(*a):
    print("[gpu-config]", *a, flush=True)


@pytest.fixture(scope="module")
def lib():
    from yoricktrtt_amd import _lib

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _lib.load()


def _p(a):
    from yoricktrtt_amd import _lib

    return a.ctypes.data_as(_lib.dp)


def _dev(a, dtype):
    return torch.tensor(a, dtype=dtype, device="cuda:0")


def _table(m=10000):
    from yoricktrtt_amd.geometry import footprint_table

    fp, cum, coord, step = footprint_table(m)
    return fp, cum, coord, step


# ---------------------------------------------------------------------------
# configuration 1: 2-D parallel beam, FP64
# ---------------------------------------------------------------------------

def _config1():
    from yoricktrtt_amd.geometry import parallel_geometry_2d
    from yoricktrtt_amd.phantoms import ellipses_2d

    n, nv, V = 256, 367, 360
    geo = parallel_geometry_2d(n, n, 1.0, nv, 1.0, np.arange(V) * (math.pi / V))
    img, _ = ellipses_2d(n, 10, seed=1)
    sino = np.random.default_rng(101).uniform(size=(V, nv))
    return geo, img, sino


@pytest.mark.parametrize("model", ["cubic", "table"])
def test_config1_full_fp64(lib, oracle, model):
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.projectors import Projector

    _say("config 1", model)
    geo, img, sino = _config1()
    n, nv, V = geo.nx, geo.nv, geo.nviews
    fp, cum, coord, step = _table()
    tab = (cum, coord, step) if model == "table" else None
    rf = oracle.views2d(img, n, n, 1.0, nv, 1.0, geo.xos, geo.xsd, False, 0, tab)
    rb = oracle.views2d(sino, n, n, 1.0, nv, 1.0, geo.xos, geo.xsd, False, 1, tab)
    # batched plan
    P = Projector(geo, model=model, precision="f64")
    fwd = P.forward(_dev(img, torch.float64)).cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float64)).cpu().numpy()
    assert P.error_flag() == 0
    P.close()
    ef, eb = rel_err(fwd, rf), rel_err(bwd, rb)
    _say(f"  plan fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < 1e-12 and eb < 1e-12
    # per-view drop-in symbols, FP64
    lib.trtt_hip_set_precision(_lib.F64)
    try:
        dfw = np.zeros((V, nv))
        dbw = np.zeros((n, n))
        for v in range(V):
            xos, xsd = np.ascontiguousarray(geo.xos[v]), np.ascontiguousarray(geo.xsd[v])
            if model == "table":
                r1 = lib.trtt_PB2D(_p(dfw[v]), _p(img), n, n, 1.0, 3, _p(fp), _p(cum), cum.size,
                                   step, _p(coord), nv, 1.0, _p(xos), _p(xsd), 0)
                r2 = lib.trtt_PB2D(_p(dbw), _p(np.ascontiguousarray(sino[v])), n, n, 1.0, 3,
                                   _p(fp), _p(cum), cum.size, step, _p(coord), nv, 1.0, _p(xos),
                                   _p(xsd), 1)
            else:
                r1 = lib.trtt_PB2D_cubic(_p(dfw[v]), _p(img), n, n, 1.0, 3, nv, 1.0, _p(xos),
                                         _p(xsd), 0)
                r2 = lib.trtt_PB2D_cubic(_p(dbw), _p(np.ascontiguousarray(sino[v])), n, n, 1.0,
                                         3, nv, 1.0, _p(xos), _p(xsd), 1)
            assert r1 == 0 and r2 == 0, v
    finally:
        lib.trtt_hip_set_precision(_lib.F32)
    ef, eb = rel_err(dfw, rf), rel_err(dbw, rb)
    _say(f"  drop-in fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < 1e-12 and eb < 1e-12


def test_config1_reference_plugin(lib):
    """The drop-in FP64 symbols against the reference C plugin itself on
    every view of configuration 1 (both footprint models)."""
    from oracle.oracle import Reference
    from yoricktrtt_amd import _lib

    if not Reference.available():
        pytest.skip("reference plugin not built (oracle/_ref)")
    ref = Reference()
    geo, img, sino = _config1()
    n, nv, V = geo.nx, geo.nv, geo.nviews
    fp, cum, coord, step = _table()
    _say("config 1 vs reference plugin")
    lib.trtt_hip_set_precision(_lib.F64)
    try:
        for name in ("cubic", "table"):
            got = np.zeros((V, nv))
            exp = np.zeros((V, nv))
            gb = np.zeros((n, n))
            eb = np.zeros((n, n))
            for v in range(V):
                xos, xsd = np.ascontiguousarray(geo.xos[v]), np.ascontiguousarray(geo.xsd[v])
                sv = np.ascontiguousarray(sino[v])
                if name == "cubic":
                    args = lambda o, i, job: (_p(o), _p(i), n, n, 1.0, 3, nv, 1.0, _p(xos),
                                              _p(xsd), job)
                    assert lib.trtt_PB2D_cubic(*args(got[v], img, 0)) == 0
                    assert ref.lib.trtt_PB2D_cubic(*args(exp[v], img, 0)) == 0
                    assert lib.trtt_PB2D_cubic(*args(gb, sv, 1)) == 0
                    assert ref.lib.trtt_PB2D_cubic(*args(eb, sv, 1)) == 0
                else:
                    args = lambda o, i, job: (_p(o), _p(i), n, n, 1.0, 3, _p(fp), _p(cum),
                                              cum.size, step, _p(coord), nv, 1.0, _p(xos),
                                              _p(xsd), job)
                    assert lib.trtt_PB2D(*args(got[v], img, 0)) == 0
                    assert ref.lib.trtt_PB2D(*args(exp[v], img, 0)) == 0
                    assert lib.trtt_PB2D(*args(gb, sv, 1)) == 0
                    assert ref.lib.trtt_PB2D(*args(eb, sv, 1)) == 0
            ef, ebk = rel_err(got, exp), rel_err(gb, eb)
            _say(f"  {name}: fwd {ef:.2e} bwd {ebk:.2e}")
            assert ef < 1e-12 and ebk < 1e-12
    finally:
        lib.trtt_hip_set_precision(_lib.F32)


# ---------------------------------------------------------------------------
# configuration 2: 2-D fan beam, FP32, and Long & Fessler
# ---------------------------------------------------------------------------

def _config2():
    from yoricktrtt_amd.geometry import fan_geometry_2d
    from yoricktrtt_amd.phantoms import ellipses_2d

    n, s, nv, vs, V, Rsc, Rsd = 512, 0.5, 736, 0.75, 720, 1000.0, 1500.0
    theta = np.arange(V) * (2 * math.pi / V)
    geo = fan_geometry_2d(n, n, s, nv, vs, theta, Rsc, Rsd)
    img, _ = ellipses_2d(n, 20, seed=2)
    sino = np.random.default_rng(202).uniform(size=(V, nv))
    return geo, theta, img, sino


@pytest.mark.parametrize("model", ["cubic", "table"])
def test_config2_full_fp32(lib, oracle, model):
    from yoricktrtt_amd.projectors import Projector

    _say("config 2", model)
    geo, _, img, sino = _config2()
    n, nv = geo.nx, geo.nv
    _, cum, coord, step = _table()
    tab = (cum, coord, step) if model == "table" else None
    rf = oracle.views2d(img, n, n, 0.5, nv, 0.75, geo.xos, geo.xsd, True, 0, tab)
    rb = oracle.views2d(sino, n, n, 0.5, nv, 0.75, geo.xos, geo.xsd, True, 1, tab)
    P = Projector(geo, model=model, precision="f32")
    fwd = P.forward(_dev(img, torch.float32)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    P.close()
    ef, eb = rel_err(fwd, rf), rel_err(bwd, rb)
    _say(f"  fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < 1e-5 and eb < 1e-5


def test_config2_long_fessler_full(lib, oracle):
    """Long & Fessler on configuration 2 (pixelSize = pixelStep = bin pitch,
    detectorDistance = Rsd - Rsc, beta = theta - pi/2): batched plan on every
    view and the drop-in symbols on every tenth view, FP32, against the
    oracle at 1e-5."""
    from oracle.oracle import LF2DParams
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.projectors import LFProjector

    _say("config 2 Long & Fessler")
    geo, theta, img, sino = _config2()
    n, nv, V = geo.nx, geo.nv, geo.nviews
    args = (0.75, 0.75, 0.5, 1000.0, 500.0, 0.0, 0.0, n, n, nv)
    beta = theta - 0.5 * math.pi
    off = np.zeros(V)
    po = LF2DParams(*args)
    rf = oracle.lf2d_views(po, beta, off, img, 0)
    rb = oracle.lf2d_views(po, beta, off, sino, 1)
    lf = LFProjector(_lib.LF2DParameters(*args), beta, off, precision="f32")
    fwd = lf.forward(_dev(img, torch.float32)).double().cpu().numpy()
    bwd = lf.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    lf.close()
    ef, eb = rel_err(fwd, rf), rel_err(bwd, rb)
    _say(f"  plan fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < 1e-5 and eb < 1e-5
    p = _lib.LF2DParameters(*args)
    sub = np.arange(0, V, 10)
    dfw = np.zeros((sub.size, nv))
    dbw = np.zeros((n, n))
    rbs = oracle.lf2d_views(po, beta[sub], off[sub], np.ascontiguousarray(sino[sub]), 1)
    for q, v in enumerate(sub):
        lib.lf2d_direct(C.byref(p), beta[v], 0.0, _p(img), _p(dfw[q]), 1)
        lib.lf2d_adjoint(C.byref(p), beta[v], 0.0, _p(np.ascontiguousarray(sino[v])), _p(dbw), 0)
    ef, eb = rel_err(dfw, rf[sub]), rel_err(dbw, rbs)
    _say(f"  drop-in fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < 1e-5 and eb < 1e-5


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_config2_long_fessler_line_integ(lib, oracle, prec):
    """The line-integral Long & Fessler model (lf2d_line_integ_direct /
    _adjoint) on configuration 2: batched plan on every view, drop-in symbols
    on every tenth view, against the oracle (1e-5 FP32, 1e-12 FP64)."""
    from oracle.oracle import LF2DParams
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.projectors import LFProjector

    _say("config 2 Long & Fessler, line integral", prec)
    tol = 1e-12 if prec == "f64" else 1e-5
    geo, theta, img, sino = _config2()
    n, nv, V = geo.nx, geo.nv, geo.nviews
    args = (0.75, 0.75, 0.5, 1000.0, 500.0, 0.0, 0.0, n, n, nv)
    beta = theta - 0.5 * math.pi
    off = np.zeros(V)
    po = LF2DParams(*args)
    rf = oracle.lf2d_views(po, beta, off, img, 0, line=True)
    rb = oracle.lf2d_views(po, beta, off, sino, 1, line=True)
    dt = torch.float64 if prec == "f64" else torch.float32
    lf = LFProjector(_lib.LF2DParameters(*args), beta, off, precision=prec, line_integ=True)
    fwd = lf.forward(_dev(img, dt)).double().cpu().numpy()
    bwd = lf.backward(_dev(sino, dt)).double().cpu().numpy()
    lf.close()
    ef, eb = rel_err(fwd, rf), rel_err(bwd, rb)
    _say(f"  plan fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < tol and eb < tol
    p = _lib.LF2DParameters(*args)
    sub = np.arange(3, V, 10)
    dfw = np.zeros((sub.size, nv))
    dbw = np.zeros((n, n))
    rbs = oracle.lf2d_views(po, beta[sub], off[sub], np.ascontiguousarray(sino[sub]), 1, line=True)
    lib.trtt_hip_set_precision(_lib.F64 if prec == "f64" else _lib.F32)
    try:
        for q, v in enumerate(sub):
            lib.lf2d_line_integ_direct(C.byref(p), beta[v], 0.0, _p(img), _p(dfw[q]), 1)
            lib.lf2d_line_integ_adjoint(C.byref(p), beta[v], 0.0,
                                        _p(np.ascontiguousarray(sino[v])), _p(dbw), 0)
    finally:
        lib.trtt_hip_set_precision(_lib.F32)
    ef, eb = rel_err(dfw, rf[sub]), rel_err(dbw, rbs)
    _say(f"  drop-in fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < tol and eb < tol


# ---------------------------------------------------------------------------
# configuration 3: 3-D cone beam 256^3, FP32
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def config3(lib):
    from yoricktrtt_amd.configs import cone_setup

    geo = cone_setup(256)
    rng = np.random.default_rng(303)
    vol = rng.uniform(size=(256, 256, 256)).astype(np.float32)
    sino = rng.uniform(size=(geo.nviews, geo.nv, geo.nu)).astype(np.float32)
    return geo, vol, sino


@pytest.fixture(scope="module")
def config3_ref(config3, oracle):
    """Oracle forward of 12 evenly spaced views and back projection of all
    360 views on the slabs of SLABS3 (cubic model)."""
    geo, vol, sino = config3
    views = np.arange(0, geo.nviews, geo.nviews // 12)
    v64 = vol.astype(np.float64)
    rf = oracle.views_fwd3d(v64, 256, 256, 256, 1.0, geo.nu, geo.nv, 1.0, 1.0, geo.xos[views],
                            geo.xsd[views], True)
    s64 = sino.astype(np.float64)
    rb = [oracle.views_bwd3d(s64, 256, 256, 256, 1.0, geo.nu, geo.nv, 1.0, 1.0, geo.xos, geo.xsd,
                             True, k0=a, k1=b) for a, b in SLABS3]
    return views, rf, rb


@pytest.mark.parametrize("shape", ["shipped", "default"])
def test_config3_full_vs_oracle(config3, config3_ref, shape):
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.projectors import Projector

    _say("config 3 cubic", shape)
    geo, vol, sino = config3
    views, rf, rb = config3_ref
    P = Projector(geo, model="cubic", precision="f32")
    assert P.path > 0
    if shape == "shipped":
        AT.set_tuning(P, **SHIPPED3)
        cur = AT.get_tuning(P)
        assert {k: cur[k] for k in SHIPPED3} == SHIPPED3
    fwd = P.forward(_dev(vol, torch.float32))[views].double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32))
    slabs = [bwd[a:b].double().cpu().numpy() for a, b in SLABS3]
    assert P.error_flag() == 0
    P.close()
    ef = max(rel_err(fwd[q], rf[q]) for q in range(views.size))
    eb = rel_err(np.concatenate(slabs), np.concatenate(rb))
    _say(f"  fwd (12 views) {ef:.2e}, bwd (360 views, 16 slices) {eb:.2e}")
    assert ef < 1e-5 and eb < 1e-5


def test_config3_adjoint_full(config3):
    """<A x, y> = <x, A' y> over all 360 views with the shipped launch
    shape, FP64 sums."""
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.projectors import Projector

    _say("config 3 adjoint identity")
    geo, vol, sino = config3
    P = Projector(geo, model="cubic", precision="f32")
    AT.set_tuning(P, **SHIPPED3)
    x = _dev(vol, torch.float32)
    y = _dev(sino, torch.float32)
    ax = P.forward(x)
    aty = P.backward(y)
    lhs = float(torch.sum(ax.double() * y.double()).item())
    rhs = float(torch.sum(x.double() * aty.double()).item())
    P.close()
    err = abs(lhs - rhs) / abs(lhs)
    _say(f"  <Ax,y> {lhs:.10e} <x,A'y> {rhs:.10e} rel {err:.2e}")
    assert err < 1e-6


def test_config3_table_model(config3, oracle):
    """Table footprint (10^5 entries, as the 3-D scripts build it) on the
    general kernels: 12 forward views, 360-view backward on two slabs."""
    from yoricktrtt_amd.projectors import Projector

    _say("config 3 table model")
    geo, vol, sino = config3
    _, cum, coord, step = _table(100000)
    tab = (cum, coord, step)
    views = np.arange(0, geo.nviews, geo.nviews // 12)
    slabs = SLABS3[::3]
    rf = oracle.views_fwd3d(vol.astype(np.float64), 256, 256, 256, 1.0, geo.nu, geo.nv, 1.0, 1.0,
                            geo.xos[views], geo.xsd[views], True, tab)
    rb = [oracle.views_bwd3d(sino.astype(np.float64), 256, 256, 256, 1.0, geo.nu, geo.nv, 1.0,
                             1.0, geo.xos, geo.xsd, True, tab, k0=a, k1=b) for a, b in slabs]
    P = Projector(geo, model="table", precision="f32", size_footprint=100000)
    fwd = P.forward(_dev(vol, torch.float32))[views].double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32))
    gb = [bwd[a:b].double().cpu().numpy() for a, b in slabs]
    P.close()
    ef = max(rel_err(fwd[q], rf[q]) for q in range(views.size))
    eb = rel_err(np.concatenate(gb), np.concatenate(rb))
    _say(f"  fwd {ef:.2e} bwd {eb:.2e}")
    assert ef < 1e-5 and eb < 1e-5


# ---------------------------------------------------------------------------
# configuration 4: 4-D cone beam, all phases in one plan
# ---------------------------------------------------------------------------

def test_config4_data_term_full(lib, oracle):
    """f = 1/2 sum_k (A_k x_p(k) - y_k)' W_k (A_k x_p(k) - y_k) and its
    gradient over 10 phases and 516 views, against the oracle loop of
    Ytrtt4D_optim.i:418-446 in FP64; then the same through the captured
    HIP graph, replayed twice."""
    from yoricktrtt_amd.configs import config4_setup, respiratory_series
    from yoricktrtt_amd.projectors import Projector
    from yoricktrtt_amd.reconstruction import DataTerm

    _say("config 4 data term")
    nt = 10
    geo, phase, counts = config4_setup(nt)
    assert geo.nviews == int(counts.sum())
    x = respiratory_series(nt)
    rng = np.random.default_rng(404)
    nz, ny, nx = x.shape[1:]
    nu, nv, V = geo.nu, geo.nv, geo.nviews
    # oracle loop
    ax = np.zeros((V, nv, nu))
    for p in range(nt):
        sel = np.nonzero(phase == p)[0]
        ax[sel] = oracle.views_fwd3d(x[p].astype(np.float64), nx, ny, nz, 2.0, nu, nv, 2.0, 2.0,
                                     geo.xos[sel], geo.xsd[sel], True)
    y = (ax * rng.uniform(0.5, 1.5, size=ax.shape)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=ax.shape).astype(np.float32)
    res = ax - y.astype(np.float64)
    r = w.astype(np.float64) * res
    f_ref = 0.5 * float(np.sum(r * res))
    g_ref = np.zeros((nt, nz, ny, nx))
    for p in range(nt):
        sel = np.nonzero(phase == p)[0]
        g_ref[p] = oracle.views_bwd3d(np.ascontiguousarray(r[sel]), nx, ny, nz, 2.0, nu, nv, 2.0,
                                      2.0, geo.xos[sel], geo.xsd[sel], True)
    P = Projector(geo, model="cubic", precision="f32", nt=nt, t0=phase, t1=phase,
                  w0=np.ones(V), w1=np.zeros(V))
    assert P.path > 0
    D = DataTerm(P, _dev(y, torch.float32), _dev(w, torch.float32))
    xd = _dev(x, torch.float32)
    g = torch.zeros_like(xd)
    f = D.eval(xd, g)
    ef, eg = abs(f - f_ref) / abs(f_ref), rel_err(g.double().cpu().numpy(), g_ref)
    _say(f"  eager f {f:.8e} (ref {f_ref:.8e}) rel {ef:.2e}, g {eg:.2e}")
    assert ef < 1e-5 and eg < 1e-5
    gs = torch.zeros_like(xd)
    D.capture(xd.clone(), gs)
    for rep in range(2):
        fr = D.eval(xd, g)
        torch.cuda.synchronize()
        er, egr = abs(fr - f_ref) / abs(f_ref), rel_err(g.double().cpu().numpy(), g_ref)
        _say(f"  replay {rep}: f rel {er:.2e}, g {egr:.2e}")
        assert er < 1e-5 and egr < 1e-5
    D.check()
    P.close()
