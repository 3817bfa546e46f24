"""GPU parity of libtrtt_hip.so against the reference golden vectors and the
FP64 oracle.  Parity measure: max|gpu - ref| / max|ref| per output.

Tolerances (written per test):
  FP32 plans / drop-ins   1e-5   (the FP32 projection tolerance)
  FP64 plans / drop-ins   1e-12  (table mode reproduces the reference's FP64
                                  coordinates and floor() indices)
  adjoint <Ax,y> = <x,A'y> with FP64 dot products: 1e-6 (FP32), 1e-12 (FP64)
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import CASES2D, CASES3D, rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def lib():
    from yoricktrtt_amd import _lib

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _lib.load()


def _p(a):
    from yoricktrtt_amd import _lib

    return a.ctypes.data_as(_lib.dp)


@pytest.fixture(params=["f32", "f64"])
def precision(request, lib):
    from yoricktrtt_amd import _lib

    lib.trtt_hip_set_precision(_lib.F64 if request.param == "f64" else _lib.F32)
    yield request.param
    lib.trtt_hip_set_precision(_lib.F32)


TOL = {"f32": 1e-5, "f64": 1e-12}


@pytest.mark.parametrize("name", CASES3D)
@pytest.mark.parametrize("mode", ["cubic", "table"])
def test_dropin_3d_vs_reference(lib, golden3d, precision, name, mode):
    g = golden3d
    _say("dropin3d", name, mode, precision)
    vol = np.ascontiguousarray(g["vol"])
    nz, ny, nx = vol.shape
    nu, nv = (int(v) for v in g[f"{name}/dims"])
    us, vs = (float(v) for v in g[f"{name}/scl"])
    cone = name.startswith("cone")
    pre = "CB3D" if cone else "PB3D"
    cum, coord, step = g["cum"], g["coord"], float(g["step"])
    fp = np.zeros_like(cum)
    for v in range(g[f"{name}/xos"].shape[0]):
        xos = np.ascontiguousarray(g[f"{name}/xos"][v])
        xsd = np.ascontiguousarray(g[f"{name}/xsd"][v])
        for job in (0, 1):
            inp = vol if job == 0 else np.ascontiguousarray(g[f"{name}/sino_in"][v])
            out = np.zeros((nv, nu)) if job == 0 else np.zeros((nz, ny, nx))
            if mode == "cubic":
                r = getattr(lib, f"trtt_{pre}_cubic")(_p(out), _p(inp), nx, ny, nz, 1.0, 3, nu, nv, us,
                                                     vs, _p(xos), _p(xsd), job)
            else:
                r = getattr(lib, f"trtt_{pre}")(_p(out), _p(inp), nx, ny, nz, 1.0, 3, _p(fp), _p(cum),
                                               cum.size, step, _p(coord), nu, nv, us, vs, _p(xos),
                                               _p(xsd), job)
            assert r == 0
            ref = g[f"{name}/{mode}/{'fwd' if job == 0 else 'bwd'}"][v]
            assert rel_err(out, ref) < TOL[precision], (v, job)


@pytest.mark.parametrize("name", CASES2D)
@pytest.mark.parametrize("mode", ["cubic", "table"])
def test_dropin_2d_vs_reference(lib, golden2d, precision, name, mode):
    g = golden2d
    _say("dropin2d", name, mode, precision)
    img = np.ascontiguousarray(g["img"])
    ny, nx = img.shape
    nv, vs = int(g[f"{name}/nv"]), float(g[f"{name}/vs"])
    pre = "FB2D" if name.startswith("fan") else "PB2D"
    cum, coord, step = g["cum"], g["coord"], float(g["step"])
    fp = np.zeros_like(cum)
    for v in range(g[f"{name}/xos"].shape[0]):
        xos = np.ascontiguousarray(g[f"{name}/xos"][v])
        xsd = np.ascontiguousarray(g[f"{name}/xsd"][v])
        for job in (0, 1):
            inp = img if job == 0 else np.ascontiguousarray(g[f"{name}/sino_in"][v])
            out = np.zeros(nv) if job == 0 else np.zeros((ny, nx))
            if mode == "cubic":
                r = getattr(lib, f"trtt_{pre}_cubic")(_p(out), _p(inp), nx, ny, 1.0, 3, nv, vs,
                                                     _p(xos), _p(xsd), job)
            else:
                r = getattr(lib, f"trtt_{pre}")(_p(out), _p(inp), nx, ny, 1.0, 3, _p(fp), _p(cum),
                                               cum.size, step, _p(coord), nv, vs, _p(xos), _p(xsd),
                                               job)
            assert r == 0
            ref = g[f"{name}/{mode}/{'fwd' if job == 0 else 'bwd'}"][v]
            assert rel_err(out, ref) < TOL[precision], (v, job)
            if mode == "cubic" and job == 0:
                assert out[-1] == 0.0  # last bin quirk


@pytest.mark.parametrize("name", ["fine", "coarse", "edge"])
def test_dropin_dd_pb3d_vs_reference(lib, oracle, precision, name):
    """trtt_DD_PB3D: forward and the reference's transpose against its golden
    vectors; trtt_hip_DD_PB3D_adjoint against the oracle's exact adjoint."""
    import os

    from conftest import GOLDEN

    _say("dd_pb3d", name, precision)
    g = np.load(os.path.join(GOLDEN, "dd3d.npz"))
    vol = np.ascontiguousarray(g["vol"])
    nz, ny, nx = vol.shape
    nu, nv = (int(v) for v in g[f"{name}/dims"])
    us, vs = (float(v) for v in g[f"{name}/scl"])
    for v in range(g[f"{name}/xos"].shape[0]):
        xos = np.ascontiguousarray(g[f"{name}/xos"][v])
        xsd = np.ascontiguousarray(g[f"{name}/xsd"][v])
        y = np.ascontiguousarray(g[f"{name}/sino_in"][v])
        f = np.zeros((nv, nu))
        b = np.zeros((nz, ny, nx))
        a = np.zeros((nz, ny, nx))
        assert lib.trtt_DD_PB3D(_p(f), _p(vol), nx, ny, nz, 1.0, nu, nv, us, vs, _p(xos), _p(xsd),
                                0) == 0
        assert lib.trtt_DD_PB3D(_p(b), _p(y), nx, ny, nz, 1.0, nu, nv, us, vs, _p(xos), _p(xsd),
                                1) == 0
        assert lib.trtt_hip_DD_PB3D_adjoint(_p(a), _p(y), nx, ny, nz, 1.0, nu, nv, us, vs, _p(xos),
                                            _p(xsd)) == 0
        assert rel_err(f, g[f"{name}/fwd"][v]) < TOL[precision]
        assert rel_err(b, g[f"{name}/bwd"][v]) < TOL[precision]
        ra = oracle.dd_pb3d(np.zeros((nz, ny, nx)), y, nx, ny, nz, 1.0, nu, nv, us, vs, xos, xsd, 1,
                            adjoint=True)
        assert rel_err(a, ra) < TOL[precision]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dd_batched_vs_oracle(lib, oracle, prec):
    """Batched distance-driven projector over 9 views (pixels smaller than
    the voxels, so the reference transpose differs from the adjoint)."""
    from yoricktrtt_amd.geometry import parallel_geometry_3d
    from yoricktrtt_amd.projectors import DDProjector

    _say("dd batched", prec)
    nx, ny, nz, nu, nv, V = 21, 19, 15, 44, 50, 9
    th = np.arange(V) * (math.pi / V) + 0.2
    geo = parallel_geometry_3d(nx, ny, nz, 1.0, nu, nv, 0.55, 0.5, th)
    P = DDProjector(geo, precision=prec)
    rng = np.random.default_rng(9)
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    dt = torch.float64 if prec == "f64" else torch.float32
    fwd = P.forward(_dev(vol, dt)).double().cpu().numpy()
    adj = P.backward(_dev(sino, dt)).double().cpu().numpy()
    ref_t = P.backward(_dev(sino, dt), exact=False).double().cpu().numpy()
    ra = np.zeros((nz, ny, nx))
    rt = np.zeros((nz, ny, nx))
    for v in range(V):
        rf = oracle.dd_pb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, 0.55, 0.5, geo.xos[v],
                            geo.xsd[v], 0)
        assert rel_err(fwd[v], rf) < TOL[prec], v
        oracle.dd_pb3d(ra, sino[v], nx, ny, nz, 1.0, nu, nv, 0.55, 0.5, geo.xos[v], geo.xsd[v], 1,
                       adjoint=True)
        oracle.dd_pb3d(rt, sino[v], nx, ny, nz, 1.0, nu, nv, 0.55, 0.5, geo.xos[v], geo.xsd[v], 1)
    assert rel_err(adj, ra) < TOL[prec]
    assert rel_err(ref_t, rt) < TOL[prec]
    a, b = float((fwd * sino).sum()), float((vol * adj).sum())
    assert abs(a - b) <= (1e-12 if prec == "f64" else 1e-6) * abs(a)
    P.close()


def test_dropin_accumulates(lib, golden3d):
    """The drop-ins add into `out` like the reference (out += A x)."""
    g = golden3d
    _say("dropin accumulate")
    vol = np.ascontiguousarray(g["vol"])
    nz, ny, nx = vol.shape
    nu, nv = (int(v) for v in g["cone_a/dims"])
    xos = np.ascontiguousarray(g["cone_a/xos"][0])
    xsd = np.ascontiguousarray(g["cone_a/xsd"][0])
    out = np.full((nv, nu), 2.5)
    assert lib.trtt_CB3D_cubic(_p(out), _p(vol), nx, ny, nz, 1.0, 3, nu, nv, 1.0, 1.0, _p(xos),
                               _p(xsd), 0) == 0
    assert rel_err(out - 2.5, g["cone_a/cubic/fwd"][0]) < 1e-5


def test_dropin_source_plane_failure(lib):
    """A voxel centre in the source plane returns TRTT_FAILURE
    (trtt_3D_projectors.c:1264-1284)."""
    from yoricktrtt_amd.geometry import cone_geometry

    _say("source-plane failure")
    nx = ny = nz = 5
    g = cone_geometry(nx, ny, nz, 1.0, 8, 8, 1.0, 1.0, [0.0], 0.0, 10.0)  # Rsc = 0
    vol = np.ones((nz, ny, nx))
    out = np.zeros((8, 8))
    r = lib.trtt_CB3D_cubic(_p(out), _p(vol), nx, ny, nz, 1.0, 3, 8, 8, 1.0, 1.0,
                            _p(np.ascontiguousarray(g.xos[0])), _p(np.ascontiguousarray(g.xsd[0])), 0)
    assert r == -1


def _lf_params(n1, n2, nd, ps, pst, vs, R, D, o1=0.0, o2=0.0):
    from yoricktrtt_amd._lib import LF2DParameters

    return LF2DParameters(ps, pst, vs, R, D, o1, o2, n1, n2, nd)


def test_dropin_lf2d_vs_oracle(lib, oracle, precision):
    from oracle.oracle import LF2DParams

    _say("lf2d dropin", precision)
    rng = np.random.default_rng(3)
    args = (23, 19, 71, 0.9, 1.0, 1.3, 60.0, 40.0, 0.7, -1.1)
    p = _lf_params(*args)
    po = LF2DParams(args[3], args[4], args[5], args[6], args[7], args[8], args[9], args[0], args[1],
                    args[2])
    for beta in (0.0, 0.9, 2.5, 5.1):
        x = rng.uniform(size=(19, 23))
        y = rng.uniform(size=71)
        hx = np.zeros(71)
        lib.lf2d_direct(C.byref(p), beta, 1.5, _p(x), _p(hx), 1)
        hty = np.zeros((19, 23))
        lib.lf2d_adjoint(C.byref(p), beta, 1.5, _p(y), _p(hty), 1)
        rhx = oracle.lf2d(po, beta, 1.5, x, np.zeros(71), 1, 0)
        rhty = oracle.lf2d(po, beta, 1.5, y, np.zeros((19, 23)), 1, 1)
        assert rel_err(hx, rhx) < TOL[precision]
        assert rel_err(hty, rhty) < TOL[precision]


def _dev(a, dtype):
    return torch.tensor(a, dtype=dtype, device="cuda:0")


@pytest.mark.parametrize("kind", ["cone", "parallel3d"])
@pytest.mark.parametrize("model", ["cubic", "table"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batched_plan_vs_oracle(lib, oracle, kind, model, prec):
    """All views in one launch: equal to the oracle view by view."""
    from yoricktrtt_amd.geometry import cone_geometry, footprint_table, parallel_geometry_3d
    from yoricktrtt_amd.projectors import Projector

    _say("batched", kind, model, prec)
    nx, ny, nz, nu, nv, V = 40, 36, 30, 44, 52, 12
    th = np.arange(V) * (2 * math.pi / V) + 0.1
    if kind == "cone":
        geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 120.0, 180.0)
    else:
        geo = parallel_geometry_3d(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th)
    rng = np.random.default_rng(11)
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    P = Projector(geo, model=model, precision=prec, size_footprint=10000)
    # FP32 plans take the axial kernels, with a cubic CDF table too
    expect_axial = prec == "f32"
    assert (P.path > 0) == expect_axial, P.path
    dt = torch.float64 if prec == "f64" else torch.float32
    fwd = P.forward(_dev(vol, dt)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, dt)).double().cpu().numpy()
    tab = None
    if model == "table":
        _, cum, coord, step = footprint_table(10000)
        tab = (cum, coord, step)
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        rf = oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v],
                         geo.xsd[v], kind == "cone", 0, tab)
        assert rel_err(fwd[v], rf) < TOL[prec], v
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v], geo.xsd[v],
                    kind == "cone", 1, tab)
    assert rel_err(bwd, rb) < TOL[prec]
    P.close()


@pytest.mark.parametrize("kind", ["fan", "parallel2d"])
@pytest.mark.parametrize("model", ["cubic", "table"])
def test_batched_plan_2d_vs_oracle(lib, oracle, kind, model):
    from yoricktrtt_amd.geometry import fan_geometry_2d, footprint_table, parallel_geometry_2d
    from yoricktrtt_amd.projectors import Projector

    _say("batched2d", kind, model)
    nx, ny, nv, V = 70, 64, 97, 40
    th = np.arange(V) * (2 * math.pi / V)
    if kind == "fan":
        geo = fan_geometry_2d(nx, ny, 1.0, nv, 1.0, th, 200.0, 300.0)
    else:
        geo = parallel_geometry_2d(nx, ny, 1.0, nv, 1.0, th)
    rng = np.random.default_rng(12)
    img = rng.uniform(size=(ny, nx))
    sino = rng.uniform(size=(V, nv))
    P = Projector(geo, model=model, precision="f32")
    fwd = P.forward(_dev(img, torch.float32)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    tab = None
    if model == "table":
        _, cum, coord, step = footprint_table(10000)
        tab = (cum, coord, step)
    rb = np.zeros((ny, nx))
    for v in range(V):
        rf = oracle.sb2d(np.zeros(nv), img, nx, ny, 1.0, nv, 1.0, geo.xos[v], geo.xsd[v],
                         kind == "fan", 0, tab)
        assert rel_err(fwd[v], rf) < 1e-5, v
        oracle.sb2d(rb, sino[v], nx, ny, 1.0, nv, 1.0, geo.xos[v], geo.xsd[v], kind == "fan", 1, tab)
    assert rel_err(bwd, rb) < 1e-5
    P.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kind", ["fan", "parallel2d"])
@pytest.mark.parametrize("pitch", [0.3, 0.05, 0.01])
def test_batched_plan_2d_detector_pitch(lib, oracle, kind, pitch, prec):
    """Detector cells much finer than the pixels: footprints outgrow a
    thread's 16-cell strip of the FP32 forward (pitch 0.3 on some views,
    0.05 on all) and a tile's window outgrows the workgroup's 2048-cell LDS
    window (pitch 0.01), so the forward also takes its per-cell LDS atomic
    and global atomic paths."""
    from yoricktrtt_amd.geometry import fan_geometry_2d, parallel_geometry_2d
    from yoricktrtt_amd.projectors import Projector

    _say("batched2d pitch", kind, pitch, prec)
    nx, ny, V = 70, 64, 12
    nv = int(math.ceil(200.0 / pitch))
    th = np.arange(V) * (2 * math.pi / V) + 0.2
    if kind == "fan":
        geo = fan_geometry_2d(nx, ny, 1.0, nv, pitch, th, 200.0, 300.0)
    else:
        geo = parallel_geometry_2d(nx, ny, 1.0, nv, pitch, th)
    rng = np.random.default_rng(21)
    img = rng.uniform(size=(ny, nx))
    sino = rng.uniform(size=(V, nv))
    dt = torch.float32 if prec == "f32" else torch.float64
    P = Projector(geo, model="cubic", precision=prec)
    fwd = P.forward(_dev(img, dt)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, dt)).double().cpu().numpy()
    rb = np.zeros((ny, nx))
    for v in range(V):
        rf = oracle.sb2d(np.zeros(nv), img, nx, ny, 1.0, nv, pitch, geo.xos[v], geo.xsd[v],
                         kind == "fan", 0)
        assert rel_err(fwd[v], rf) < TOL[prec], v
        oracle.sb2d(rb, sino[v], nx, ny, 1.0, nv, pitch, geo.xos[v], geo.xsd[v], kind == "fan", 1)
    assert rel_err(bwd, rb) < TOL[prec]
    P.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_adjoint_identity(lib, prec):
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("adjoint", prec)
    nx, ny, nz, nu, nv, V = 64, 64, 48, 80, 96, 30
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, np.arange(V) * (2 * math.pi / V),
                        200.0, 300.0)
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device="cuda:0").manual_seed(5)
    x = torch.rand((nz, ny, nx), generator=g, device="cuda:0", dtype=dt)
    y = torch.rand((V, nv, nu), generator=g, device="cuda:0", dtype=dt)
    P = Projector(geo, model="cubic", precision=prec)
    Ax = P.forward(x)
    Aty = P.backward(y)
    lhs = torch.sum(Ax.double() * y.double()).item()
    rhs = torch.sum(x.double() * Aty.double()).item()
    tol = 1e-6 if prec == "f32" else 1e-12
    assert abs(lhs - rhs) <= tol * abs(lhs), (lhs, rhs)
    P.close()


def test_time_interpolated_phases(lib, oracle):
    """4-D series: view k sees w0 x[t0] + w1 x[t1]; back-projection splits
    with the same weights (one-hot and fractional weights)."""
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("phases")
    nx, ny, nz, nu, nv, nt, V = 24, 22, 20, 30, 34, 4, 14
    rng = np.random.default_rng(21)
    th = rng.uniform(0, 2 * math.pi, size=V)
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 90.0, 140.0)
    t = rng.uniform(0, nt - 1, size=V)
    t[:5] = np.floor(t[:5])  # integer (phase-binned) views
    t0 = np.floor(t).astype(np.int32)
    t1 = np.minimum(t0 + 1, nt - 1).astype(np.int32)
    w1 = t - t0
    w1[t0 == t1] = 0.0
    w0 = 1.0 - w1
    x = rng.uniform(size=(nt, nz, ny, nx))
    y = rng.uniform(size=(V, nv, nu))
    P = Projector(geo, model="cubic", precision="f64", nt=nt, t0=t0, t1=t1, w0=w0, w1=w1)
    fwd = P.forward(_dev(x, torch.float64)).cpu().numpy()
    bwd = P.backward(_dev(y, torch.float64)).cpu().numpy()
    rb = np.zeros((nt, nz, ny, nx))
    for v in range(V):
        xv = w0[v] * x[t0[v]] + w1[v] * x[t1[v]]
        rf = oracle.sb3d(np.zeros((nv, nu)), xv, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v],
                         geo.xsd[v], True, 0)
        assert rel_err(fwd[v], rf) < 1e-12
        bv = oracle.sb3d(np.zeros((nz, ny, nx)), y[v], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v],
                         geo.xsd[v], True, 1)
        rb[t0[v]] += w0[v] * bv
        rb[t1[v]] += w1[v] * bv
    assert rel_err(bwd, rb) < 1e-12
    P.close()


def test_weighted_residual(lib):
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("residual")
    geo = cone_geometry(8, 8, 8, 1.0, 10, 12, 1.0, 1.0, [0.0, 1.0, 2.0], 50.0, 80.0)
    P = Projector(geo, precision="f32")
    g = torch.Generator(device="cuda:0").manual_seed(2)
    s = torch.rand(P.sino_shape, generator=g, device="cuda:0")
    y = torch.rand(P.sino_shape, generator=g, device="cuda:0")
    w = torch.rand(P.sino_shape, generator=g, device="cuda:0")
    fx = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    d = (s - y).double()
    ref_fx = 0.5 * torch.sum(w.double() * d * d).item()
    ref_r = (w.double() * d)
    P.weighted_residual(s, y, w, fx)
    assert abs(fx.item() - ref_fx) < 1e-6 * ref_fx
    assert rel_err(s.double().cpu().numpy(), ref_r.cpu().numpy()) < 1e-6
    P.close()


@pytest.mark.parametrize("kind", ["cone", "parallel3d"])
def test_tilted_orbit_uses_general_kernels(lib, oracle, kind):
    """A rotation axis tilted away from z is not axial: the general kernels
    run and still match the oracle."""
    from yoricktrtt_amd.geometry import (Geometry, XForm3D, detector_xi, object_xi,
                                         projector_coefs, view_transforms_3d)
    from yoricktrtt_amd.projectors import Projector

    _say("tilted", kind)
    nx, ny, nz, nu, nv, V = 20, 18, 16, 30, 34, 5
    oxi = object_xi(nx, ny, nz, 1.0)
    dxi = detector_xi(nu, nv, 1.0, 1.0)
    xs, ds = [], []
    for th in np.linspace(0.2, 5.0, V):
        Rsc, Rsd = (80.0, 130.0) if kind == "cone" else (0.0, 0.0)
        xos, xot, xsd = view_transforms_3d(th, Rsc, Rsd)
        tilt = XForm3D().rotate(math.cos(0.15), math.sin(0.15), 0.0, 0.0)
        xos = XForm3D.combine(xos, tilt)
        a, b = projector_coefs(oxi, dxi, xos, xot, xsd)
        xs.append(a)
        ds.append(b)
    geo = Geometry(kind, nx, ny, nz, nu, nv, 1.0, 1.0, 1.0, np.array(xs), np.array(ds))
    P = Projector(geo, model="cubic", precision="f32")
    assert P.path == 0
    rng = np.random.default_rng(4)
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    fwd = P.forward(_dev(vol, torch.float32)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        rf = oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v],
                         geo.xsd[v], kind == "cone", 0)
        assert rel_err(fwd[v], rf) < 1e-5
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v], geo.xsd[v],
                    kind == "cone", 1)
    assert rel_err(bwd, rb) < 1e-5
    P.close()


def _tab(model, m=10000):
    """(cum, coord, step) of an m-entry cubic table for the oracle, or None."""
    if model != "table":
        return None
    from yoricktrtt_amd.geometry import footprint_table

    _, cum, coord, step = footprint_table(m)
    return cum, coord, step


@pytest.mark.parametrize("model", ["cubic", "table"])
def test_axial_wide_footprints_and_truncation(lib, oracle, model):
    """Large magnification (W = 13 kernels) and a detector smaller than the
    projected volume (windows clipped at every edge, hulls at the borders)."""
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("wide footprints")
    nx, ny, nz, nu, nv, V = 33, 29, 70, 41, 45, 6
    us, vs = 1.0, 1.0
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, us, vs, np.linspace(0.0, 6.0, V), 70.0, 120.0,
                        u_off=3.0, v_off=-2.0)
    P = Projector(geo, model=model, precision="f32")
    assert P.path == 13
    rng = np.random.default_rng(8)
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    fwd = P.forward(_dev(vol, torch.float32)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        rf = oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, us, vs, geo.xos[v],
                         geo.xsd[v], True, 0, _tab(model))
        assert rel_err(fwd[v], rf) < 1e-5
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, us, vs, geo.xos[v], geo.xsd[v], True, 1,
                    _tab(model))
    assert rel_err(bwd, rb) < 1e-5
    P.close()


@pytest.mark.parametrize("kind,us", [("parallel", 1.5), ("parallel", 2.5), ("parallel", 3.5),
                                     ("parallel", 4.5), ("cone", 2.5), ("cone", 5.0)])
@pytest.mark.parametrize("model", ["cubic", "table"])
def test_axial_coarse_detector_vs_oracle(lib, oracle, kind, us, model):
    """Detector pixels along the rotation axis coarser than the projected
    slice pitch: voxels of neighbouring slices start their windows on the
    same pixel, so the forward accumulates its column profiles in 2, 3 or 4
    interleaved plain passes (parallel pitch 1.5 / 2.5 / 3.5, cone 2.5 / 5)
    or with LDS atomics (parallel 4.5)."""
    from yoricktrtt_amd.geometry import cone_geometry, parallel_geometry_3d
    from yoricktrtt_amd.projectors import Projector

    _say("axial coarse detector", kind, us)
    nx, ny, nz, nv, V = 24, 22, 64, 48, 5
    th = np.linspace(0.1, 5.0, V)
    if kind == "cone":
        nu = int(math.ceil(64 * 1.8 / us)) + 6
        geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, us, 1.0, th, 70.0, 120.0)
    else:
        nu = int(math.ceil(64 / us)) + 6
        geo = parallel_geometry_3d(nx, ny, nz, 1.0, nu, nv, us, 1.0, th)
    P = Projector(geo, model=model, precision="f32")
    assert P.path in (9, 13)
    rng = np.random.default_rng(int(us * 10))
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    fwd = P.forward(_dev(vol, torch.float32)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        rf = oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, us, 1.0, geo.xos[v],
                         geo.xsd[v], kind == "cone", 0, _tab(model))
        assert rel_err(fwd[v], rf) < 1e-5, v
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, us, 1.0, geo.xos[v], geo.xsd[v],
                    kind == "cone", 1, _tab(model))
    assert rel_err(bwd, rb) < 1e-5
    P.close()


@pytest.mark.parametrize("nt", [1, 3])
@pytest.mark.parametrize("model", ["cubic", "table"])
def test_axial_backward_tiles_vs_oracle(lib, oracle, nt, model):
    """Every backward launch shape, including the tiles of 128 slices (two
    voxels per lane), against the FP64 oracle: a slice count that leaves a
    partial last tile for both k extents, and (nt = 3) views interpolated
    between two phases next to one-hot ones."""
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    nx, ny, nz, nu, nv, V = 20, 18, 150, 250, 40, 8
    _say("backward tiles", nt)
    rng = np.random.default_rng(40 + nt)
    th = rng.uniform(0, 2 * math.pi, size=V)
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 300.0, 450.0)
    kw = {}
    t0 = np.zeros(V, np.int32)
    t1 = np.zeros(V, np.int32)
    w0 = np.ones(V)
    w1 = np.zeros(V)
    if nt > 1:
        t = rng.uniform(0, nt - 1, size=V)
        t[:3] = np.floor(t[:3])
        t0 = np.floor(t).astype(np.int32)
        t1 = np.minimum(t0 + 1, nt - 1).astype(np.int32)
        w1 = t - t0
        w1[t0 == t1] = 0.0
        w0 = 1.0 - w1
        kw = dict(nt=nt, t0=t0, t1=t1, w0=w0, w1=w1)
    P = Projector(geo, model=model, precision="f32", **kw)
    assert P.path > 0
    y = rng.uniform(size=(V, nv, nu))
    rb = np.zeros((nt, nz, ny, nx))
    for v in range(V):
        bv = oracle.sb3d(np.zeros((nz, ny, nx)), y[v], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0,
                         geo.xos[v], geo.xsd[v], True, 1, _tab(model))
        rb[t0[v]] += w0[v] * bv
        rb[t1[v]] += w1[v] * bv
    for tile in sorted(AT.BWD_TILES):
        AT.set_tuning(P, bwd_tile=tile, bwd_window_bytes=32768)
        bwd = P.backward(_dev(y, torch.float32)).double().cpu().numpy().reshape(rb.shape)
        assert rel_err(bwd, rb) < 1e-5, tile
    P.close()



@pytest.mark.parametrize("model", ["cubic", "table"])
def test_axial_direct_path_and_edges_vs_oracle(lib, oracle, model):
    """The smallest LDS windows send most tiles down the direct path of both
    kernels (the block entered when a lane's window does not fit), on a
    detector smaller than the projected volume, so that windows cross both
    edges (margin entries in front of the profile and the Q row)."""
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("direct path and edges")
    nx, ny, nz, nu, nv, V = 26, 23, 60, 52, 44, 5
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, np.linspace(0.3, 5.9, V), 110.0, 170.0,
                        u_off=-2.5, v_off=1.5)
    P = Projector(geo, model=model, precision="f32")
    assert P.path > 0
    rng = np.random.default_rng(17)
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    rf = np.stack([oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0,
                               geo.xos[v], geo.xsd[v], True, 0, _tab(model)) for v in range(V)])
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v], geo.xsd[v], True, 1,
                    _tab(model))
    for wf, wb in [(4096, 4096), (8192, 16384)]:
        AT.set_tuning(P, fwd_window_bytes=wf, bwd_window_bytes=wb)
        fwd = P.forward(_dev(vol, torch.float32)).double().cpu().numpy()
        bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
        assert rel_err(fwd, rf) < 1e-5, wf
        assert rel_err(bwd, rb) < 1e-5, wb
    P.close()


@pytest.mark.parametrize("case", ["quadratic", "perturbed", "short", "mismatch"])
def test_table_not_cubic_cdf_uses_general_kernels(lib, oracle, case):
    """Only a table that holds the B-spline CDF of the plan's degree at its
    nodes takes the axial kernels: a 10^3-entry degree-2 table (trapezoid
    error 3e-6), a cubic table with a perturbation of 1e-4, a 500-entry
    cubic table (trapezoid error 1.4e-5), and a cubic table on a degree-2
    plan run the general kernels, and all match the oracle with their own
    table."""
    from yoricktrtt_amd.geometry import cone_geometry, footprint_table
    from yoricktrtt_amd.projectors import Projector

    _say("table fallback", case)
    nx, ny, nz, nu, nv, V = 22, 20, 18, 30, 34, 4
    deg = 2 if case in ("quadratic", "mismatch") else 3
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, np.linspace(0.1, 4.0, V), 80.0, 130.0,
                        s_deg=deg)
    m = {"short": 500, "quadratic": 1000}.get(case, 10000)
    _, cum, coord, step = footprint_table(m, 3 if case == "mismatch" else deg)
    if case == "perturbed":
        cum = cum + 1e-4 * np.sin(np.arange(cum.size) * 0.01)
    P = Projector(geo, model="table", precision="f32", table=(cum, coord, step))
    assert P.path == 0, case
    rng = np.random.default_rng(21)
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    fwd = P.forward(_dev(vol, torch.float32)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        rf = oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v],
                         geo.xsd[v], True, 0, (cum, coord, step), deg)
        assert rel_err(fwd[v], rf) < 1e-5
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v], geo.xsd[v], True, 1,
                    (cum, coord, step), deg)
    assert rel_err(bwd, rb) < 1e-5
    P.close()


@pytest.mark.parametrize("m", [10000, 100000])
def test_axial_table_cone_vs_oracle(lib, oracle, m):
    """The axial table path on a cone plan with the 10^4 (2-D / 4-D scripts)
    and 10^5 (3-D script) tables, every view forward and back."""
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("axial table", m)
    nx, ny, nz, nu, nv, V = 40, 36, 70, 64, 72, 9
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, np.linspace(0.0, 6.0, V), 120.0, 180.0)
    P = Projector(geo, model="table", precision="f32", size_footprint=m)
    assert P.path > 0
    rng = np.random.default_rng(m % 97)
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    fwd = P.forward(_dev(vol, torch.float32)).double().cpu().numpy()
    bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
    tab = _tab("table", m)
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        rf = oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v],
                         geo.xsd[v], True, 0, tab)
        assert rel_err(fwd[v], rf) < 1e-5, v
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v], geo.xsd[v], True, 1,
                    tab)
    assert rel_err(bwd, rb) < 1e-5
    P.close()


@pytest.mark.parametrize("m", [1000, 10000])
def test_deg2_tables_vs_reference(lib, golden3d, golden2d, precision, m):
    """s_deg = 2 with a degree-2 table (the reference's 4-D benchmark
    setting): the drop-in symbols and the batched plans against the
    reference C outputs of tests/golden/deg2.npz."""
    import os

    from conftest import GOLDEN
    from yoricktrtt_amd.geometry import Geometry
    from yoricktrtt_amd.projectors import Projector

    _say("deg2", m, precision)
    d = dict(np.load(os.path.join(GOLDEN, "deg2.npz")))
    cum, coord, step = d[f"m{m}/cum"], d[f"m{m}/coord"], float(d[f"m{m}/step"])
    fp = np.zeros_like(cum)
    tol = TOL[precision]
    dt = torch.float64 if precision == "f64" else torch.float32
    vol = np.ascontiguousarray(golden3d["vol"])
    nz, ny, nx = vol.shape
    for name in ("cone_a", "cone_edge", "par_a"):
        nu, nv = (int(v) for v in golden3d[f"{name}/dims"])
        us, vs = (float(v) for v in golden3d[f"{name}/scl"])
        xos, xsd = golden3d[f"{name}/xos"], golden3d[f"{name}/xsd"]
        sino_in = golden3d[f"{name}/sino_in"]
        cone = name.startswith("cone")
        fn = lib.trtt_CB3D if cone else lib.trtt_PB3D
        rf, rb = d[f"m{m}/3d/{name}/fwd"], d[f"m{m}/3d/{name}/bwd"]
        for v in range(xos.shape[0]):
            xo, xs = np.ascontiguousarray(xos[v]), np.ascontiguousarray(xsd[v])
            of = np.zeros((nv, nu))
            ob = np.zeros((nz, ny, nx))
            assert fn(_p(of), _p(vol), nx, ny, nz, 1.0, 2, _p(fp), _p(cum), m, step, _p(coord), nu,
                      nv, us, vs, _p(xo), _p(xs), 0) == 0
            assert fn(_p(ob), _p(np.ascontiguousarray(sino_in[v])), nx, ny, nz, 1.0, 2, _p(fp),
                      _p(cum), m, step, _p(coord), nu, nv, us, vs, _p(xo), _p(xs), 1) == 0
            assert rel_err(of, rf[v]) < tol and rel_err(ob, rb[v]) < tol, (name, v)
        geo = Geometry("cone" if cone else "parallel3d", nx, ny, nz, nu, nv, 1.0, us, vs,
                       np.ascontiguousarray(xos), np.ascontiguousarray(xsd), s_deg=2)
        P = Projector(geo, model="table", precision=precision, table=(cum, coord, step))
        # only FP32 plans whose table holds the quadratic B-spline CDF at its
        # nodes within 1e-6 (10^4 entries: 3e-8; 10^3 entries: 3e-6) may
        # take the axial kernels
        if precision == "f64" or m == 1000:
            assert P.path == 0
        fwd = P.forward(_dev(vol, dt)).double().cpu().numpy()
        bwd = P.backward(_dev(sino_in, dt)).double().cpu().numpy()
        assert rel_err(fwd, rf) < tol and rel_err(bwd, rb.sum(axis=0)) < tol, name
        P.close()
    img = np.ascontiguousarray(golden2d["img"])
    ny2, nx2 = img.shape
    for name in ("fan_a", "fan_edge", "par_a"):
        nv, vs = int(golden2d[f"{name}/nv"]), float(golden2d[f"{name}/vs"])
        xos, xsd = golden2d[f"{name}/xos"], golden2d[f"{name}/xsd"]
        sino_in = golden2d[f"{name}/sino_in"]
        fan = name.startswith("fan")
        fn = lib.trtt_FB2D if fan else lib.trtt_PB2D
        rf, rb = d[f"m{m}/2d/{name}/fwd"], d[f"m{m}/2d/{name}/bwd"]
        for v in range(xos.shape[0]):
            xo, xs = np.ascontiguousarray(xos[v]), np.ascontiguousarray(xsd[v])
            of = np.zeros(nv)
            ob = np.zeros((ny2, nx2))
            assert fn(_p(of), _p(img), nx2, ny2, 1.0, 2, _p(fp), _p(cum), m, step, _p(coord), nv,
                      vs, _p(xo), _p(xs), 0) == 0
            assert fn(_p(ob), _p(np.ascontiguousarray(sino_in[v])), nx2, ny2, 1.0, 2, _p(fp),
                      _p(cum), m, step, _p(coord), nv, vs, _p(xo), _p(xs), 1) == 0
            assert rel_err(of, rf[v]) < tol and rel_err(ob, rb[v]) < tol, (name, v)


@pytest.mark.parametrize("case", ["fan_f32", "lf2d_f32", "cone_f64", "tilted_f32"])
def test_atomic_forwards_run_to_run(lib, case):
    """The general forwards sum each tile's window in LDS with atomics and add
    it to the sinogram with global atomics, so the summation order, and the
    last bits of an FP32 result, vary between runs (a few FP32 roundings per
    pixel): three runs agree within 1e-6 relative (FP32) or 1e-14 (FP64)."""
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.geometry import (Geometry, XForm3D, cone_geometry, detector_xi,
                                         fan_geometry_2d, object_xi, projector_coefs,
                                         view_transforms_3d)
    from yoricktrtt_amd.projectors import LFProjector, Projector

    _say("run to run", case)
    rng = np.random.default_rng(77)
    if case == "fan_f32":
        geo = fan_geometry_2d(300, 280, 1.0, 420, 1.0, np.arange(180) * (2 * math.pi / 180), 500.0,
                              800.0)
        P = Projector(geo, model="cubic", precision="f32")
        x = rng.uniform(size=(280, 300))
    elif case == "lf2d_f32":
        V = 180
        P = LFProjector(_lib.LF2DParameters(0.75, 0.75, 0.5, 400.0, 200.0, 0.0, 0.0, 300, 280, 420),
                        np.arange(V) * (2 * math.pi / V), np.zeros(V), precision="f32")
        x = rng.uniform(size=(280, 300))
    elif case == "cone_f64":
        geo = cone_geometry(60, 56, 50, 1.0, 80, 90, 1.0, 1.0, np.arange(24) * 0.26, 150.0, 220.0)
        P = Projector(geo, model="cubic", precision="f64")
        x = rng.uniform(size=(50, 56, 60))
    else:
        nx, ny, nz, nu, nv, V = 48, 44, 40, 64, 70, 12
        xs, ds = [], []
        for th in np.linspace(0.2, 5.0, V):
            xos, xot, xsd = view_transforms_3d(th, 120.0, 190.0)
            xos = XForm3D.combine(xos, XForm3D().rotate(math.cos(0.1), math.sin(0.1), 0.0, 0.0))
            a, b = projector_coefs(object_xi(nx, ny, nz, 1.0), detector_xi(nu, nv, 1.0, 1.0), xos,
                                   xot, xsd)
            xs.append(a)
            ds.append(b)
        geo = Geometry("cone", nx, ny, nz, nu, nv, 1.0, 1.0, 1.0, np.array(xs), np.array(ds))
        P = Projector(geo, model="cubic", precision="f32")
        x = rng.uniform(size=(nz, ny, nx))
    assert P.path == 0  # the general kernels
    dt = P.dtype
    xd = _dev(x, dt)
    runs = [P.forward(xd).double().cpu().numpy() for _ in range(3)]
    tol = 1e-14 if dt == torch.float64 else 1e-6
    worst = max(rel_err(r, runs[0]) for r in runs[1:])
    _say(f"  run-to-run {worst:.2e}")
    assert worst < tol
    P.close()


@pytest.mark.parametrize("nt", [1, 3])
@pytest.mark.parametrize("model", ["cubic", "table"])
def test_axial_forward_tiles_vs_oracle(lib, oracle, nt, model):
    """Every forward launch shape with several views per workgroup, both
    ways of adding the tile windows to the projections (global atomics, or
    scratch slots summed by the reduction kernel), against the FP64 oracle:
    a slice count that leaves a partial last tile, a detector smaller than
    the projected volume, views interpolated between two phases (nt = 3),
    and the smallest window (most tiles take the direct path)."""
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    nx, ny, nz, nu, nv, V = 21, 18, 150, 230, 36, 7
    _say("forward tiles", nt, model)
    rng = np.random.default_rng(60 + nt)
    th = rng.uniform(0, 2 * math.pi, size=V)
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 300.0, 450.0, v_off=1.5)
    t0 = np.zeros(V, np.int32)
    t1 = np.zeros(V, np.int32)
    w0 = np.ones(V)
    w1 = np.zeros(V)
    kw = {}
    if nt > 1:
        t = rng.uniform(0, nt - 1, size=V)
        t[:3] = np.floor(t[:3])
        t0 = np.floor(t).astype(np.int32)
        t1 = np.minimum(t0 + 1, nt - 1).astype(np.int32)
        w1 = t - t0
        w1[t0 == t1] = 0.0
        w0 = 1.0 - w1
        kw = dict(nt=nt, t0=t0, t1=t1, w0=w0, w1=w1)
    P = Projector(geo, model=model, precision="f32", **kw)
    assert P.path > 0
    x = rng.uniform(size=(nt, nz, ny, nx))
    rf = np.zeros((V, nv, nu))
    for v in range(V):
        xv = w0[v] * x[t0[v]] + w1[v] * x[t1[v]]
        oracle.sb3d(rf[v], xv, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[v], geo.xsd[v], True, 0,
                    _tab(model))
    xd = _dev(x.reshape(P.vol_shape), torch.float32)
    for tile in sorted(AT.FWD_TILES):
        # fwd_xcds: XCD-aware block order (8: MI355X; 3: a count that leaves
        # padding blocks in the 1-D grid)
        for wb, vpb, flush, xcds in [(24576, 0, 1, 0), (16384, 3, 0, 0), (16384, 3, 1, 0),
                                     (4096, 1, 1, 0), (8192, 4, 1, 8), (16384, 3, 1, 3)]:
            AT.set_tuning(P, fwd_tile=tile, fwd_window_bytes=wb, fwd_views_per_block=vpb,
                          fwd_flush=flush, fwd_xcds=xcds)
            fwd = P.forward(xd).double().cpu().numpy()
            assert rel_err(fwd, rf) < 1e-5, (tile, wb, vpb, flush, xcds)
    assert P.error_flag() == 0
    P.close()


def test_weighted_residual_refuses_mismatched_arguments(lib):
    """Data of another type or shape, and an fx that is not one FP64
    element, are refused before any launch (the kernel would read them as
    the plan's type and shape)."""
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("residual argument checks")
    geo = cone_geometry(8, 8, 8, 1.0, 10, 12, 1.0, 1.0, [0.0, 1.0], 50.0, 80.0)
    P = Projector(geo, precision="f32")
    s = torch.zeros(P.sino_shape, device="cuda:0")
    fx = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError):
        P.weighted_residual(s, torch.zeros(P.sino_shape, dtype=torch.float64, device="cuda:0"),
                            None, fx)
    with pytest.raises(ValueError):
        P.weighted_residual(s, s.clone(), torch.zeros((2, 12, 9), device="cuda:0"), fx)
    with pytest.raises(ValueError):
        P.weighted_residual(s, s.clone(), None, torch.zeros(1, device="cuda:0"))
    with pytest.raises(ValueError):
        P.weighted_residual(s, s.clone(), None, torch.zeros(2, dtype=torch.float64, device="cuda:0"))
    P.close()


def test_cubic_deg2_vs_reference(lib, golden3d, golden2d, precision):
    """The closed-form cubic symbols called with s_deg = 2 (a truncated
    cubic footprint: the pixel window of a 1.5-voxel half support) against
    the reference C outputs of tests/golden/deg2.npz: drop-ins and plans
    (such plans run the general kernels)."""
    import os

    from conftest import GOLDEN
    from yoricktrtt_amd.geometry import Geometry
    from yoricktrtt_amd.projectors import Projector

    _say("cubic deg2", precision)
    d = dict(np.load(os.path.join(GOLDEN, "deg2.npz")))
    tol = TOL[precision]
    dt = torch.float64 if precision == "f64" else torch.float32
    vol = np.ascontiguousarray(golden3d["vol"])
    nz, ny, nx = vol.shape
    for name in ("cone_a", "cone_edge", "par_a"):
        nu, nv = (int(v) for v in golden3d[f"{name}/dims"])
        us, vs = (float(v) for v in golden3d[f"{name}/scl"])
        xos, xsd = golden3d[f"{name}/xos"], golden3d[f"{name}/xsd"]
        sino_in = golden3d[f"{name}/sino_in"]
        cone = name.startswith("cone")
        fn = lib.trtt_CB3D_cubic if cone else lib.trtt_PB3D_cubic
        rf, rb = d[f"cubic/3d/{name}/fwd"], d[f"cubic/3d/{name}/bwd"]
        for v in range(xos.shape[0]):
            xo, xs = np.ascontiguousarray(xos[v]), np.ascontiguousarray(xsd[v])
            of = np.zeros((nv, nu))
            ob = np.zeros((nz, ny, nx))
            assert fn(_p(of), _p(vol), nx, ny, nz, 1.0, 2, nu, nv, us, vs, _p(xo), _p(xs), 0) == 0
            assert fn(_p(ob), _p(np.ascontiguousarray(sino_in[v])), nx, ny, nz, 1.0, 2, nu, nv, us,
                      vs, _p(xo), _p(xs), 1) == 0
            assert rel_err(of, rf[v]) < tol and rel_err(ob, rb[v]) < tol, (name, v)
        geo = Geometry("cone" if cone else "parallel3d", nx, ny, nz, nu, nv, 1.0, us, vs,
                       np.ascontiguousarray(xos), np.ascontiguousarray(xsd), s_deg=2)
        P = Projector(geo, model="cubic", precision=precision)
        assert P.path == 0
        fwd = P.forward(_dev(vol, dt)).double().cpu().numpy()
        bwd = P.backward(_dev(sino_in, dt)).double().cpu().numpy()
        assert rel_err(fwd, rf) < tol and rel_err(bwd, rb.sum(axis=0)) < tol, name
        P.close()
    img = np.ascontiguousarray(golden2d["img"])
    ny2, nx2 = img.shape
    for name in ("fan_a", "fan_edge", "par_a"):
        nv, vs = int(golden2d[f"{name}/nv"]), float(golden2d[f"{name}/vs"])
        xos, xsd = golden2d[f"{name}/xos"], golden2d[f"{name}/xsd"]
        sino_in = golden2d[f"{name}/sino_in"]
        fan = name.startswith("fan")
        fn = lib.trtt_FB2D_cubic if fan else lib.trtt_PB2D_cubic
        rf, rb = d[f"cubic/2d/{name}/fwd"], d[f"cubic/2d/{name}/bwd"]
        for v in range(xos.shape[0]):
            xo, xs = np.ascontiguousarray(xos[v]), np.ascontiguousarray(xsd[v])
            of = np.zeros(nv)
            ob = np.zeros((ny2, nx2))
            assert fn(_p(of), _p(img), nx2, ny2, 1.0, 2, nv, vs, _p(xo), _p(xs), 0) == 0
            assert fn(_p(ob), _p(np.ascontiguousarray(sino_in[v])), nx2, ny2, 1.0, 2, nv, vs,
                      _p(xo), _p(xs), 1) == 0
            assert rel_err(of, rf[v]) < tol and rel_err(ob, rb[v]) < tol, (name, v)


@pytest.mark.parametrize("kind", ["cone", "parallel3d"])
def test_axial_deg2_table_vs_oracle(lib, oracle, kind):
    """A degree-2 table of 10^4 entries (within 3e-8 of the quadratic
    B-spline CDF at its nodes) on circular orbits runs the axial kernels with
    the quadratic CDF evaluated at the table nodes: every forward tile and
    both backward tile kinds against the FP64 oracle of the same table (the
    oracle is pinned to the reference C by the degree-2 golden vectors)."""
    from yoricktrtt_amd import autotune as AT
    from yoricktrtt_amd.geometry import cone_geometry, footprint_table, parallel_geometry_3d
    from yoricktrtt_amd.projectors import Projector

    _say("axial deg2 table", kind)
    nx, ny, nz, nu, nv, V = 22, 19, 70, 90, 40, 6
    rng = np.random.default_rng(88)
    th = rng.uniform(0, 2 * math.pi, size=V)
    if kind == "cone":
        geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 120.0, 190.0, v_off=0.7,
                            s_deg=2)
    else:
        geo = parallel_geometry_3d(nx, ny, nz, 1.0, nu, nv, 0.8, 0.9, th, s_deg=2)
    _, cum, coord, step = footprint_table(10000, 2)
    P = Projector(geo, model="table", precision="f32", table=(cum, coord, step))
    assert P.path > 0
    vol = rng.uniform(size=(nz, ny, nx))
    sino = rng.uniform(size=(V, nv, nu))
    tab = (cum, coord, step)
    us, vs = geo.u_scl, geo.v_scl
    rf = np.stack([oracle.sb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, us, vs,
                               geo.xos[v], geo.xsd[v], kind == "cone", 0, tab, deg=2)
                   for v in range(V)])
    rb = np.zeros((nz, ny, nx))
    for v in range(V):
        oracle.sb3d(rb, sino[v], nx, ny, nz, 1.0, nu, nv, us, vs, geo.xos[v], geo.xsd[v],
                    kind == "cone", 1, tab, deg=2)
    for tile in sorted(AT.FWD_TILES):
        AT.set_tuning(P, fwd_tile=tile, fwd_window_bytes=16384)
        fwd = P.forward(_dev(vol, torch.float32)).double().cpu().numpy()
        assert rel_err(fwd, rf) < 1e-5, tile
    for tile in (0, 3):
        AT.set_tuning(P, bwd_tile=tile, bwd_window_bytes=32768)
        bwd = P.backward(_dev(sino, torch.float32)).double().cpu().numpy()
        assert rel_err(bwd, rb) < 1e-5, tile
    assert P.error_flag() == 0
    P.close()
