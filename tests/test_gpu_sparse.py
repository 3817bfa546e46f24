"""GPU sparse system matrix of the 2-D plans (trtt_hip_sparse_*) against the
oracle: the entries in the reference's order (view, pixel, bin) with its
6-decimal coefficients or exact ones, and the two products against the
plan's own projectors.

Tolerances: entries of FP64 table plans equal the oracle's (coefficients
1e-12 relative, 6-decimal coefficients equal but for rounding-boundary
flips of one unit of the sixth decimal); products 1e-12 (FP64) and 1e-5
(FP32) relative to the projectors' results.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import rel_err

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


def _geometry(kind, nx, ny, nv, V, **kw):
    from yoricktrtt_amd.geometry import fan_geometry_2d, parallel_geometry_2d

    th = np.arange(V) * (2 * math.pi / V) + 0.05
    if kind == "fan":
        return fan_geometry_2d(nx, ny, 1.0, nv, 1.0, th, kw.get("Rsc", 40.0), kw.get("Rsd", 60.0),
                               voff=kw.get("voff"))
    return parallel_geometry_2d(nx, ny, 1.0, nv, 1.0, th, voff=kw.get("voff"))


def _oracle_coo(oracle, geo, tab, decimals):
    nx, ny, nv = geo.nx, geo.ny, geo.nv
    fan = geo.kind == "fan"

    def column(v, i, j):
        img = np.zeros((ny, nx))
        img[j, i] = 1.0
        return oracle.sb2d(np.zeros(nv), img, nx, ny, 1.0, nv, 1.0, geo.xos[v], geo.xsd[v], fan, 0,
                           tab)
    return oracle.sparse2d_coo(column, geo.nviews, nx, ny, decimals)


def _dense(coefs, rows, cols, nrows, ncols):
    m = np.zeros((nrows, ncols))
    np.add.at(m, (rows - 1, cols - 1), coefs)
    return m


@pytest.mark.parametrize("kind", ["fan", "parallel2d"])
@pytest.mark.parametrize("decimals", [6, -1])
def test_sparse_entries_table_vs_oracle(lib, oracle, kind, decimals):
    from yoricktrtt_amd.geometry import footprint_table
    from yoricktrtt_amd.projectors import Projector

    _say("sparse entries", kind, decimals)
    # a detector narrower than the object: footprints cut at both edges
    geo = _geometry(kind, 21, 17, 23, 5, voff=np.array([0.0, 2.5, -3.0, 0.7, 0.0]))
    P = Projector(geo, model="table", precision="f64", size_footprint=10000)
    M = P.sparse_matrix(decimals)
    coefs, rows, cols = M.coo()
    _, cum, coord, step = footprint_table(10000)
    rc, rr, rcol = _oracle_coo(oracle, geo, (cum, coord, step), decimals)
    assert M.nnz == rc.size > 0
    assert np.array_equal(rows, rr) and np.array_equal(cols, rcol)
    if decimals == 6:
        diff = np.abs(coefs - rc)
        assert diff.max() <= 1.0000001e-6
        assert np.count_nonzero(diff) <= max(1, rc.size // 1000)
    else:
        assert rel_err(coefs, rc) < 1e-12
    M.close()
    P.close()


@pytest.mark.parametrize("kind", ["fan", "parallel2d"])
def test_sparse_entries_cubic_vs_oracle(lib, oracle, kind):
    from yoricktrtt_amd.projectors import Projector

    _say("sparse entries cubic", kind)
    geo = _geometry(kind, 19, 22, 29, 4, voff=np.array([1.5, 0.0, -2.0, 0.0]))
    P = Projector(geo, model="cubic", precision="f64")
    M = P.sparse_matrix(-1)
    coefs, rows, cols = M.coo()
    rc, rr, rcol = _oracle_coo(oracle, geo, None, -1)
    n = (geo.nviews * geo.nv, geo.nx * geo.ny)
    assert rel_err(_dense(coefs, rows, cols, *n), _dense(rc, rr, rcol, *n)) < 1e-12
    # the reference's order: view-major, then pixel, then bin
    key = (cols - 1) + (rows - 1) // geo.nv * (geo.nx * geo.ny)
    assert np.all(np.diff(key) >= 0)
    M.close()
    P.close()


@pytest.mark.parametrize("line", [False, True])
def test_sparse_entries_lf2d_vs_oracle(lib, oracle, line):
    """Long & Fessler matrix (lf2d_sparser, or lf2d_line_integ_sparser for
    the line-integral model) against the oracle's one-hot columns."""
    from oracle.oracle import LF2DParams
    from yoricktrtt_amd._lib import LF2DParameters
    from yoricktrtt_amd.projectors import LFProjector

    _say("sparse entries lf2d", "line" if line else "")
    n1, n2, nd = 15, 13, 41
    args = (0.9, 1.0, 1.3, 60.0, 40.0, 0.7, -1.1)
    beta = np.array([0.0, 0.9, 2.5, 5.1])
    off = np.array([1.5, 0.0, -0.5, 0.2])
    P = LFProjector(LF2DParameters(*args, n1, n2, nd), beta, off, precision="f64",
                    line_integ=line)
    M = P.sparse_matrix(6)
    coefs, rows, cols = M.coo()
    po = LF2DParams(*args, n1, n2, nd)

    def column(v, i, j):
        img = np.zeros((n2, n1))
        img[j, i] = 1.0
        return oracle.lf2d(po, beta[v], off[v], img, np.zeros(nd), 1, 0, line=line)
    rc, rr, rcol = oracle.sparse2d_coo(column, beta.size, n1, n2, 6)
    assert np.array_equal(rows, rr) and np.array_equal(cols, rcol)
    diff = np.abs(coefs - rc)
    assert diff.max() <= 1.0000001e-6 and np.count_nonzero(diff) <= max(1, rc.size // 1000)
    M.close()
    P.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind,model", [("fan", "table"), ("fan", "cubic"), ("parallel2d", "table"),
                                        ("lf", "lf")])
def test_sparse_products_vs_projector(lib, kind, model, prec):
    """Exact-coefficient matrix: M x and M' y equal the plan's projections;
    <M x, y> = <x, M' y>."""
    from yoricktrtt_amd._lib import LF2DParameters
    from yoricktrtt_amd.projectors import LFProjector, Projector

    _say("sparse products", kind, model, prec)
    V = 36
    if kind == "lf":
        P = LFProjector(LF2DParameters(0.75, 0.75, 0.5, 200.0, 100.0, 0.0, 0.0, 160, 150, 230),
                        np.arange(V) * (2 * math.pi / V), precision=prec)
    else:
        P = Projector(_geometry(kind, 160, 150, 230, V, Rsc=200.0, Rsd=300.0), model=model,
                      precision=prec)
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device="cuda:0").manual_seed(4)
    x = torch.rand(P.vol_shape, generator=g, device="cuda:0", dtype=dt)
    y = torch.rand(P.sino_shape, generator=g, device="cuda:0", dtype=dt)
    M = P.sparse_matrix(-1)
    assert M.nnz > 0
    tol = 1e-12 if prec == "f64" else 1e-5
    fx, bx = P.forward(x), P.backward(y)
    mx, my = M.mvmult(x), M.mvmult(y, 1)
    assert rel_err(mx.double().cpu().numpy(), fx.double().cpu().numpy()) < tol
    assert rel_err(my.double().cpu().numpy(), bx.double().cpu().numpy()) < tol
    # accumulation into given outputs
    acc = torch.ones_like(fx)
    M.forward(x, acc)
    assert rel_err((acc - 1).double().cpu().numpy(), mx.double().cpu().numpy()) < 10 * tol
    a = float((mx.double() * y.double()).sum())
    b = float((x.double() * my.double()).sum())
    assert abs(a - b) <= tol * abs(a)
    M.close()
    P.close()


def test_sparse_rejects_unsupported_plans(lib):
    from yoricktrtt_amd import _lib
    from yoricktrtt_amd.geometry import cone_geometry
    from yoricktrtt_amd.projectors import Projector

    _say("sparse rejects")
    geo = cone_geometry(8, 8, 8, 1.0, 10, 10, 1.0, 1.0, np.arange(3) * 1.0, 50.0, 80.0)
    P = Projector(geo, model="cubic", precision="f32")
    assert not lib.trtt_hip_sparse_create(C.c_void_p(P.handle), 6)
    assert "2-D plans" in _lib.last_error()
    P.close()
    geo2 = _geometry("fan", 10, 10, 15, 4)
    P2 = Projector(geo2, model="table", precision="f64", nt=2, t0=[0, 1, 0, 1], t1=[0, 1, 0, 1],
                   w0=[1.0] * 4, w1=[0.0] * 4)
    assert not lib.trtt_hip_sparse_create(C.c_void_p(P2.handle), 6)
    assert "one phase" in _lib.last_error()
    P3 = Projector(geo2, model="table", precision="f64")
    assert not lib.trtt_hip_sparse_create(C.c_void_p(P3.handle), 3)
    assert "decimals" in _lib.last_error()
    P2.close()
    P3.close()


def test_sparse_matrix_without_entries(lib):
    """A detector that no footprint reaches: an empty matrix whose products
    leave their outputs unchanged."""
    from yoricktrtt_amd.projectors import Projector

    _say("sparse empty")
    # parallel beam with the detector shifted far off the object
    geo = _geometry("parallel2d", 12, 10, 8, 3, voff=np.array([200.0, 200.0, 200.0]))
    P = Projector(geo, model="table", precision="f64")
    M = P.sparse_matrix(6)
    assert M.nnz == 0
    coefs, rows, cols = M.coo()
    assert coefs.size == rows.size == cols.size == 0
    x = torch.rand(P.vol_shape, dtype=torch.float64, device="cuda:0")
    y = torch.full(P.sino_shape, 2.0, dtype=torch.float64, device="cuda:0")
    M.forward(x, y)
    assert torch.all(y == 2.0)
    g = torch.full(P.vol_shape, 3.0, dtype=torch.float64, device="cuda:0")
    M.backward(y, g)
    assert torch.all(g == 3.0)
    M.close()
    P.close()
