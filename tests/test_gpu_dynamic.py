"""GPU 2-D+t dynamic projector C_k = A_k B_k (Tomography_2D_dyn): the
temporal interpolator against numpy, and the combined operator against the
oracle view by view on the interpolated frames.

Tolerances: FP64 1e-12 relative; the adjoint identity 1e-12 (FP64).
"""
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


def _dev(a):
    return torch.tensor(a, dtype=torch.float64, device="cuda:0")


@pytest.mark.parametrize("nt,deg", [(6, 1), (6, 3), (2, 3), (5, 2)])
def test_time_interpolator_vs_numpy(lib, nt, deg):
    from yoricktrtt_amd.projectors import TimeInterpolator, temporal_weights

    _say("tinterp", nt, deg)
    rng = np.random.default_rng(nt * 10 + deg)
    V, shape = 23, (13, 17)
    t = rng.uniform(0, nt, size=V)
    idx, w = temporal_weights(t, nt, deg)
    B = TimeInterpolator(shape, nt, idx, w, precision="f64")
    x = rng.uniform(size=(nt,) + shape)
    y = rng.uniform(size=(V,) + shape)
    fx = B.forward(_dev(x)).cpu().numpy()
    by = B.backward(_dev(y)).cpu().numpy()
    ref_f = np.zeros((V,) + shape)
    ref_b = np.zeros((nt,) + shape)
    for k in range(V):
        for q in range(deg + 1):
            ref_f[k] += w[k, q] * x[idx[k, q]]
            ref_b[idx[k, q]] += w[k, q] * y[k]  # repeated phases add up
    assert rel_err(fx, ref_f) < 1e-12
    assert rel_err(by, ref_b) < 1e-12
    a, b = float((fx * y).sum()), float((x * by).sum())
    assert abs(a - b) <= 1e-12 * abs(a)
    B.close()


@pytest.mark.parametrize("deg", [1, 3])
@pytest.mark.parametrize("kind", ["fan", "parallel2d"])
def test_dynamic_projector_vs_oracle(lib, oracle, kind, deg):
    from yoricktrtt_amd.geometry import fan_geometry_2d, parallel_geometry_2d
    from yoricktrtt_amd.projectors import DynamicProjector, temporal_weights

    _say("dynamic", kind, deg)
    nx, ny, nv, V, nt = 31, 27, 45, 40, 6
    # views over half a turn spread over three cycles of the motion
    # (Ytrtt2D_system_dyn.i:500-522: t_index = (t % tcycl) nt / tcycl)
    th = np.arange(V) * (math.pi / V)
    tac = np.arange(V) * (3.0 / (V - 1))
    t_index = (tac % 1.0) * nt
    if kind == "fan":
        geo = fan_geometry_2d(nx, ny, 1.0, nv, 1.0, th, 70.0, 110.0)
    else:
        geo = parallel_geometry_2d(nx, ny, 1.0, nv, 1.0, th)
    P = DynamicProjector(geo, nt, t_index, deg, model="cubic", precision="f64")
    rng = np.random.default_rng(deg)
    x = rng.uniform(size=(nt, ny, nx))
    y = rng.uniform(size=(V, nv))
    fwd = P.forward(_dev(x)).cpu().numpy()
    bwd = P.backward(_dev(y)).cpu().numpy()
    idx, w = temporal_weights(t_index, nt, deg)
    fan = kind == "fan"
    rb = np.zeros((nt, ny, nx))
    for k in range(V):
        xk = sum(w[k, q] * x[idx[k, q]] for q in range(deg + 1))
        rf = oracle.sb2d(np.zeros(nv), xk, nx, ny, 1.0, nv, 1.0, geo.xos[k], geo.xsd[k], fan, 0)
        assert rel_err(fwd[k], rf) < 1e-12, k
        bk = oracle.sb2d(np.zeros((ny, nx)), y[k], nx, ny, 1.0, nv, 1.0, geo.xos[k], geo.xsd[k],
                         fan, 1)
        for q in range(deg + 1):
            rb[idx[k, q]] += w[k, q] * bk
    assert rel_err(bwd, rb) < 1e-12
    P.close()


@pytest.mark.parametrize("op", ["sparse", "dynamic"])
def test_data_term_graph_replay_sparse_and_dynamic(lib, op):
    """The data term over the sparse matrix or the dynamic projector captured
    in one HIP graph: every replay, with eager work in between, equals the
    eager evaluation (FP32, 1e-6)."""
    from conftest import rel_err as rerr
    from yoricktrtt_amd.geometry import fan_geometry_2d
    from yoricktrtt_amd.projectors import DynamicProjector, Projector
    from yoricktrtt_amd.reconstruction import DataTerm

    _say("graph replay", op)
    nx, ny, nv, V, nt = 48, 44, 70, 60, 5
    th = np.arange(V) * (2 * math.pi / V)
    geo = fan_geometry_2d(nx, ny, 1.0, nv, 1.0, th, 90.0, 140.0)
    rng = np.random.default_rng(11)
    if op == "sparse":
        P = Projector(geo, model="table", precision="f32")
        A = P.sparse_matrix(6)
        shape = (ny, nx)
    else:
        t_index = (np.arange(V) * (2.7 / (V - 1)) % 1.0) * nt
        A = DynamicProjector(geo, nt, t_index, 3, model="cubic", precision="f32")
        shape = (nt, ny, nx)
    y = torch.tensor(rng.uniform(size=(V, nv)), dtype=torch.float32, device="cuda")
    w = torch.tensor(rng.uniform(0.5, 1.5, size=(V, nv)), dtype=torch.float32, device="cuda")
    D = DataTerm(A, y, w)
    xs = [torch.tensor(rng.uniform(size=shape), dtype=torch.float32, device="cuda")
          for _ in range(3)]
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
            assert rerr(g_static.double().cpu().numpy(), ge.double().cpu().numpy()) < 1e-6
    A.close()
    if op == "sparse":
        P.close()
