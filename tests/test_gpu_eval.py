"""Host-array evaluation of the data term (trtt_hip_eval) and the cached
one-view plans of the drop-in symbols.

* trtt_hip_eval against the FP64 oracle loop of Ytrtt4D_optim.i:418-446 on a
  two-phase cone plan (one-hot and interpolated views, weights), in both
  gradient forms: sum A'W(Ax - y) and the reference's sum A'WAx - WCyT;
  1e-5 (FP32) / 1e-12 (FP64) on f and g; and the device data term
  (DataTerm) on the same plan agrees with it.
* Drop-in symbols called view after view, alternating geometry families
  (more families than the cache holds), and a footprint table changed in
  place between calls: every result equals the oracle of its own call.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def lib():
    from yoricktrtt_amd import _lib

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _lib.load()


def _p(a):
    from yoricktrtt_amd import _lib

    return a.ctypes.data_as(_lib.dp)


def _two_phase(seed):
    from yoricktrtt_amd.geometry import cone_geometry

    nx, ny, nz, nu, nv, V, nt = 30, 28, 26, 36, 40, 10, 2
    rng = np.random.default_rng(seed)
    geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, np.linspace(0, 5.8, V), 90.0, 140.0)
    t = rng.uniform(0, nt - 1, size=V)
    t[:4] = np.round(t[:4])
    t0 = np.floor(t).astype(np.int32)
    t1 = np.minimum(t0 + 1, nt - 1).astype(np.int32)
    w1 = t - t0
    w1[t0 == t1] = 0.0
    w0 = 1.0 - w1
    return geo, nt, t0, t1, w0, w1, rng


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("split", [False, True])
def test_eval_host_vs_oracle(lib, oracle, prec, split):
    from yoricktrtt_amd.projectors import Projector
    from yoricktrtt_amd.reconstruction import DataTerm

    print(f"[gpu] eval host {prec} split={split}", flush=True)
    geo, nt, t0, t1, w0, w1, rng = _two_phase(31)
    nx, ny, nz, nu, nv, V = geo.nx, geo.ny, geo.nz, geo.nu, geo.nv, geo.nviews
    x = rng.uniform(size=(nt, nz, ny, nx))
    # oracle loop
    ax = np.zeros((V, nv, nu))
    for k in range(V):
        xk = w0[k] * x[t0[k]] + w1[k] * x[t1[k]]
        oracle.sb3d(ax[k], xk, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[k], geo.xsd[k], True, 0)
    y = ax * rng.uniform(0.5, 1.5, size=ax.shape)
    w = rng.uniform(0.5, 2.0, size=ax.shape)
    res = ax - y
    f_ref = 0.5 * float(np.sum(w * res * res))
    g_ref = np.zeros_like(x)
    for k in range(V):
        b = np.zeros((nz, ny, nx))
        oracle.sb3d(b, w[k] * res[k], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, geo.xos[k], geo.xsd[k],
                    True, 1)
        g_ref[t0[k]] += w0[k] * b
        g_ref[t1[k]] += w1[k] * b
    P = Projector(geo, model="cubic", precision=prec, nt=nt, t0=t0, t1=t1, w0=w0, w1=w1)
    P.eval_set_data(y, w)
    tol = 1e-12 if prec == "f64" else 1e-5
    for rep in range(2):  # the second call reuses the device copies
        f, g = P.eval_host(x, split=split)
        ef, eg = abs(f - f_ref) / f_ref, rel_err(g, g_ref)
        print(f"[gpu]   f rel {ef:.2e}, g {eg:.2e}", flush=True)
        assert ef < tol and eg < tol
    # the device data term on the same plan
    dt = torch.float64 if prec == "f64" else torch.float32
    D = DataTerm(P, torch.tensor(y, dtype=dt, device="cuda"), torch.tensor(w, dtype=dt, device="cuda"))
    gd = torch.zeros(P.vol_shape, dtype=dt, device="cuda")
    fd = D.eval(torch.tensor(x, dtype=dt, device="cuda"), gd)
    assert abs(fd - f) / f_ref < tol and rel_err(gd.double().cpu().numpy(), g_ref) < tol
    # unit weights
    P.eval_set_data(y)
    f1, g1 = P.eval_host(x, split=split)
    assert abs(f1 - 0.5 * float(np.sum(res * res))) / f1 < tol
    P.close()


def test_eval_host_rejects_missing_data(lib):
    from yoricktrtt_amd.projectors import Projector

    geo, nt, t0, t1, w0, w1, rng = _two_phase(32)
    P = Projector(geo, model="cubic", precision="f32", nt=nt, t0=t0, t1=t1, w0=w0, w1=w1)
    with pytest.raises(RuntimeError, match="eval_set_data"):
        P.eval_host(np.zeros(P.vol_shape))
    P.close()


def test_dropin_plan_cache(lib, oracle):
    """Ten geometry families (the cache holds eight) called in turns, views
    changing every call, both directions; then a table rewritten in place."""
    from yoricktrtt_amd.geometry import cone_geometry, footprint_table, parallel_geometry_3d

    print("[gpu] drop-in plan cache", flush=True)
    rng = np.random.default_rng(5)
    fams = []
    for f in range(10):
        nx, ny, nz = 14 + f, 12 + (f % 3), 10 + (f % 4)
        nu, nv = 20 + f, 22 + f
        th = rng.uniform(0, 2 * math.pi, size=3)
        if f % 2 == 0:
            geo = cone_geometry(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th, 60.0, 95.0)
        else:
            geo = parallel_geometry_3d(nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, th)
        fams.append(geo)
    fp, cum, coord, step = footprint_table(5000)
    for rnd in range(2):
        for f, geo in enumerate(fams):
            cone = f % 2 == 0
            for v in range(geo.nviews):
                vol = rng.uniform(size=(geo.nz, geo.ny, geo.nx))
                det = rng.uniform(size=(geo.nv, geo.nu))
                xos, xsd = np.ascontiguousarray(geo.xos[v]), np.ascontiguousarray(geo.xsd[v])
                table = (f % 3 == 0)
                tab = (cum, coord, step) if table else None
                for job, inp, shape in ((0, vol, (geo.nv, geo.nu)), (1, det, (geo.nz, geo.ny, geo.nx))):
                    out = np.zeros(shape)
                    if table:
                        fn = lib.trtt_CB3D if cone else lib.trtt_PB3D
                        r = fn(_p(out), _p(inp), geo.nx, geo.ny, geo.nz, 1.0, 3, _p(fp), _p(cum),
                               cum.size, step, _p(coord), geo.nu, geo.nv, 1.0, 1.0, _p(xos),
                               _p(xsd), job)
                    else:
                        fn = lib.trtt_CB3D_cubic if cone else lib.trtt_PB3D_cubic
                        r = fn(_p(out), _p(inp), geo.nx, geo.ny, geo.nz, 1.0, 3, geo.nu, geo.nv,
                               1.0, 1.0, _p(xos), _p(xsd), job)
                    assert r == 0
                    ref = oracle.sb3d(np.zeros(shape), inp, geo.nx, geo.ny, geo.nz, 1.0, geo.nu,
                                      geo.nv, 1.0, 1.0, xos, xsd, cone, job, tab)
                    assert rel_err(out, ref) < 1e-5, (rnd, f, v, job)
    # the same table array, new content: the cached plan must not be reused
    geo = fams[0]
    vol = rng.uniform(size=(geo.nz, geo.ny, geo.nx))
    xos, xsd = np.ascontiguousarray(geo.xos[0]), np.ascontiguousarray(geo.xsd[0])
    for scale in (1.0, 0.5):
        cum *= scale
        out = np.zeros((geo.nv, geo.nu))
        assert lib.trtt_CB3D(_p(out), _p(vol), geo.nx, geo.ny, geo.nz, 1.0, 3, _p(fp), _p(cum),
                             cum.size, step, _p(coord), geo.nu, geo.nv, 1.0, 1.0, _p(xos), _p(xsd),
                             0) == 0
        ref = oracle.sb3d(np.zeros((geo.nv, geo.nu)), vol, geo.nx, geo.ny, geo.nz, 1.0, geo.nu,
                          geo.nv, 1.0, 1.0, xos, xsd, True, 0, (cum, coord, step))
        assert rel_err(out, ref) < 1e-5, scale
