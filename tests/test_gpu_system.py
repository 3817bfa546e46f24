"""GPU projector set of a 4-D system after a save/load round trip
(yoricktrtt_amd.system): the loaded system's C_k = A_k B_k forward and
transpose equal the original system's (1e-14 relative: the run-to-run
spread of the atomic FP64 kernels; the plan inputs themselves are equal bit
for bit, checked on the CPU) and the FP64 oracle view by view on the
linearly interpolated phases (1e-12 relative)."""
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


def test_loaded_system_projectors(lib, oracle, tmp_path):
    from yoricktrtt_amd import system as S
    from yoricktrtt_amd.projectors import TRTT_CONE_BEAM, temporal_weights

    print("[gpu] system save/load projector set", flush=True)
    nx, ny, nz, nt, nu, nv, V = 16, 14, 12, 3, 20, 18, 9
    rng = np.random.default_rng(4)
    sy = S.create_simu_system_4d(rng.uniform(size=(V, nv, nu)), nx, ny, nz, nt, 1.0, nu, nv, 1.0,
                                 1.0, V, 2.0, 5.0, TRTT_CONE_BEAM, Rsc=60.0, Rsd=100.0,
                                 theta=np.linspace(0, 2 * math.pi, V, endpoint=False), v_off=0.4)
    S.save_system(tmp_path / "h.npz", sy)
    lo = S.load_system(tmp_path / "h.npz")
    x = rng.uniform(size=(nt, nz, ny, nx))
    xd = torch.tensor(x, device="cuda")
    outs = []
    for s in (sy, lo):
        P = s.projectors(precision="f64")
        sino = P.forward(xd, torch.zeros(P.sino_shape, dtype=torch.float64, device="cuda"))
        back = P.backward(sino, torch.zeros(P.vol_shape, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        outs.append((sino.cpu().numpy(), back.cpu().numpy()))
        P.close()
    # the same plan inputs; the general FP64 kernels accumulate with
    # atomics, so two runs agree to their run-to-run spread, not bit for bit
    assert rel_err(outs[1][0], outs[0][0]) < 1e-14 and rel_err(outs[1][1], outs[0][1]) < 1e-14
    g = lo.geometry()
    idx, w = temporal_weights(lo.t_index(), nt, 1)
    ref = np.zeros((V, nv, nu))
    for k in range(V):
        xk = w[k, 0] * x[idx[k, 0]] + w[k, 1] * x[idx[k, 1]]
        oracle.sb3d(ref[k], xk, nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, g.xos[k], g.xsd[k], True, 0)
    assert rel_err(outs[1][0], ref) < 1e-12
    # transpose: sum_k B_k' A_k' sino_k
    bref = np.zeros_like(x)
    for k in range(V):
        b = np.zeros((nz, ny, nx))
        oracle.sb3d(b, outs[1][0][k], nx, ny, nz, 1.0, nu, nv, 1.0, 1.0, g.xos[k], g.xsd[k], True, 1)
        bref[idx[k, 0]] += w[k, 0] * b
        bref[idx[k, 1]] += w[k, 1] * b
    assert rel_err(outs[1][1], bref) < 1e-12
