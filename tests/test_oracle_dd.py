"""Pins the oracle's distance-driven projector (trtt_DD_PB3D) before it is
trusted: golden vectors of the reference C (tests/golden/make_golden.py,
forward and the reference's transpose), and the adjoint identity for the
exact adjoint.  The reference transpose is shown not to be the adjoint once
footprints span three or more pixels along v (trtt_3D_projectors.c:909-916).
"""
import numpy as np
import pytest

from conftest import GOLDEN, rel_err

CASES = ["fine", "coarse", "edge"]


@pytest.fixture(scope="module")
def dd():
    import os

    return dict(np.load(os.path.join(GOLDEN, "dd3d.npz")))


def _args(g, name):
    nu, nv = (int(v) for v in g[f"{name}/dims"])
    us, vs = (float(v) for v in g[f"{name}/scl"])
    return nu, nv, us, vs


@pytest.mark.parametrize("name", CASES)
def test_oracle_dd_matches_reference(oracle, dd, name):
    vol = dd["vol"]
    nz, ny, nx = vol.shape
    nu, nv, us, vs = _args(dd, name)
    for v in range(dd[f"{name}/xos"].shape[0]):
        xos, xsd = dd[f"{name}/xos"][v], dd[f"{name}/xsd"][v]
        f = oracle.dd_pb3d(np.zeros((nv, nu)), vol, nx, ny, nz, 1.0, nu, nv, us, vs, xos, xsd, 0)
        b = oracle.dd_pb3d(np.zeros((nz, ny, nx)), dd[f"{name}/sino_in"][v], nx, ny, nz, 1.0, nu,
                           nv, us, vs, xos, xsd, 1)
        assert rel_err(f, dd[f"{name}/fwd"][v]) < 1e-14
        assert rel_err(b, dd[f"{name}/bwd"][v]) < 1e-14


@pytest.mark.parametrize("name", CASES)
def test_oracle_dd_adjoint(oracle, dd, name):
    vol = dd["vol"]
    nz, ny, nx = vol.shape
    nu, nv, us, vs = _args(dd, name)
    for v in range(dd[f"{name}/xos"].shape[0]):
        xos, xsd = dd[f"{name}/xos"][v], dd[f"{name}/xsd"][v]
        y = dd[f"{name}/sino_in"][v]
        ax = dd[f"{name}/fwd"][v]
        aty = oracle.dd_pb3d(np.zeros((nz, ny, nx)), y, nx, ny, nz, 1.0, nu, nv, us, vs, xos, xsd,
                             1, adjoint=True)
        lhs, rhs = float((ax * y).sum()), float((vol * aty).sum())
        assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
        ref = float((vol * dd[f"{name}/bwd"][v]).sum())
        if name == "coarse":
            # footprints of at most two pixels along v: the two transposes agree
            assert abs(lhs - ref) <= 1e-13 * abs(lhs)
        elif name == "fine":
            assert abs(lhs - ref) > 0.1 * abs(lhs)
