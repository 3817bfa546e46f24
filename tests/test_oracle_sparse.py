"""Pins the oracle's sparse system matrix of the 2-D projectors before it is
trusted as the checker of the device-built matrix.

* columns: the oracle's projection of a one-hot image equals the reference
  trtt_PB2D_sparser / trtt_FB2D_sparser output of that pixel
  (golden vectors from the reference C, tests/golden/make_golden.py);
* assembly (the Yorick loop of Ytrtt_2D_projectors.i:252-350, not runnable
  here): the product of the assembled matrix with an image equals the
  oracle's projection of that image on every view, and its transpose the
  back projection; the 6-decimal form differs from the exact one by at most
  half a unit of the sixth decimal per coefficient.
"""
import numpy as np
import pytest

from conftest import CASES2D, rel_err


def _table(g):
    return (g["cum"], g["coord"], float(g["step"]))


def _column_fn(oracle, g2, gs, name):
    nx, ny, s = int(gs["nx"]), int(gs["ny"]), float(gs["s"])
    nv, vs = int(gs[f"{name}/nv"]), float(gs[f"{name}/vs"])
    fan = name.startswith("fan")
    tab = _table(g2)

    def column(v, i, j):
        img = np.zeros((ny, nx))
        img[j, i] = 1.0
        return oracle.sb2d(np.zeros(nv), img, nx, ny, s, nv, vs, gs[f"{name}/xos"][v],
                           gs[f"{name}/xsd"][v], fan, 0, tab)
    return column


@pytest.mark.parametrize("name", CASES2D)
def test_oracle_columns_match_reference_sparser(oracle, golden2d, golden_sparse2d, name):
    gs = golden_sparse2d
    cols = gs[f"{name}/cols"]
    column = _column_fn(oracle, golden2d, gs, name)
    nview, ny, nx, nv = cols.shape
    assert np.count_nonzero(cols) > 0
    for v in range(nview):
        got = np.array([[column(v, i, j) for i in range(nx)] for j in range(ny)])
        assert rel_err(got, cols[v]) < 1e-14, (name, v)
        # same non-zero pattern: the assembly keeps exactly these entries
        assert np.array_equal(got != 0.0, cols[v] != 0.0)


def _dense(coefs, rows, cols, nrows, ncols):
    m = np.zeros((nrows, ncols))
    np.add.at(m, (rows - 1, cols - 1), coefs)
    return m


@pytest.mark.parametrize("name", CASES2D)
def test_oracle_assembly_is_the_projector(oracle, golden2d, golden_sparse2d, name):
    gs = golden_sparse2d
    nx, ny, s = int(gs["nx"]), int(gs["ny"]), float(gs["s"])
    nv, vs = int(gs[f"{name}/nv"]), float(gs[f"{name}/vs"])
    xos, xsd = gs[f"{name}/xos"], gs[f"{name}/xsd"]
    nview = xos.shape[0]
    fan = name.startswith("fan")
    tab = _table(golden2d)
    column = _column_fn(oracle, golden2d, gs, name)
    exact = oracle.sparse2d_coo(column, nview, nx, ny, decimals=-1)
    text = oracle.sparse2d_coo(column, nview, nx, ny, decimals=6)
    # same entries in the same order; 6-decimal coefficients within half a unit
    assert np.array_equal(exact[1], text[1]) and np.array_equal(exact[2], text[2])
    assert np.abs(exact[0] - text[0]).max() <= 0.5e-6 + 1e-15
    assert np.all(exact[0] != 0.0)
    # rows run view-major, then pixel, then bin (the reference's order)
    key = (exact[2] - 1) + (exact[1] - 1) // nv * (nx * ny)
    assert np.all(np.diff(key) >= 0)
    M = _dense(*exact, nview * nv, nx * ny)
    img = golden2d["img"]
    sino = np.zeros((nview, nv))
    back = np.zeros((ny, nx))
    y = np.random.default_rng(5).uniform(size=(nview, nv))
    for v in range(nview):
        oracle.sb2d(sino[v], img, nx, ny, s, nv, vs, xos[v], xsd[v], fan, 0, tab)
        oracle.sb2d(back, y[v], nx, ny, s, nv, vs, xos[v], xsd[v], fan, 1, tab)
    assert rel_err(M @ img.ravel(), sino.ravel()) < 1e-13
    assert rel_err(M.T @ y.ravel(), back.ravel()) < 1e-13
