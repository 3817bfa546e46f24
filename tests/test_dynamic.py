"""Taps of the temporal interpolation B_k (host side, CPU only).

Degree 1 is pinned by the reference's linear interpolation
(Ytrtt_3D_projectors.i:301-311: phases floor(t) and ceil(t) modulo nt,
weights 1 - frac and frac).  Degrees 2 and 3 stand in for SPL's
spl_interp_coefs_new_spline, which is not in the reference tree: parity
unpinned; they are checked against the B-spline definition (partition of
unity, the values of beta^d, symmetry, reproduction of linear functions).
"""
import math

import numpy as np
import pytest

from yoricktrtt_amd.projectors import temporal_weights


def test_linear_taps_are_the_reference_interpolation():
    nt = 7
    t = np.array([0.0, 0.25, 1.0, 3.5, 5.999, 6.0, 6.75])
    idx, w = temporal_weights(t, nt, 1)
    for k, tk in enumerate(t):
        lo, hi = math.floor(tk) % nt, math.ceil(tk) % nt
        f = tk - math.floor(tk)
        dense = np.zeros(nt)
        np.add.at(dense, idx[k], w[k])
        ref = np.zeros(nt)
        ref[lo] += 1.0 - f
        ref[hi] += f
        assert np.allclose(dense, ref, atol=1e-15), tk


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_partition_of_unity_and_linear_reproduction(deg):
    nt = 10
    t = np.random.default_rng(1).uniform(0, nt, size=200)
    idx, w = temporal_weights(t, nt, deg)
    assert idx.shape == (200, deg + 1)
    assert np.allclose(w.sum(axis=1), 1.0, atol=1e-14)
    assert np.all(w >= 0)
    # unwrapped phases: sum_j w_j j = t (B-splines reproduce linear functions)
    j0 = np.floor(t - 0.5 * (deg + 1)) + 1
    j = j0[:, None] + np.arange(deg + 1)[None, :]
    assert np.allclose((w * j).sum(axis=1), t, atol=1e-12)


def test_cubic_taps_values_and_wrap():
    idx, w = temporal_weights([0.0, 2.5], 4, 3)
    # t = 0: phases -1, 0, 1, 2 -> 3, 0, 1, 2 with beta3(1), beta3(0), beta3(1), beta3(2)
    assert list(idx[0]) == [3, 0, 1, 2]
    assert np.allclose(w[0], [1 / 6, 2 / 3, 1 / 6, 0.0], atol=1e-15)
    # t = 2.5: phases 1..4 -> 1, 2, 3, 0; symmetric weights
    assert list(idx[1]) == [1, 2, 3, 0]
    assert np.allclose(w[1], w[1][::-1], atol=1e-15)
    assert np.allclose(w[1], [1 / 48, 23 / 48, 23 / 48, 1 / 48], atol=1e-15)


def test_out_of_range_time_is_refused():
    with pytest.raises(ValueError):
        temporal_weights([-0.1], 5, 1)
    with pytest.raises(ValueError):
        temporal_weights([1.0], 5, 4)


def test_data_term_converts_and_checks_its_data():
    """DataTerm stores y and w in the plan's type and refuses another shape
    (the residual kernel reads them as the plan's type and shape); the
    projector's own checks refuse host tensors and a non-FP64 fx."""
    import torch

    from yoricktrtt_amd.projectors import Projector
    from yoricktrtt_amd.reconstruction import DataTerm

    class Stub:
        sino_shape = (3, 4, 5)
        dtype = torch.float32

    D = DataTerm(Stub(), torch.ones((3, 4, 5), dtype=torch.float64),
                 torch.ones((3, 4, 5), dtype=torch.float64))
    assert D.y.dtype == torch.float32 and D.w.dtype == torch.float32
    with pytest.raises(ValueError, match="shape"):
        DataTerm(Stub(), torch.ones((3, 4, 6)))
    with pytest.raises(ValueError, match="shape"):
        DataTerm(Stub(), torch.ones((3, 4, 5)), torch.ones((3, 5, 4)))
    P = Projector.__new__(Projector)  # argument checks only, no plan
    P.torch, P.dtype = torch, torch.float32
    with pytest.raises(TypeError):
        P._check(torch.ones((3, 4, 5)), (3, 4, 5))
    with pytest.raises(ValueError, match="fx"):
        P._check_fx(torch.zeros(1, dtype=torch.float32))
