"""sliceinv_3x3 and its pull-back on the GPU against the oracle (pinned by numpy's inverse and
by finite differences), on 2D and 3D tensor fields."""
import numpy as np
import pytest

import oracle as O
import omx

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3, 3) + shape) * 0.3
    for a in range(3):
        A[a, a] += 2.0 + a  # well conditioned, not symmetric
    return A


@pytest.mark.parametrize("shape", [(7, 5), (64, 48), (12, 10, 6)], ids=lambda s: "x".join(map(str, s)))
def test_sliceinv_matches_oracle(shape):
    A = _field(shape, 3)
    Y = omx.sliceinv_3x3(A)
    ref = O.sliceinv_3x3(A)
    assert np.max(np.abs(Y - ref)) < 1e-14 * np.max(np.abs(ref))
    # A Y = I per pixel
    Am, Ym = np.moveaxis(A, (0, 1), (-2, -1)), np.moveaxis(Y, (0, 1), (-2, -1))
    assert np.max(np.abs(Am @ Ym - np.eye(3))) < 1e-14


@pytest.mark.parametrize("shape", [(7, 5), (12, 10, 6)], ids=lambda s: "x".join(map(str, s)))
def test_sliceinv_pullback_matches_oracle_and_fd(shape):
    A = _field(shape, 4)
    Ybar = np.random.default_rng(5).standard_normal((3, 3) + shape)
    Y, pb = omx.rrule_sliceinv_3x3(A)
    Abar = pb(Ybar)
    ref = O.sliceinv_3x3_pullback(O.sliceinv_3x3(A), Ybar)
    assert np.max(np.abs(Abar - ref)) < 1e-13 * np.max(np.abs(ref))
    # d/dt <Ybar, inv(A + t dA)> = <Abar, dA>
    dA = np.random.default_rng(6).standard_normal(A.shape)
    h = 1e-6
    fd = (np.sum(Ybar * omx.sliceinv_3x3(A + h * dA)) - np.sum(Ybar * omx.sliceinv_3x3(A - h * dA))) / (2 * h)
    assert abs(fd - np.sum(Abar * dA)) < 1e-7 * abs(fd)
