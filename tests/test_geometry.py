"""Host Kottke smoothing (omx.geometry): the kernel-value limits of the reference's own
tests (lib/DielectricSmoothing/test/runtests.jl:62-82), the smoothed-map invariants of
its primal test (runtests.jl:112-128), and closed forms of planar interfaces through
the whole supersampling pipeline.  The C ABI's argument checks run without a GPU."""
import ctypes
import os

import numpy as np
import pytest

from omx import _lib
from omx import geometry as G


class _Grid:
    def __init__(self, Dx, Dy, Nx, Ny):
        self.Dx, self.Dy, self.Nx, self.Ny = Dx, Dy, Nx, Ny


def harm(a, b, f=0.5):
    return 1.0 / (f / a + (1 - f) / b)


def test_tau_round_trip():
    rng = np.random.default_rng(0)
    B = rng.standard_normal((3, 3))
    e = B @ B.T + 3 * np.eye(3)
    assert np.allclose(G._tau_inv(G._tau(e)), e, rtol=1e-14, atol=1e-14)


def test_kottke_kernel_limits():
    # runtests.jl:62-82
    e1 = 3.0 * np.eye(3)
    e2 = np.diag([5.0, 4.0, 4.5])
    assert np.allclose(G.kottke_average(e1, e1, (1.0, 0.0), 0.3), e1, rtol=1e-14)
    assert np.allclose(G.kottke_average(e1, e2, (1.0, 0.0), 1.0), e1, rtol=1e-14)
    assert np.allclose(G.kottke_average(e1, e2, (1.0, 0.0), 0.0), e2, rtol=1e-14)
    es = G.kottke_average(e1, e2, (0.0, 1.0), 0.5)
    assert es[1, 1] == pytest.approx(harm(3.0, 4.0), rel=1e-12)       # normal: harmonic
    assert es[0, 0] == pytest.approx((3.0 + 5.0) / 2, rel=1e-12)       # tangential: arithmetic
    assert es[2, 2] == pytest.approx((3.0 + 4.5) / 2, rel=1e-12)


@pytest.mark.parametrize("theta", [0.0, 0.3, 1.1, 2.5, -0.7])
def test_kottke_isotropic_any_normal(theta):
    # isotropic materials: eps = eps_par (I - n n^T) + eps_perp n n^T, eps_perp the harmonic
    # and eps_par the arithmetic mean; the sign of n does not matter
    a, b, f = 2.085, 3.96, 0.37
    n = np.array([np.cos(theta), np.sin(theta), 0.0])
    nn = np.outer(n, n)
    ref = (f * a + (1 - f) * b) * (np.eye(3) - nn) + harm(a, b, f) * nn
    for s in (1, -1):
        got = G.kottke_average(a * np.eye(3), b * np.eye(3), (s * n[0], s * n[1]), f)
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-14)


def test_kottke_rotates_anisotropic_tensors_covariantly():
    # rotating both materials and the normal about z rotates the average the same way
    rng = np.random.default_rng(3)
    B1, B2 = rng.standard_normal((2, 3, 3))
    e1, e2 = B1 @ B1.T + 2 * np.eye(3), B2 @ B2.T + 2 * np.eye(3)
    th = 0.83
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    n = np.array([0.6, -0.8])
    Rn = R[:2, :2] @ n
    a = G.kottke_average(e1, e2, n, 0.41)
    b = G.kottke_average(R @ e1 @ R.T, R @ e2 @ R.T, Rn, 0.41)
    assert np.allclose(R @ a @ R.T, b, rtol=1e-12, atol=1e-13)


def _einv_to_eps(ei):
    return np.linalg.inv(np.moveaxis(ei, (0, 1), (-2, -1)))


def test_uniform_background():
    g = _Grid(2.0, 1.0, 8, 4)
    bg = np.array([[2.0, 0.1, 0], [0.1, 3.0, 0], [0, 0, 4.0]])
    ei = G.smooth_epsinv([], bg, g)
    assert np.allclose(np.moveaxis(ei, (0, 1), (-2, -1)), np.linalg.inv(bg), rtol=1e-14)


def test_axis_aligned_interface_through_pixel_centres():
    """A horizontal and a vertical edge exactly through a row / column of pixel centres:
    half of each cut pixel's samples lie on either side, so the smoothed tensor is the
    normal-harmonic / tangential-arithmetic mean of the two (anisotropic) materials."""
    g = _Grid(4.0, 4.0, 16, 16)
    d = 0.25
    e_core = np.diag([5.0, 4.0, 4.5])
    e_clad = 3.0
    # core occupies y < y_c(row 9) (edge through the row 9 centres), x < x_c(column 11)
    yc, xc = -2.0 + 9 * d, -2.0 + 11 * d
    core = G.Polygon([(-10.0, -10.0), (xc, -10.0), (xc, yc), (-10.0, yc)], e_core)
    eps = _einv_to_eps(G.smooth_epsinv([core], e_clad, g, nsub=8))
    # pixel (4, 9): horizontal edge, normal along y
    p = eps[4, 9]
    assert p[1, 1] == pytest.approx(harm(4.0, 3.0), rel=1e-12)
    assert p[0, 0] == pytest.approx((5.0 + 3.0) / 2, rel=1e-12)
    assert p[2, 2] == pytest.approx((4.5 + 3.0) / 2, rel=1e-12)
    assert abs(p[0, 1]) < 1e-13
    # pixel (11, 3): vertical edge, normal along x
    p = eps[11, 3]
    assert p[0, 0] == pytest.approx(harm(5.0, 3.0), rel=1e-12)
    assert p[1, 1] == pytest.approx((4.0 + 3.0) / 2, rel=1e-12)
    # bulk pixels exact
    assert np.allclose(eps[2, 2], e_core, rtol=1e-14)
    assert np.allclose(eps[14, 14], 3.0 * np.eye(3), rtol=1e-14)


def test_diagonal_interface_normal():
    """A 45-degree edge through a pixel centre: the integer first moment gives the exact
    normal (1, -1)/sqrt(2); the fill fraction counts the samples on or below the diagonal."""
    g = _Grid(4.0, 4.0, 16, 16)
    d = 0.25
    c = -2.0 + 7 * d
    # half plane y <= x - 0 through (c, c), as a large counter-clockwise triangle
    tri = G.Polygon([(c - 20, c - 20), (c + 20, c - 20), (c + 20, c + 20)], 4.0)
    eps = _einv_to_eps(G.smooth_epsinv([tri], 2.0, g, nsub=8))
    f = 36 / 64  # 28 samples strictly below the diagonal and 8 on it
    n = np.array([1.0, -1.0, 0.0]) / np.sqrt(2)
    nn = np.outer(n, n)
    ref = (f * 4.0 + (1 - f) * 2.0) * (np.eye(3) - nn) + harm(4.0, 2.0, f) * nn
    assert np.allclose(eps[7, 7], ref, rtol=1e-12, atol=1e-13)


def test_smoothed_map_invariants():
    # runtests.jl:112-128 on a ridge waveguide
    g = _Grid(6.0, 4.0, 32, 24)
    e_ln = G.EPS_LN
    ei = G.ridge_waveguide(g, 1.7, 0.7, 75.0, G.EPS_SI3N4, eps_sub=G.diag_eps(*e_ln), eps_clad=1.0)
    eps = _einv_to_eps(ei)
    assert np.allclose(eps[16, 4], np.diag(e_ln), rtol=1e-12)        # deep in the substrate
    assert np.allclose(eps[16, 13], G.EPS_SI3N4 * np.eye(3), rtol=1e-12)  # deep in the core
    diags = np.stack([eps[..., i, i] for i in range(3)])
    assert diags.min() >= 1.0 - 1e-12
    assert diags.max() <= max(max(e_ln), G.EPS_SI3N4) + 1e-12
    assert np.abs(ei - np.swapaxes(ei, 0, 1)).max() < 1e-15
    # every tensor is positive definite
    assert np.linalg.eigvalsh(eps).min() > 0


def _call_gpu_smooth(nshapes, nsub, nx=8, ny=8):
    L = _lib.lib()
    nverts = (ctypes.c_int * max(1, nshapes))(*([4] * nshapes))
    verts = np.zeros(8 * max(1, nshapes))
    eps_s = np.tile(np.eye(3).ravel(), max(1, nshapes))
    bg = np.eye(3).ravel().copy()
    out = np.empty(9 * nx * ny)
    return L.omx_smooth_epsinv(0, nx, ny, 1.0, 1.0, nshapes, nverts, _lib.dptr(verts), _lib.dptr(eps_s),
                               _lib.dptr(bg), nsub, _lib.dptr(out))


def test_c_abi_smoothing_rejects_bad_arguments():
    # argument checks run before any device call
    assert _call_gpu_smooth(16, 8) == _lib.OMX_ERR_UNSUPPORTED
    assert _call_gpu_smooth(1, 0) == _lib.OMX_ERR_INVALID_ARG
    assert _call_gpu_smooth(1, 65) == _lib.OMX_ERR_INVALID_ARG
    assert _call_gpu_smooth(1, 4, nx=1) == _lib.OMX_ERR_BAD_SHAPE
    with pytest.raises(ValueError):
        G.smooth_epsinv([], 1.0, _Grid(1.0, 1.0, 4, 4), backend="cpu")


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_gpu_smoothing_without_device_fails_loudly():
    with pytest.raises(_lib.OmxError) as exc:
        G.rect_waveguide(_Grid(4.0, 3.0, 16, 16), backend="gpu")
    assert exc.value.status == _lib.OMX_ERR_NO_DEVICE
