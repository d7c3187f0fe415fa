"""The allocation-free operator entry point (omx_apply_helmholtz_noalloc) against the
allocating one (omx_apply_helmholtz) through the C ABI: the same kernels on the same inputs,
so the outputs must agree (FP64 to 1e-13, FP32 to 1e-6 relative), on consecutive calls of
decreasing size that share the library's fixed per-device scratch array; a call larger than
that array is refused with OMX_ERR_BAD_SHAPE before any GPU work."""
import ctypes

import numpy as np
import pytest

import oracle as O
import omx
from omx import _lib

pytestmark = pytest.mark.gpu

P_WG = [4.2, 2.1, 1.0, 0.6]
SCRATCH = 64 << 20


def _apply(fn, h, k, H, prec):
    X = np.ascontiguousarray(H.T, dtype=np.complex128)       # (m, 2N), vector j contiguous
    Y = np.full_like(X, np.nan)
    st = fn(h.h, float(k), _lib.dptr(X), _lib.dptr(Y), X.shape[0], prec)
    return st, Y.T


def _relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("size", [(8.0, 4.0, 256, 256), (5.0, 5.0, 60, 30)], ids=["256x256", "60x30"])
def test_noalloc_matches_apply_helmholtz(size):
    Dx, Dy, nx, ny = size
    print(f"noalloc vs apply_helmholtz {nx}x{ny}", flush=True)
    og = O.Grid(Dx, Dy, nx, ny)
    g = omx.Grid(Dx, Dy, nx, ny)
    ei = O.gaussian_wg_epsi(P_WG, og)
    solver = omx.HIPSolver()
    h = solver.handle(g, ei)
    L = _lib.lib()
    N = nx * ny
    rng = np.random.default_rng(nx + ny)
    # calls in a row, each smaller than the one before (largest that fits first)
    m64 = min(9, (SCRATCH - 4096) // (112 * N))
    m32 = min(11, (SCRATCH - 4096) // (88 * N))
    calls = [(1.31, m64, omx.OMX_PREC_FP64, 1e-13), (1.12, 3, omx.OMX_PREC_FP64, 1e-13),
             (1.27, m32, omx.OMX_PREC_FP32, 1e-6), (0.98, 2, omx.OMX_PREC_FP32, 1e-6)]
    for k, m, prec, tol in calls:
        H = rng.standard_normal((2 * N, m)) + 1j * rng.standard_normal((2 * N, m))
        st_ref, Y_ref = _apply(L.omx_apply_helmholtz, h, k, H, prec)
        st, Y = _apply(L.omx_apply_helmholtz_noalloc, h, k, H, prec)
        assert st_ref == 0 and st == 0, L.omx_last_error(h.h)
        err = _relerr(Y, Y_ref)
        print(f"  k={k} m={m} prec={prec}: rel diff {err:.3e}, bit-identical {np.array_equal(Y, Y_ref)}",
              flush=True)
        assert np.all(np.isfinite(Y))
        assert err <= tol
    # against the oracle on the last FP64-sized input as well
    H = rng.standard_normal((2 * N, 2)) + 1j * rng.standard_normal((2 * N, 2))
    st, Y = _apply(L.omx_apply_helmholtz_noalloc, h, 1.2, H, omx.OMX_PREC_FP64)
    assert st == 0
    assert _relerr(Y, O.HelmholtzOracle(1.2, ei, og).M(H)) < 1e-12


def test_noalloc_refuses_calls_beyond_the_scratch_limit():
    print("noalloc size limit", flush=True)
    og = O.Grid(8.0, 4.0, 256, 256)
    g = omx.Grid(8.0, 4.0, 256, 256)
    ei = O.gaussian_wg_epsi(P_WG, og)
    solver = omx.HIPSolver()
    h = solver.handle(g, ei)
    L = _lib.lib()
    N = g.N
    H = np.zeros((2 * N, 10), dtype=np.complex128)
    st, _ = _apply(L.omx_apply_helmholtz_noalloc, h, 1.2, H, omx.OMX_PREC_FP64)   # 112 N m > 64 MiB
    assert st == _lib.OMX_ERR_BAD_SHAPE
    assert b"scratch" in L.omx_last_error(h.h)
    # the handle still works afterwards
    H = np.ones((2 * N, 1), dtype=np.complex128)
    st, Y = _apply(L.omx_apply_helmholtz_noalloc, h, 1.2, H, omx.OMX_PREC_FP64)
    assert st == 0 and np.all(np.isfinite(Y))
