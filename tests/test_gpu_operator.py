"""GPU parity of the fused operator kernels against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
import omx

pytestmark = pytest.mark.gpu

P_WG = [4.2, 2.1, 1.0, 0.6]


def _rand(rng, n, m):
    return rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))


def _aniso(grid, seed=3):
    xs, ys = grid.x(), grid.y()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    bump = np.exp(-((X - 0.3 * grid.Dx / 6) ** 2 / 1.1 + (Y + 0.1 * grid.Dy) ** 2 / 0.5))
    eps = np.zeros((3, 3, grid.Nx, grid.Ny))
    for a, base in enumerate((2.0, 2.3, 2.1)):
        eps[a, a] = base + 1.8 * bump * (1 + 0.1 * a)
    eps[0, 1] = eps[1, 0] = 0.3 * bump * np.sign(X + 1e-3)
    eps[0, 2] = eps[2, 0] = 0.1 * bump
    eps[1, 2] = eps[2, 1] = -0.05 * bump * (Y > 0)
    return O.sliceinv_3x3(eps)


def _relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


SIZES = [(6.0, 4.0, 24, 16), (4.0, 3.0, 12, 10), (6.0, 4.0, 128, 128), (6.0, 4.0, 192, 192),
         (8.0, 4.0, 256, 256), (8.0, 4.0, 384, 384), (8.0, 4.0, 512, 256), (5.0, 5.0, 60, 30),
         (4.0, 4.0, 64, 64), (6.0, 4.0, 128, 256), (8.0, 3.0, 256, 64)]


@pytest.mark.parametrize("size", SIZES, ids=lambda s: f"{s[2]}x{s[3]}")
def test_operator_family_matches_oracle(size):
    print(f"operator parity {size}", flush=True)
    Dx, Dy, nx, ny = size
    og = O.Grid(Dx, Dy, nx, ny)
    g = omx.Grid(Dx, Dy, nx, ny)
    ei = _aniso(og)
    k = 1.37
    rng = np.random.default_rng(nx * 1000 + ny)
    H = _rand(rng, 2 * og.N, 3)
    ref = O.HelmholtzOracle(k, ei, og)
    M_ref, P_ref = ref.M(H), ref.P(H)
    hk_ref = np.array([ref.HMkH(H[:, j]) for j in range(3)])
    solver = omx.HIPSolver()
    for prec, tol in ((omx.OMX_PREC_FP64, 1e-12), (omx.OMX_PREC_FP32, 2e-6)):
        Mg = omx.HelmholtzMap(k, ei, g, solver, precision=prec)
        assert _relerr(Mg * H, M_ref) < tol
        assert _relerr(Mg.precondition(H), P_ref) < tol
        assert np.allclose(Mg.HMkH(H), hk_ref, rtol=10 * tol, atol=0)


def test_fp64_hermitian_and_homogeneous():
    og = O.Grid(6.0, 4.0, 24, 16)
    g = omx.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi(P_WG, og)
    Mg = omx.HelmholtzMap(1.2, ei, g, omx.HIPSolver())
    rng = np.random.default_rng(0)
    v, w = _rand(rng, 2 * og.N, 1)[:, 0], _rand(rng, 2 * og.N, 1)[:, 0]
    assert np.vdot(w, Mg * v) == pytest.approx(np.conj(np.vdot(v, Mg * w)), rel=1e-12)
    # homogeneous medium: plane waves are eigenvectors with |k - G|^2 / eps
    eps = 2.25
    eh = np.zeros((3, 3) + og.shape)
    for a in range(3):
        eh[a, a] = 1 / eps
    Mh = omx.HelmholtzMap(0.9, eh, g, omx.HIPSolver())
    gx, gy = O.gvec(og)
    for (ix, iy, c) in [(0, 0, 0), (3, 5, 1), (23, 15, 0), (12, 8, 1)]:
        h = np.zeros((2,) + og.shape, complex)
        h[c, ix, iy] = 1.0
        vv = O.to_flat(h)
        lam = (gx[ix] ** 2 + gy[iy] ** 2 + 0.81) / eps
        assert np.allclose(Mh * vv, lam * vv, atol=1e-13)
        assert np.allclose(Mh.precondition(Mh * vv), vv, atol=1e-13)


@pytest.mark.parametrize("size", [(4.0, 4.0, 64, 64), (6.0, 4.0, 128, 256), (8.0, 4.0, 256, 256)],
                         ids=lambda s: f"{s[2]}x{s[3]}")
def test_fp32_register_and_stockham_operators_agree(size):
    """The two FP32 operator implementations (register FFT, LDS Stockham FFT) give the same
    M, P and HMkH to FP32 rounding on power-of-two grids."""
    print(f"register vs Stockham {size}", flush=True)
    Dx, Dy, nx, ny = size
    og = O.Grid(Dx, Dy, nx, ny)
    g = omx.Grid(Dx, Dy, nx, ny)
    ei = _aniso(og)
    rng = np.random.default_rng(7)
    H = _rand(rng, 2 * og.N, 4)
    out = {}
    for reg in (True, False):
        solver = omx.HIPSolver(reg_ops=reg)
        Mg = omx.HelmholtzMap(0.9, ei, g, solver, precision=omx.OMX_PREC_FP32)
        out[reg] = (Mg * H, Mg.precondition(H), Mg.HMkH(H))
    assert _relerr(out[True][0], out[False][0]) < 2e-6
    assert _relerr(out[True][1], out[False][1]) < 2e-6
    assert np.allclose(out[True][2], out[False][2], rtol=2e-5, atol=0)


@pytest.mark.parametrize("size", [(4.0, 4.0, 64, 64), (6.0, 4.0, 128, 256), (8.0, 4.0, 256, 256),
                                  (8.0, 4.0, 512, 256)], ids=lambda s: f"{s[2]}x{s[3]}")
def test_register_path_block_shapes_match_oracle(size):
    """The register path with half the rows per block in passes 1/3 (8 in FP32, 4 in FP64)
    and half the threads per block in pass 2 (options reg_rows32, reg_rows64, reg_nt2, which
    the autotuner may pick): M, P and HMkH against the oracle, including the HMkH partial
    sums, whose count follows the rows per block."""
    print(f"register block shapes {size}", flush=True)
    Dx, Dy, nx, ny = size
    og, g = O.Grid(Dx, Dy, nx, ny), omx.Grid(Dx, Dy, nx, ny)
    ei = _aniso(og)
    k = 1.21
    rng = np.random.default_rng(nx + 3 * ny)
    H = _rand(rng, 2 * og.N, 3)
    ref = O.HelmholtzOracle(k, ei, og)
    M_ref, P_ref = ref.M(H), ref.P(H)
    hk_ref = np.array([ref.HMkH(H[:, j]) for j in range(3)])
    for rows32, rows64, nt2 in ((8, 4, 128), (16, 8, 128), (8, 4, 256)):
        solver = omx.HIPSolver()
        solver.options.update(reg_rows32=float(rows32), reg_rows64=float(rows64), reg_nt2=float(nt2))
        for prec, tol in ((omx.OMX_PREC_FP64, 1e-12), (omx.OMX_PREC_FP32, 2e-6)):
            Mg = omx.HelmholtzMap(k, ei, g, solver, precision=prec)
            assert _relerr(Mg * H, M_ref) < tol, (rows32, rows64, nt2, prec)
            assert _relerr(Mg.precondition(H), P_ref) < tol, (rows32, rows64, nt2, prec)
            assert np.allclose(Mg.HMkH(H), hk_ref, rtol=10 * tol, atol=0), (rows32, rows64, nt2, prec)


def test_register_block_shape_options_are_checked():
    g = omx.Grid(4.0, 4.0, 64, 64)
    ei = _aniso(O.Grid(4.0, 4.0, 64, 64))
    solver = omx.HIPSolver()
    h = solver.handle(g, ei)
    L = omx._lib.lib()
    for key, bad in (("reg_rows32", 4), ("reg_rows64", 16), ("reg_nt2", 64)):
        assert L.omx_set_option(h.h, key.encode(), float(bad)) == omx._lib.OMX_ERR_INVALID_ARG
