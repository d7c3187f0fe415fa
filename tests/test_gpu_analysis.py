"""GPU HMH and group index (rows a8, a15) against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
import omx

pytestmark = pytest.mark.gpu

P_WG = [4.2, 2.1, 1.0, 0.6]
W0 = 1 / 1.55


def _sym_field(og, seed):
    """A smooth symmetric per-pixel 3x3 field (not diagonal, not eps^-1)."""
    rng = np.random.default_rng(seed)
    xs, ys = og.x(), og.y()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    B = rng.standard_normal((3, 3))
    A = np.einsum("ab,xy->abxy", B + B.T, 0.3 + np.exp(-(X ** 2 + Y ** 2)))
    A += np.einsum("ab,xy->abxy", np.eye(3), 1.5 + 0.2 * np.cos(X))
    return A


def test_hmh_matches_oracle_and_rayleigh_quotient():
    print("HMH vs oracle", flush=True)
    og = O.Grid(6.0, 4.0, 24, 16)
    g = omx.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi(P_WG, og)
    k = 1.2
    solver = omx.HIPSolver(precision="fp64")
    Mg = omx.HelmholtzMap(k, ei, g, solver)
    rng = np.random.default_rng(5)
    H = rng.standard_normal((2 * og.N, 3)) + 1j * rng.standard_normal((2 * og.N, 3))
    A = _sym_field(og, 1)
    ref = np.array([O.HMH(H[:, j], k, A, og) for j in range(3)])
    assert np.allclose(Mg.HMH(H, A), ref, rtol=1e-12, atol=0)
    # A = eps^-1: <H, M H>; on a unit eigenvector it is the eigenvalue (runtests.jl:61)
    ref_e = np.array([O.HMH(H[:, j], k, ei, og) for j in range(3)])
    assert np.allclose(Mg.HMH(H), ref_e, rtol=1e-12, atol=0)
    w2, evs = omx.solve_omega2(k, ei, g, solver, nev=2, tol=1e-11, maxiter=300)
    for lam, v in zip(w2, evs):
        assert Mg.HMH(v) == pytest.approx(lam, rel=1e-11)


def test_group_index_matches_oracle():
    print("group index vs oracle", flush=True)
    og = O.Grid(6.0, 4.0, 24, 16)
    g = omx.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi(P_WG, og)
    solver = omx.HIPSolver(precision="fp64")
    km, evs = omx.solve_k(W0, ei, g, solver, nev=2, k_tol=1e-12, eig_tol=1e-11)
    # non-dispersive: n_g = ω / HMkH
    ng0 = omx.group_index(km, evs, W0, ei, None, g, solver)
    zero = np.zeros_like(ei)
    ref0 = np.array([O.group_index(k, v, W0, ei, zero, og) for k, v in zip(km, evs)])
    assert np.allclose(ng0, ref0, rtol=1e-11, atol=0)
    # a synthetic dispersive term: symmetric and dominated by a positive multiple of ε
    eps = O.sliceinv_3x3(ei)
    deps = 0.05 * eps + 0.005 * _sym_field(og, 2)
    ng = omx.group_index(km, evs, W0, ei, deps, g, solver)
    ref = np.array([O.group_index(k, v, W0, ei, deps, og) for k, v in zip(km, evs)])
    assert np.allclose(ng, ref, rtol=1e-11, atol=0)
    assert np.all(ng > ng0)  # positive material dispersion raises the group index
    # single-mode call form
    assert omx.group_index(km[0], evs[0], W0, ei, deps, g, solver) == pytest.approx(ng[0], rel=1e-14)
