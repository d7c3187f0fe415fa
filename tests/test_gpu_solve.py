"""GPU eigensolves (solve_omega2, solve_k) against dense brute force and the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import omx

pytestmark = pytest.mark.gpu

P_WG = [4.2, 2.1, 1.0, 0.6]
W0 = 1 / 1.55


def _dense(k, ei, og):
    op = O.HelmholtzOracle(k, ei, og)
    n2 = 2 * og.N
    A = np.stack([op.M(np.eye(n2)[:, j]) for j in range(n2)], 1)
    return (A + A.conj().T) / 2


@pytest.mark.parametrize("precision,nev", [("mixed", 4), ("fp64", 4), ("mixed", 16), ("fp64", 13)])
def test_solve_omega2_matches_dense(precision, nev):
    """nev = 16 / 13 exercise the widest blocks: 48 / 39 Rayleigh-Ritz columns (three Gram
    tiles, several Jacobi items per thread, an odd Jacobi size)."""
    print(f"solve_omega2 dense {precision} nev={nev}", flush=True)
    og = O.Grid(6.0, 4.0, 24, 16)
    g = omx.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi(P_WG, og)
    k = 1.1
    w = np.linalg.eigvalsh(_dense(k, ei, og))[:nev]
    solver = omx.HIPSolver(precision=precision)
    w2, evs, info = omx.solve_omega2(k, ei, g, solver, nev=nev, tol=1e-10, maxiter=300, return_info=True)
    print(info, flush=True)
    assert np.allclose(w2, w, rtol=1e-11, atol=0)
    ref = O.HelmholtzOracle(k, ei, og)
    for lam, v in zip(w2, evs):
        assert np.linalg.norm(v) == pytest.approx(1.0, abs=1e-12)
        assert np.linalg.norm(ref.M(v) - lam * v) < 1e-9


@pytest.mark.parametrize("precision", ["mixed", "fp64", "plain"])
def test_solve_k_matches_oracle(precision):
    """mixed / fp64: the optimised path; plain: the reference's algorithm on the same kernels
    (the bench's comparison baseline)."""
    print(f"solve_k oracle {precision}", flush=True)
    og = O.Grid(6.0, 4.0, 24, 16)
    g = omx.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi(P_WG, og)
    km_ref, ev_ref = O.solve_k(W0, ei, og, nev=2, k_tol=1e-12, eig_tol=1e-11)
    solver = (omx.HIPSolver(precision="fp64", plain=True) if precision == "plain"
              else omx.HIPSolver(precision=precision))
    km, evs, info = omx.solve_k(W0, ei, g, solver, nev=2, k_tol=1e-12, eig_tol=1e-11, return_info=True)
    print(km, km_ref, info, flush=True)
    assert np.allclose(km, km_ref, rtol=1e-10, atol=0)
    for k, v, vr in zip(km, evs, ev_ref):
        # eigenvalue at the returned k equals ω² (round trip, runtests.jl:80-82)
        lam = np.linalg.eigvalsh(_dense(k, ei, og))
        assert np.min(np.abs(lam - W0 ** 2)) / W0 ** 2 < 1e-10
        # both phase-canonicalised: fields agree directly
        assert abs(abs(np.vdot(v, vr)) - 1) < 1e-9
        assert np.linalg.norm(v - vr) < 1e-5
