"""GPU adjoint of solve_k (rrule pull-back) against the oracle's pull-back on the same
primal solution, and against central finite differences of the GPU solve itself."""
import numpy as np
import pytest

import oracle as O
import omx

pytestmark = pytest.mark.gpu
P_WG = [4.2, 2.1, 1.0, 0.6]
W0 = 1 / 1.55


def _setup(nx=16, ny=12, nev=2):
    og = O.Grid(6.0, 4.0, nx, ny)
    g = omx.Grid(6.0, 4.0, nx, ny)
    ei = O.gaussian_wg_epsi(P_WG, og)
    solver = omx.HIPSolver()
    (km, evs), pb = omx.rrule_solve_k(W0, ei, g, solver, nev=nev, k_tol=1e-13, eig_tol=1e-11)
    return og, g, ei, solver, km, evs, pb


def test_adjoint_kbar_channel_matches_oracle():
    print("adjoint kbar", flush=True)
    og, g, ei, solver, km, evs, pb = _setup(nev=2)
    kbar = [1.0, -0.5]
    _, wbar, eibar, _, _ = pb((kbar, None))
    wr, er = O.solve_k_pullback(W0, ei, og, km, evs, kbar, None)
    assert wbar == pytest.approx(wr, rel=1e-9)
    assert np.allclose(eibar, er, rtol=0, atol=1e-9 * np.abs(er).max())
    assert np.allclose(eibar, np.transpose(eibar, (1, 0, 2, 3)))


def test_adjoint_evbar_channel_matches_oracle():
    print("adjoint evbar", flush=True)
    og, g, ei, solver, km, evs, pb = _setup(nev=2)
    rng = np.random.default_rng(3)
    c = [rng.standard_normal(2 * og.N) + 1j * rng.standard_normal(2 * og.N) for _ in range(2)]
    evbar = [2 * c[j] * np.vdot(c[j], evs[j]) for j in range(2)]   # d|<ev,c>|^2 / d conj(ev)
    kbar = [0.3, None]
    _, wbar, eibar, _, _ = pb((kbar, evbar))
    wr, er = O.solve_k_pullback(W0, ei, og, km, evs, kbar, evbar, adj_rtol=1e-13)
    assert wbar == pytest.approx(wr, rel=1e-6)
    assert np.allclose(eibar, er, rtol=0, atol=1e-6 * np.abs(er).max())


def test_adjoint_dk_domega_fd_through_gpu():
    print("adjoint fd", flush=True)
    og, g, ei, solver, km, evs, pb = _setup(nev=1)
    _, wbar, _, _, _ = pb(([1.0], None))
    h = 1e-4
    kp, _ = omx.solve_k(W0 + h, ei, g, solver, nev=1, k_tol=1e-13, eig_tol=1e-11)
    kmn, _ = omx.solve_k(W0 - h, ei, g, solver, nev=1, k_tol=1e-13, eig_tol=1e-11)
    fd = (kp[0] - kmn[0]) / (2 * h)
    assert wbar == pytest.approx(fd, rel=1e-6)        # runtests.jl:91-94 asks 1e-4
    # n_g bound (runtests.jl:96)
    assert np.sqrt(P_WG[1]) < wbar < 1.5 * np.sqrt(P_WG[0])
