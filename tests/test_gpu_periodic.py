"""3D cells periodic along z on the GPU: the operator family against the 3D oracle, solve_ω²
and solve_k on a Bragg-grating cell, the period derivatives and the solve_k_periodic
pull-back against the oracle and against finite differences of GPU re-solves."""
import numpy as np
import pytest

import oracle as O
from oracle import periodic as P
import omx

pytestmark = pytest.mark.gpu


def _aniso3(g, seed=5):
    """Smooth full-tensor ε (symmetric positive definite per voxel) varying in x, y and z."""
    X, Y, Z = np.meshgrid(g.x(), g.y(), g.z(), indexing="ij")
    bump = np.exp(-(X ** 2 / 0.6 + Y ** 2 / 0.3)) * (1.0 + 0.3 * np.cos(2 * np.pi * Z / g.Dz))
    eps = np.zeros((3, 3) + g.shape)
    for a, base in enumerate((2.0, 2.3, 2.1)):
        eps[a, a] = base + 1.7 * bump * (1 + 0.1 * a)
    eps[0, 1] = eps[1, 0] = 0.25 * bump * np.sin(2 * np.pi * Z / g.Dz)
    eps[0, 2] = eps[2, 0] = 0.1 * bump
    eps[1, 2] = eps[2, 1] = -0.05 * bump * (Y > 0)
    return np.linalg.inv(eps.transpose(2, 3, 4, 0, 1)).transpose(3, 4, 0, 1, 2).copy()


def _grids(shape, D=(2.4, 2.0, 0.9)):
    return P.Grid3(*D, *shape), omx.Grid3(*D, *shape)


def _relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


SIZES = [(12, 10, 6), (16, 16, 8), (24, 16, 12), (20, 18, 5), (64, 64, 16)]


@pytest.mark.parametrize("slab", [1, 0], ids=["slab", "five_pass"])
@pytest.mark.parametrize("shape", SIZES, ids=lambda s: "x".join(map(str, s)))
def test_operator_family_matches_oracle_3d(shape, slab):
    print(f"3D operator {shape} slab {slab}", flush=True)
    og, g = _grids(shape)
    ei = _aniso3(og)
    rng = np.random.default_rng(2)
    m = 3
    H = rng.standard_normal((2 * og.N, m)) + 1j * rng.standard_normal((2 * og.N, m))
    k = 0.83
    basis = P.mag_mn3(k, og)
    refM = P.apply_M3(H, k, ei, og, basis)
    refP = P.apply_P3(H, k, ei, og, basis)
    refK = np.array([P.HMkH3(H[:, j], k, ei, og, basis) for j in range(m)])
    solver = omx.HIPSolver()
    h = solver.handle(g, ei)
    omx._lib.check(omx._lib.lib().omx_set_option(h.h, b"slab3d", float(slab)), h.h)
    for prec, tol in ((omx.OMX_PREC_FP64, 1e-12), (omx.OMX_PREC_FP32, 3e-6)):
        Mh = omx.HelmholtzMap(k, ei, g, solver, precision=prec)
        assert _relerr(Mh * H, refM) < tol
        assert _relerr(Mh.precondition(H), refP) < tol
        assert np.allclose(Mh.HMkH(H), refK, rtol=10 * tol, atol=0)


def test_period_derivatives_match_oracle():
    og, g = _grids((16, 12, 10))
    ei = _aniso3(og)
    rng = np.random.default_rng(4)
    ks = [0.31, 0.77]
    evs = []
    for _ in ks:
        v = rng.standard_normal(2 * og.N) + 1j * rng.standard_normal(2 * og.N)
        evs.append(v / np.linalg.norm(v))
    dk, dL = omx.period_derivatives(ks, evs, ei, g, omx.HIPSolver())
    for j, k in enumerate(ks):
        assert abs(dk[j] - 2 * P.HMkH3(evs[j], k, ei, og)) < 1e-11 * abs(dk[j])
        ref = P.dw2_dL(evs[j], k, ei, og)
        assert abs(dL[j] - ref) < 1e-11 * abs(ref)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_solve_omega2_3d_matches_dense(precision):
    og, g = _grids((6, 6, 4), D=(2.0, 1.6, 0.9))
    ei = P.bragg_wg_epsi(og, w0=0.5, h0=0.4)
    k = 0.45
    A = P.apply_M3(np.eye(2 * og.N), k, ei, og)
    w = np.linalg.eigvalsh(0.5 * (A + A.conj().T))
    w2, evs = omx.solve_omega2(k, ei, g, omx.HIPSolver(precision=precision), nev=4, tol=1e-10, maxiter=300)
    assert np.allclose(w2, w[:4], rtol=1e-9)
    for j in range(4):
        r = P.apply_M3(evs[j], k, ei, og) - w2[j] * evs[j]
        assert np.linalg.norm(r) < 1e-7


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_solve_k_3d_matches_oracle(precision):
    print(f"3D solve_k {precision}", flush=True)
    # period 0.45 um: the modes at omega = 0.5 lie inside the first Brillouin zone (k < 1/(2 Dz))
    og, g = _grids((16, 16, 8), D=(2.4, 2.0, 0.45))
    ei = P.bragg_wg_epsi(og, eps_core=4.0, eps_bg=2.1, w0=0.5, h0=0.4, mod=0.5)
    omega = 0.5
    kr, evr = O.solve_k(omega, ei, og, nev=2, k_tol=1e-12, eig_tol=1e-10, maxiter=300)
    km, evs = omx.solve_k(omega, ei, g, omx.HIPSolver(precision=precision), nev=2, k_tol=1e-11, eig_tol=1e-9,
                          maxiter=300)
    print(f"  k gpu {km} oracle {kr}", flush=True)
    # |Δω²| from |Δk| through the band slope
    for j in range(2):
        slope = 2 * P.HMkH3(evr[j], kr[j], ei, og)
        assert abs(slope * (km[j] - kr[j])) < 1e-9 * omega ** 2
        assert abs(np.linalg.norm(evs[j]) - 1) < 1e-12
        assert abs(np.vdot(evs[j], evr[j])) > 1 - 1e-7
        # the same canonical phase as the oracle's canonicalize_phase
        assert np.linalg.norm(evs[j] - evr[j]) < 1e-3


def test_solve_k_periodic_pullback_matches_oracle_and_finite_differences():
    og, g = _grids((12, 12, 6), D=(2.0, 1.6, 0.45))
    ei = P.bragg_wg_epsi(og, eps_core=4.0, eps_bg=2.1, w0=0.5, h0=0.4, mod=0.5)
    omega, L = 0.5, 0.45
    solver = omx.HIPSolver(precision="fp64")
    kw = dict(nev=2, k_tol=1e-13, eig_tol=1e-11, maxiter=400)
    (km, evs), pb = omx.rrule_solve_k_periodic(omega, ei, L, g, solver, **kw)
    kbar = [1.0, -0.5]
    wbar, Lbar = pb((kbar, None))
    wref, Lref = P.solve_k_periodic_kbar_pullback(omega, ei, L, og, km, evs, kbar)
    assert abs(wbar - wref) < 1e-9 * abs(wref) and abs(Lbar - Lref) < 1e-9 * abs(Lref)
    h = 1e-4

    def f(w, Lp):
        k, _ = omx.solve_k_periodic(w, ei, Lp, g, solver, fetch_evecs=False, **kw)
        return float(np.dot(kbar, k))

    fd_w = (f(omega + h, L) - f(omega - h, L)) / (2 * h)
    fd_L = (f(omega, L + h) - f(omega, L - h)) / (2 * h)
    print(f"  wbar {wbar} fd {fd_w}; Lbar {Lbar} fd {fd_L}", flush=True)
    assert abs(wbar - fd_w) < 1e-6 * abs(fd_w)
    assert abs(Lbar - fd_L) < 1e-5 * abs(fd_L)


def test_3d_handles_reject_2d_only_paths():
    og, g = _grids((8, 8, 4), D=(2.4, 2.0, 0.45))
    ei = P.bragg_wg_epsi(og, eps_core=4.0, eps_bg=2.1, w0=0.5, h0=0.4, mod=0.5)
    solver = omx.HIPSolver()
    km, evs = omx.solve_k(0.5, ei, g, solver, nev=1, maxiter=300)
    with pytest.raises(omx.OmxError) as e:
        omx.rrule_solve_k(0.5, ei, g, solver, nev=1)[1](([1.0], None))
    assert e.value.status == omx._lib.OMX_ERR_UNSUPPORTED
    with pytest.raises(omx.OmxError) as e:  # Nz = 7: a prime factor the FFT does not take
        omx.solve_k(0.5, np.repeat(ei[..., :1], 7, axis=-1), omx.Grid3(2.4, 2.0, 0.9, 8, 8, 7), solver, nev=1)
    assert e.value.status == omx._lib.OMX_ERR_BAD_SHAPE
