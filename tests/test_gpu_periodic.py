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
    _, wbar, eibar, Lbar, _, _ = pb((kbar, None))
    wref, Lref = P.solve_k_periodic_kbar_pullback(omega, ei, L, og, km, evs, kbar)
    assert abs(wbar - wref) < 1e-9 * abs(wref) and abs(Lbar - Lref) < 1e-9 * abs(Lref)
    _, eref, _ = P.solve_k_periodic_pullback(omega, ei, L, og, km, evs, kbar, None)
    assert eibar.shape == (3, 3) + og.shape
    assert np.max(np.abs(eibar - eref)) < 1e-9 * np.max(np.abs(eref))
    h = 1e-4

    def f(w, Lp):
        k, _ = omx.solve_k_periodic(w, ei, Lp, g, solver, fetch_evecs=False, **kw)
        return float(np.dot(kbar, k))

    fd_w = (f(omega + h, L) - f(omega - h, L)) / (2 * h)
    fd_L = (f(omega, L + h) - f(omega, L - h)) / (2 * h)
    print(f"  wbar {wbar} fd {fd_w}; Lbar {Lbar} fd {fd_L}", flush=True)
    assert abs(wbar - fd_w) < 1e-6 * abs(fd_w)
    assert abs(Lbar - fd_L) < 1e-5 * abs(fd_L)


def _full_pullback_case():
    og, g = _grids((12, 12, 6), D=(2.0, 1.6, 0.45))
    ei = P.bragg_wg_epsi(og, eps_core=4.0, eps_bg=2.1, w0=0.5, h0=0.4, mod=0.5)
    return og, g, ei


def test_solve_k_periodic_full_pullback_matches_oracle():
    # both channels: k̄ on band 1, ēv = 2 c <c, ev> on both bands (the cotangent of |<ev, c>|^2)
    print("3D full period pull-back", flush=True)
    og, g, ei = _full_pullback_case()
    omega, L = 0.5, 0.45
    solver = omx.HIPSolver(precision="fp64")
    kw = dict(nev=2, k_tol=1e-13, eig_tol=1e-11, maxiter=400)
    (km, evs), pb = omx.rrule_solve_k_periodic(omega, ei, L, g, solver, **kw)
    rng = np.random.default_rng(21)
    c = rng.standard_normal(2 * og.N) + 1j * rng.standard_normal(2 * og.N)
    evbar = [2 * c * np.vdot(c, evs[j]) for j in range(2)]
    for kbar, eb in (([0.7, None], evbar), (None, [None, evbar[1]])):
        _, wbar, eibar, Lbar, _, _ = pb((kbar, eb))
        wref, eref, Lref = P.solve_k_periodic_pullback(omega, ei, L, og, km, evs, kbar or [None, None], eb)
        print(f"  wbar {wbar} ref {wref}; Lbar {Lbar} ref {Lref}", flush=True)
        assert abs(wbar - wref) < 1e-7 * abs(wref)
        assert abs(Lbar - Lref) < 1e-7 * abs(Lref)
        assert np.max(np.abs(eibar - eref)) < 1e-7 * np.max(np.abs(eref))
        # the 3D convention: nothing below the diagonal
        for a, b in ((1, 0), (2, 0), (2, 1)):
            assert not np.any(eibar[a, b])


def test_solve_k_periodic_eigenvector_pullback_matches_finite_differences():
    # J = |<ev_2, c>|^2 through GPU re-solves, in ω and in Λ
    og, g, ei = _full_pullback_case()
    omega, L = 0.5, 0.45
    solver = omx.HIPSolver(precision="fp64")
    kw = dict(nev=2, k_tol=1e-13, eig_tol=1e-12, maxiter=500)
    (km, evs), pb = omx.rrule_solve_k_periodic(omega, ei, L, g, solver, **kw)
    c = np.random.default_rng(22).standard_normal(2 * og.N) + 0j
    _, wbar, _, Lbar, _, _ = pb((None, [None, 2 * c * np.vdot(c, evs[1])]))

    def J(w, Lp):
        _, ev = omx.solve_k_periodic(w, ei, Lp, g, solver, **kw)
        return abs(np.vdot(ev[1], c)) ** 2

    h = 1e-4
    fd_w = (J(omega + h, L) - J(omega - h, L)) / (2 * h)
    fd_L = (J(omega, L + h) - J(omega, L - h)) / (2 * h)
    print(f"  wbar {wbar} fd {fd_w}; Lbar {Lbar} fd {fd_L}", flush=True)
    assert abs(wbar - fd_w) < 1e-4 * abs(fd_w)
    assert abs(Lbar - fd_L) < 1e-4 * abs(fd_L)


def test_solve_k_adjoint_on_a_3d_handle_matches_oracle():
    # rrule(solve_k) on a 3D cell: ω̄ and ε̄⁻¹ of the periodic pull-back, no period cotangent
    og, g, ei = _full_pullback_case()
    solver = omx.HIPSolver(precision="fp64")
    (km, evs), pb = omx.rrule_solve_k(0.5, ei, g, solver, nev=1, k_tol=1e-13, eig_tol=1e-11, maxiter=400)
    c = np.random.default_rng(23).standard_normal(2 * og.N) + 0j
    eb = [2 * c * np.vdot(c, evs[0])]
    _, wbar, eibar, _, _ = pb(([1.0], eb))
    wref, eref, _ = P.solve_k_periodic_pullback(0.5, ei, og.Dz, og, km, evs, [1.0], eb)
    assert abs(wbar - wref) < 1e-7 * abs(wref)
    assert np.max(np.abs(eibar - eref)) < 1e-7 * np.max(np.abs(eref))


def test_3d_handles_reject_2d_only_paths():
    og, g = _grids((8, 8, 4), D=(2.4, 2.0, 0.45))
    ei = P.bragg_wg_epsi(og, eps_core=4.0, eps_bg=2.1, w0=0.5, h0=0.4, mod=0.5)
    solver = omx.HIPSolver()
    km, evs = omx.solve_k(0.5, ei, g, solver, nev=1, maxiter=300)
    d = np.zeros_like(ei)
    with pytest.raises(omx.OmxError) as e:
        omx.ng_gvd(0.5, km[0], evs[0], ei, d, d, g, solver)
    assert e.value.status == omx._lib.OMX_ERR_UNSUPPORTED
    # the period cotangent needs a 3D handle
    og2 = O.Grid(2.4, 2.0, 8, 8)
    g2 = omx.Grid(2.4, 2.0, 8, 8)
    ei2 = O.gaussian_wg_epsi([4.0, 2.1, 0.5, 0.4], og2)
    k2, ev2 = omx.solve_k(0.5, ei2, g2, solver, nev=1, maxiter=300)
    h2 = solver.handle(g2, ei2)
    with pytest.raises(omx.OmxError) as e:
        omx._adjoint(h2, g2, 0.5, 1, k2, ev2, ([1.0], None), period=True)
    assert e.value.status == omx._lib.OMX_ERR_UNSUPPORTED
    with pytest.raises(omx.OmxError) as e:  # Nz = 7: a prime factor the FFT does not take
        omx.solve_k(0.5, np.repeat(ei[..., :1], 7, axis=-1), omx.Grid3(2.4, 2.0, 0.9, 8, 8, 7), solver, nev=1)
    assert e.value.status == omx._lib.OMX_ERR_BAD_SHAPE
