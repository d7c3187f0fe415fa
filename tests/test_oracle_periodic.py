"""The 3D (z-periodic) oracle pinned by closed forms, by the already-pinned 2D oracle, by
dense eigendecompositions and by finite differences in k, in the period and in ω."""
import numpy as np
import pytest

import oracle as O
from oracle import periodic as P

RNG = np.random.default_rng(11)


def _rand(n):
    return RNG.standard_normal(n) + 1j * RNG.standard_normal(n)


def _homog(grid, eps):
    ei = np.zeros((3, 3) + grid.shape)
    for a in range(3):
        ei[a, a] = 1.0 / eps
    return ei


def test_homogeneous_closed_form():
    # every plane wave is an eigenvector: M = |k - G|^2 / eps, P = eps / |k - G|^2
    g = P.Grid3(2.0, 1.5, 0.8, 6, 5, 4)
    eps, k = 2.3, 0.7
    ei = _homog(g, eps)
    H = _rand(2 * g.N)
    mag, _, _ = P.mag_mn3(k, g)
    Hg = P.to_grid3(H, g)
    want_m = P.to_flat3(Hg * (mag ** 2 / eps)[None])
    want_p = P.to_flat3(Hg * (eps / mag ** 2)[None])
    assert np.abs(P.apply_M3(H, k, ei, g) - want_m).max() < 1e-13 * np.abs(want_m).max()
    assert np.abs(P.apply_P3(H, k, ei, g) - want_p).max() < 1e-13 * np.abs(want_p).max()


def test_basis_is_orthonormal_and_transverse():
    g = P.Grid3(2.0, 1.5, 0.8, 6, 5, 4)
    k = 0.3
    mag, m, n = P.mag_mn3(k, g)
    gx, gy, gz = P.gvec3(g)
    w = np.stack(np.broadcast_arrays(-gx[:, None, None], -gy[None, :, None], k - gz[None, None, :]))
    assert np.allclose((m * m).sum(0), 1) and np.allclose((n * n).sum(0), 1)
    assert np.allclose((m * n).sum(0), 0, atol=1e-15)
    assert np.allclose((m * w).sum(0), 0, atol=1e-14) and np.allclose((n * w).sum(0), 0, atol=1e-14)
    assert np.allclose(mag, np.sqrt((w * w).sum(0)))


@pytest.mark.parametrize("nz", [1, 4])
def test_z_invariant_cell_equals_2d_planes(nz):
    # with eps^-1 constant along z, plane iz (g_z = gz[iz]) of the 3D operator is the 2D
    # operator at k - gz[iz]; Nz = 1 is the 2D operator itself
    g2 = O.Grid(2.0, 1.5, 6, 5)
    g = P.Grid3(2.0, 1.5, 0.8, 6, 5, nz)
    ei2 = O.gaussian_wg_epsi([4.0, 2.1, 0.5, 0.4], g2)
    ei3 = np.repeat(ei2[..., None], nz, axis=-1)
    k = 0.7
    H = _rand(2 * g.N)
    Hg = P.to_grid3(H, g)
    gz = P.gvec3(g)[2]
    for fn3, fn2 in ((P.apply_M3, O.apply_M), (P.apply_P3, O.apply_P)):
        Y3 = P.to_grid3(fn3(H, k, ei3, g), g)
        for iz in range(nz):
            Y2 = O.to_grid(fn2(O.to_flat(Hg[..., iz]), k - gz[iz], ei2, g2), g2)
            assert np.abs(Y2 - Y3[..., iz]).max() < 1e-13 * np.abs(Y2).max()
    if nz == 1:
        assert abs(P.HMkH3(H, k, ei3, g) - O.HMkH(H, k, ei2, g2)) < 1e-12 * abs(O.HMkH(H, k, ei2, g2))


def _dense(k, eps_inv, grid):
    A = P.apply_M3(np.eye(2 * grid.N), k, eps_inv, grid)
    return 0.5 * (A + A.conj().T)


def test_derivatives_in_k_and_period_match_finite_differences():
    g = P.Grid3(2.0, 1.6, 0.9, 6, 6, 4)
    ei = P.bragg_wg_epsi(g, w0=0.5, h0=0.4)
    k, h = 0.7, 1e-5
    A = P.apply_M3(np.eye(2 * g.N), k, ei, g)
    assert np.abs(A - A.conj().T).max() < 1e-13 * np.abs(A).max()  # Hermitian
    w, V = np.linalg.eigh(_dense(k, ei, g))
    assert w[0] > 0
    for j in (0, 2):
        ev = V[:, j]
        fd_k = (np.linalg.eigvalsh(_dense(k + h, ei, g))[j] - np.linalg.eigvalsh(_dense(k - h, ei, g))[j]) / (2 * h)
        assert abs(fd_k - 2 * P.HMkH3(ev, k, ei, g)) < 1e-8 * abs(fd_k)
        gp, gm = P.grid_with_period(g, g.Dz + h), P.grid_with_period(g, g.Dz - h)
        fd_L = (np.linalg.eigvalsh(_dense(k, ei, gp))[j] - np.linalg.eigvalsh(_dense(k, ei, gm))[j]) / (2 * h)
        assert abs(fd_L - P.dw2_dL(ev, k, ei, g)) < 1e-8 * abs(fd_L)
    # the weighted form with weight 1 is HMkH; with weight 0 it vanishes
    ev = V[:, 0]
    assert abs(P.HMkH_weighted(ev, k, ei, g, np.zeros(g.shape))) == 0.0


def test_lobpcg_and_newton_on_a_3d_cell():
    g = P.Grid3(2.0, 1.6, 0.9, 6, 6, 4)
    ei = P.bragg_wg_epsi(g, w0=0.5, h0=0.4)
    k = 0.45
    w = np.linalg.eigvalsh(_dense(k, ei, g))
    H0 = RNG.standard_normal((2 * g.N, 3)) + 0j
    lam, X, _ = O.solve_omega2(k, ei, g, H0, tol=1e-10, maxiter=300)
    assert np.allclose(lam, w[:3], rtol=1e-10)
    # Newton: omega^2(k_j) = omega^2 for each band, checked on the dense operator
    omega = float(np.sqrt(w[1]))
    km, evs = O.solve_k(omega, ei, g, nev=2, k_tol=1e-12, eig_tol=1e-10, maxiter=300)
    for j, kj in enumerate(km):
        assert abs(np.linalg.eigvalsh(_dense(kj, ei, g))[j] - omega ** 2) < 1e-9
        assert abs(np.linalg.norm(evs[j]) - 1) < 1e-12


def test_period_pullback_matches_finite_differences_of_resolves():
    g = P.Grid3(2.0, 1.6, 0.9, 6, 6, 4)
    ei = P.bragg_wg_epsi(g, w0=0.5, h0=0.4)
    omega, L = 0.42, g.Dz
    kw = dict(nev=1, k_tol=1e-13, eig_tol=1e-11, maxiter=400)
    km, evs = P.solve_k_periodic(omega, ei, L, g, **kw)
    wbar, Lbar = P.solve_k_periodic_kbar_pullback(omega, ei, L, g, km, evs, [1.0])
    h = 1e-5
    fd_w = (P.solve_k_periodic(omega + h, ei, L, g, **kw)[0][0] - P.solve_k_periodic(omega - h, ei, L, g, **kw)[0][0]) / (2 * h)
    fd_L = (P.solve_k_periodic(omega, ei, L + h, g, **kw)[0][0] - P.solve_k_periodic(omega, ei, L - h, g, **kw)[0][0]) / (2 * h)
    assert abs(wbar - fd_w) < 1e-6 * abs(fd_w)
    assert abs(Lbar - fd_L) < 1e-6 * abs(fd_L)


def test_epsinv_bar3_convention():
    d = RNG.standard_normal((3, 2, 2, 2)) + 1j * RNG.standard_normal((3, 2, 2, 2))
    ld = RNG.standard_normal((3, 2, 2, 2)) + 1j * RNG.standard_normal((3, 2, 2, 2))
    e = P.epsinv_bar3(d, ld)
    e2 = O.epsinv_bar(d.reshape(3, 2, 4), ld.reshape(3, 2, 4)).reshape(3, 3, 2, 2, 2)
    # grads/solve.jl:464-487: the 2D values on and above the diagonal, zeros below
    for a in range(3):
        for b in range(3):
            assert np.allclose(e[a, b], e2[a, b] if a <= b else 0.0)


def test_wbar_z_sums_to_the_k_gradient_in_2d():
    # Σ w̄_z over the plane waves is ∇ₖmag_m_n with dk = z (grads/period.jl:97-119), checked
    # against the pinned 2D restatement on a single plane
    g2 = O.Grid(2.0, 1.5, 6, 5)
    mag, m, n = O.mag_mn(0.7, g2)
    magbar = RNG.standard_normal(mag.shape)
    mbar, nbar = RNG.standard_normal(m.shape), RNG.standard_normal(n.shape)
    ref = O.grad_k_mag_m_n(magbar, mbar, nbar, mag, m, n)
    assert abs(np.sum(P.wbar_z(magbar, mbar, nbar, mag, m, n)) - ref) < 1e-12 * max(1.0, abs(ref))


def _bragg_small():
    g = P.Grid3(2.4, 2.0, 0.45, 8, 8, 4)
    return g, P.bragg_wg_epsi(g, eps_core=4.0, eps_bg=2.1, w0=0.5, h0=0.4, mod=0.5)


def _k1(omega, ei, L, g):
    km, ev = P.solve_k_periodic(omega, ei, L, g, nev=1, k_tol=1e-14, eig_tol=1e-12, maxiter=500)
    return km[0], ev[0]


def test_period_pullback_full_matches_finite_differences():
    g, ei = _bragg_small()
    omega, L = 0.5, g.Dz
    k0, ev0 = _k1(omega, ei, L, g)
    # eigenvalue channel alone: dk/dω, dk/dΛ and a symmetric eps^-1 direction
    wbar, eibar, Lbar = P.solve_k_periodic_pullback(omega, ei, L, g, [k0], [ev0], [1.0], None)
    h = 1e-4
    fd_w = (_k1(omega + h, ei, L, g)[0] - _k1(omega - h, ei, L, g)[0]) / (2 * h)
    fd_L = (_k1(omega, ei, L + h, g)[0] - _k1(omega, ei, L - h, g)[0]) / (2 * h)
    assert abs(wbar - fd_w) < 1e-6 * abs(fd_w) and abs(Lbar - fd_L) < 1e-5 * abs(fd_L)
    rng = np.random.default_rng(9)
    d = np.zeros_like(ei)
    for a in range(3):
        d[a, a] = rng.standard_normal(g.shape) * 1e-2
    s = rng.standard_normal(g.shape) * 1e-2
    d[0, 1] = d[1, 0] = s
    he = 1e-4  # the off-diagonal direction has a large curvature: keep the step small
    fd_e = (_k1(omega, ei + he * d, L, g)[0] - _k1(omega, ei - he * d, L, g)[0]) / (2 * he)
    assert abs(np.sum(eibar * d) - fd_e) < 1e-5 * abs(fd_e)
    # eigenvector channel: J = |<ev, c>|^2 (phase invariant)
    c = rng.standard_normal(2 * g.N) + 1j * rng.standard_normal(2 * g.N)

    def J(w, e, Lp):
        return abs(np.vdot(_k1(w, e, Lp, g)[1], c)) ** 2

    evbar = 2 * c * np.vdot(c, ev0)
    wbar2, eibar2, Lbar2 = P.solve_k_periodic_pullback(omega, ei, L, g, [k0], [ev0], [None], [evbar])
    fdj_w = (J(omega + h, ei, L) - J(omega - h, ei, L)) / (2 * h)
    fdj_L = (J(omega, ei, L + h) - J(omega, ei, L - h)) / (2 * h)
    fdj_e = (J(omega, ei + he * d, L) - J(omega, ei - he * d, L)) / (2 * he)
    assert abs(wbar2 - fdj_w) < 1e-4 * abs(fdj_w)
    assert abs(Lbar2 - fdj_L) < 1e-4 * abs(fdj_L)
    assert abs(np.sum(eibar2 * d) - fdj_e) < 1e-4 * abs(fdj_e)
