"""Pins the FP64 oracle against closed forms, an independent dense assembly, brute force
and finite differences (the reference ships no golden values for this path; its own
tests are invariants: lib/MaxwellEigenmodes/test/runtests.jl:50-285)."""
import numpy as np
import pytest

import oracle as O

GRID = O.Grid(6.0, 4.0, 24, 16)          # runtests.jl:43 (Nx != Ny catches x/y mix-ups)
P_WG = [4.2, 2.1, 1.0, 0.6]              # runtests.jl:44
W0 = 1 / 1.55                            # runtests.jl:46


def _rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _aniso_epsinv(grid, seed=3):
    """Symmetric, positive-definite, spatially asymmetric ε⁻¹ with xy/xz/yz couplings, so
    a swapped FFT direction or transposed index would change the operator."""
    rng = np.random.default_rng(seed)
    xs, ys = grid.x(), grid.y()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    bump = np.exp(-((X - 0.7) ** 2 / 1.1 + (Y + 0.4) ** 2 / 0.5))
    eps = np.zeros((3, 3, grid.Nx, grid.Ny))
    for a, base in enumerate((2.0, 2.3, 2.1)):
        eps[a, a] = base + 1.8 * bump * (1 + 0.1 * a)
    eps[0, 1] = eps[1, 0] = 0.3 * bump * np.sign(X + 1e-3)
    eps[0, 2] = eps[2, 0] = 0.1 * bump
    eps[1, 2] = eps[2, 1] = -0.05 * bump * (Y > 0)
    eps += 0.0 * rng.standard_normal(eps.shape)
    return O.sliceinv_3x3(eps)


def _dense(op, n2):
    return np.stack([op.M(np.eye(n2)[:, j]) for j in range(n2)], 1)


def _independent_dense(k, epsinv, grid):
    """Plane-wave matrix assembled without the FFT pipeline: M = (1/N) U^T Ê U with
    Ê(G',G) = sum_r ε⁻¹(r)^T exp(+2πi r.(G'-G)) (explicit DFT sums) and
    u_{G,b} = (G - k ẑ) x ê_b, ê_1 = m̂, ê_2 = n̂ from np.cross.  Recipe of
    legacy/explicit.jl:263-275."""
    Nx, Ny, N = grid.Nx, grid.Ny, grid.N
    gx, gy = O.gvec(grid)
    jx, jy = np.arange(Nx), np.arange(Ny)
    # G list in flat order ix + Nx*iy
    GX = np.tile(gx, Ny)
    GY = np.repeat(gy, Nx)
    JX = np.tile(jx, Ny)
    JY = np.repeat(jy, Nx)
    U = np.zeros((N, 2, 3))
    for p in range(N):
        w = np.array([-GX[p], -GY[p], k])
        zxw = np.cross([0.0, 0.0, 1.0], w)
        n = zxw / np.linalg.norm(zxw) if (w[0] ** 2 + w[1] ** 2) > 0 else np.array([0.0, 1.0, 0.0])
        m = np.cross(n, w)
        m /= np.linalg.norm(m)
        U[p, 0] = np.cross(-w, m)
        U[p, 1] = np.cross(-w, n)
    # Ê(G', G) = sum_r ε⁻¹(r)^T exp(2πi (rx (j'x - jx)/Nx + ry (j'y - jy)/Ny))
    rx, ry = np.arange(Nx), np.arange(Ny)
    ex = np.exp(2j * np.pi * np.outer(rx, np.arange(Nx)) / Nx)   # (rx, j)
    ey = np.exp(2j * np.pi * np.outer(ry, np.arange(Ny)) / Ny)
    EiT = np.transpose(epsinv, (1, 0, 2, 3))                        # [a,b] <- ε⁻¹[b,a]
    M = np.zeros((2 * N, 2 * N), dtype=complex)
    for p2 in range(N):
        for p in range(N):
            phx = ex[:, JX[p2]] * np.conj(ex[:, JX[p]])
            phy = ey[:, JY[p2]] * np.conj(ey[:, JY[p]])
            Ehat = np.einsum("abxy,x,y->ab", EiT, phx, phy)
            blk = U[p2] @ Ehat @ U[p].T
            M[2 * p2:2 * p2 + 2, 2 * p:2 * p + 2] = blk / N
    return M


def test_fftfreq_and_gvec():
    assert np.allclose(O.my_fftfreq(4, 4.0), [0, 1, -2, -1])
    assert np.allclose(O.my_fftfreq(5, 5.0), [0, 1, 2, -2, -1])
    gx, gy = O.gvec(GRID)
    assert gx[1] == pytest.approx(1 / 6.0) and gy[1] == pytest.approx(1 / 4.0)
    assert gx[12] == pytest.approx(-12 / 6.0)


def test_basis_orthonormal_and_gimbal():
    k = 1.3
    mag, m, n = O.mag_mn(k, GRID)
    gx, gy = O.gvec(GRID)
    w = np.stack([-gx[:, None] + 0 * gy[None], 0 * gx[:, None] - gy[None], k + 0 * gx[:, None] * gy[None]])
    assert np.allclose((m * n).sum(0), 0, atol=1e-14)
    assert np.allclose((m * w).sum(0), 0, atol=1e-13) and np.allclose((n * w).sum(0), 0, atol=1e-13)
    assert np.allclose((m * m).sum(0), 1) and np.allclose((n * n).sum(0), 1)
    assert np.allclose(mag, np.sqrt((w * w).sum(0)))
    # gimbal point: w = (0,0,k) -> n = y, m = n x w / |.| = x
    assert np.allclose(n[:, 0, 0], [0, 1, 0]) and np.allclose(m[:, 0, 0], [1, 0, 0])


def test_operator_matches_independent_dense_assembly():
    g = O.Grid(3.0, 2.0, 6, 4)
    ei = _aniso_epsinv(g)
    k = 1.1
    A = _dense(O.HelmholtzOracle(k, ei, g), 2 * g.N)
    B = _independent_dense(k, ei, g)
    assert np.abs(A - B).max() < 1e-12 * np.abs(B).max()
    # a mirrored structure (what a swapped FFT direction would give) is a different operator
    C = _independent_dense(k, ei[:, :, ::-1, ::-1], g)
    assert np.abs(A - C).max() > 1e-3 * np.abs(B).max()


def test_operator_hermitian_psd():
    ei = O.gaussian_wg_epsi(P_WG, GRID)
    op = O.HelmholtzOracle(O.k_guess(W0, ei), ei, GRID)
    rng = np.random.default_rng(0)
    v, w = _rand_c(rng, 2 * GRID.N), _rand_c(rng, 2 * GRID.N)
    assert np.vdot(w, op.M(v)) == pytest.approx(np.conj(np.vdot(v, op.M(w))), rel=1e-9)   # runtests.jl:58
    assert np.real(np.vdot(v, op.M(v))) > 0
    assert O.HMH(v, op.k, ei, GRID) == pytest.approx(np.real(np.vdot(v, op.M(v))), rel=1e-6)   # runtests.jl:61


def test_homogeneous_medium_plane_waves():
    """ε⁻¹ = I/ε: every plane wave is an eigenvector with eigenvalue |k - G|²/ε (twice),
    and P̂ = M̂⁻¹ exactly."""
    eps = 2.25
    ei = np.zeros((3, 3) + GRID.shape)
    for a in range(3):
        ei[a, a] = 1 / eps
    k = 0.9
    op = O.HelmholtzOracle(k, ei, GRID)
    gx, gy = O.gvec(GRID)
    lam = ((gx[:, None] ** 2 + gy[None] ** 2 + k ** 2) / eps)
    rng = np.random.default_rng(1)
    for (ix, iy) in [(0, 0), (3, 5), (23, 15), (12, 8)]:
        for c in range(2):
            h = np.zeros((2,) + GRID.shape, complex)
            h[c, ix, iy] = 1.0
            v = O.to_flat(h)
            assert np.allclose(op.M(v), lam[ix, iy] * v, atol=1e-12)
    v = _rand_c(rng, 2 * GRID.N)
    assert np.allclose(op.P(op.M(v)), v, atol=1e-11)
    # full spectrum, twofold degenerate
    A = _dense(op, 2 * GRID.N)
    assert np.allclose(np.sort(np.linalg.eigvalsh(A)), np.sort(np.repeat(lam.ravel(), 2)), atol=1e-11)


def test_lobpcg_matches_dense_eigh():
    ei = O.gaussian_wg_epsi(P_WG, GRID)
    op = O.HelmholtzOracle(1.1, ei, GRID)
    A = _dense(op, 2 * GRID.N)
    ref = np.linalg.eigvalsh((A + A.conj().T) / 2)[:4]
    rng = np.random.default_rng(2)
    res = O.lobpcg(op.M, _rand_c(rng, 2 * GRID.N, 4), precon=op.P, tol=1e-10, maxiter=300)
    assert np.allclose(res["lam"], ref, rtol=1e-12)
    assert np.all(res["residual_norms"] < 1e-9)
    X = res["X"]
    assert np.allclose(X.conj().T @ X, np.eye(4), atol=1e-12)


def test_ortho_shifted_cholesky_rank_deficient():
    """Rank-deficient start block: Cholesky fails and the shifted retry (dftk.jl:168-192)
    must still produce an orthonormal block."""
    rng = np.random.default_rng(4)
    X = _rand_c(rng, 200, 3)
    X = np.hstack([X, X[:, :1] * (1 + 1e-17)])
    Y = X.copy()
    O.ortho(Y)
    assert np.allclose(Y.conj().T @ Y, np.eye(4), atol=1e-8)


def test_hmkh_is_half_derivative_of_omega2():
    """Hellmann-Feynman: dω²/dk = 2 HMkH (solve.jl:82), checked by central FD of dense
    eigenvalues."""
    ei = _aniso_epsinv(O.Grid(4.0, 3.0, 12, 10))
    g = O.Grid(4.0, 3.0, 12, 10)
    k, h = 1.2, 1e-5
    def lowest(kk):
        A = _dense(O.HelmholtzOracle(kk, ei, g), 2 * g.N)
        w, V = np.linalg.eigh((A + A.conj().T) / 2)
        return w[0], V[:, 0]
    w0, v0 = lowest(k)
    fd = (lowest(k + h)[0] - lowest(k - h)[0]) / (2 * h)
    assert 2 * O.HMkH(v0, k, ei, g) == pytest.approx(fd, rel=1e-7)


def test_solve_k_round_trip_and_bounds():
    ei = O.gaussian_wg_epsi(P_WG, GRID)
    km, evs = O.solve_k(W0, ei, GRID, nev=2, k_tol=1e-12, eig_tol=1e-11)
    neff = km / W0
    assert np.all((np.sqrt(P_WG[1]) < neff) & (neff < np.sqrt(P_WG[0])))       # runtests.jl:78
    for k, ev in zip(km, evs):
        op = O.HelmholtzOracle(k, ei, GRID)
        assert np.linalg.norm(op.M(ev) - W0 ** 2 * ev) < 1e-9                 # runtests.jl:73
        A = _dense(op, 2 * GRID.N)
        lam = np.linalg.eigvalsh((A + A.conj().T) / 2)
        assert np.min(np.abs(lam - W0 ** 2)) / W0 ** 2 < 1e-11                    # runtests.jl:82
    # band ordering: band i is the i-th eigenvalue at its own k
    A = _dense(O.HelmholtzOracle(km[1], ei, GRID), 2 * GRID.N)
    lam = np.linalg.eigvalsh((A + A.conj().T) / 2)
    assert abs(lam[1] - W0 ** 2) < 1e-10


def test_canonicalize_phase_makes_max_E_real_positive():
    ei = O.gaussian_wg_epsi(P_WG, GRID)
    rng = np.random.default_rng(5)
    v = _rand_c(rng, 2 * GRID.N)
    k = 1.2
    c = O.canonicalize_phase(v, k, ei, GRID)
    E = O.E_field(k, c, ei, GRID).reshape(-1, order="F")
    i = np.argmax(np.abs(E) ** 2)
    assert abs(np.angle(E[i])) < 1e-12 and E[i].real > 0
    assert np.allclose(np.abs(c), np.abs(v))


def test_eig_adjt_dense_fd():
    """runtests.jl:267-285: dense 40x40 Hermitian eigenvector adjoint against FD."""
    rng = np.random.default_rng(6)
    n = 40
    A0 = rng.standard_normal((n, n)); A = (A0 + A0.T) / 2
    w, V = np.linalg.eigh(A)
    a, v = w[0], V[:, 0]
    B0 = rng.standard_normal((n, n)); B = (B0 + B0.T) / 2
    xbar = 2 * B @ v
    lam = O.eig_adjt(A, a, v.astype(complex), 0.0, xbar.astype(complex))
    dA0 = rng.standard_normal((n, n)); dA = (dA0 + dA0.T) / 2 * 1e-2
    def f(t):
        wt, Vt = np.linalg.eigh(A + t * dA)
        vt = Vt[:, 0] * np.sign(np.dot(Vt[:, 0], v))
        return vt @ B @ vt
    h = 1e-4
    fd = (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)
    adj = -np.real(np.vdot(lam, dA @ v))
    assert adj == pytest.approx(fd, rel=1e-6)


def _kfun(omega, ei, g, nev=1):
    km, _ = O.solve_k(omega, ei, g, nev=nev, k_tol=1e-13, eig_tol=1e-12, seed=11)
    return km


def test_adjoint_dk_domega_fd():
    g = O.Grid(6.0, 4.0, 16, 12)
    ei = O.gaussian_wg_epsi(P_WG, g)
    km, evs = O.solve_k(W0, ei, g, nev=1, k_tol=1e-13, eig_tol=1e-12)
    wbar, eibar = O.solve_k_pullback(W0, ei, g, km, evs, [1.0], None)
    h = 1e-4
    fd = (_kfun(W0 + h, ei, g)[0] - _kfun(W0 - h, ei, g)[0]) / (2 * h)
    assert wbar == pytest.approx(fd, rel=1e-6)                                 # runtests.jl:91-94
    # group index from HMH/HMkH equals dk/dω for non-dispersive ε
    ng = O.group_index(km[0], evs[0], W0, ei, np.zeros_like(ei), g)
    assert ng == pytest.approx(fd, rel=1e-6)


def test_adjoint_epsinv_directional_fd():
    """runtests.jl:98-110: ε⁻¹ gradient of k along a diagonal-only direction, through the
    eigenvalue channel (k̄) and, with an eigenvector cotangent, through the adjoint solve."""
    g = O.Grid(6.0, 4.0, 16, 12)
    ei = O.gaussian_wg_epsi(P_WG, g)
    rng = np.random.default_rng(7)
    d = np.zeros_like(ei)
    for a in range(3):
        d[a, a] = rng.standard_normal(g.shape) * 1e-3
    km, evs = O.solve_k(W0, ei, g, nev=1, k_tol=1e-13, eig_tol=1e-12)
    _, eibar = O.solve_k_pullback(W0, ei, g, km, evs, [1.0], None)
    h = 1e-3
    fd = (_kfun(W0, ei + h * d, g)[0] - _kfun(W0, ei - h * d, g)[0]) / (2 * h)
    assert np.sum(eibar * d) == pytest.approx(fd, rel=1e-5)

    # eigenvector channel: objective J = |<ev, c>|^2 (phase invariant) for a fixed c
    c = _rand_c(rng, 2 * g.N)
    def J(e):
        _, ev = O.solve_k(W0, e, g, nev=1, k_tol=1e-13, eig_tol=1e-12, seed=11)
        return abs(np.vdot(ev[0], c)) ** 2
    evbar = 2 * c * np.vdot(c, evs[0])       # ∂J/∂conj(ev) convention used by the rrule
    _, eibar2 = O.solve_k_pullback(W0, ei, g, km, evs, [None], [evbar])
    fd2 = (J(ei + h * d) - J(ei - h * d)) / (2 * h)
    assert np.sum(eibar2 * d) == pytest.approx(fd2, rel=1e-4)


@pytest.mark.parametrize("pair", [(0, 1), (0, 2), (1, 2), (0, 0), (2, 2)])
def test_epsinv_bar_full_tensor_directional_fd(pair):
    """The 2D ε̄⁻¹ convention (grads/solve.jl:498-511) pinned by finite differences on a
    full-tensor ε⁻¹: off-diagonal (i, j) and (j, i) both hold the derivative with respect to
    the symmetric pair.  So a perturbation δ·(e_ij + e_ji) per pixel changes k (k̄ channel)
    and |<ev, c>|² (ēv channel) by Σ_pixels ε̄⁻¹[i, j]·δ = ½ <ε̄⁻¹, D> -- half the full
    contraction.  A diagonal direction e_ii changes them by <ε̄⁻¹, D>.  A dropped mirror term,
    a factor of two or a transposed index fails one of the two channels."""
    i, j = pair
    g = O.Grid(4.0, 3.0, 16, 12)
    ei = _aniso_epsinv(g)
    rng = np.random.default_rng(20 + 3 * i + j)
    r = rng.standard_normal(g.shape) * 0.02
    D = np.zeros_like(ei)
    D[i, j] = r
    D[j, i] = r
    km, evs = O.solve_k(W0, ei, g, nev=1, k_tol=1e-13, eig_tol=1e-12, seed=11)
    c = _rand_c(rng, 2 * g.N)
    evbar = 2 * c * np.vdot(c, evs[0])
    _, eb_k = O.solve_k_pullback(W0, ei, g, km, evs, [1.0], None)
    _, eb_v = O.solve_k_pullback(W0, ei, g, km, evs, [None], [evbar])
    assert np.allclose(eb_k, np.transpose(eb_k, (1, 0, 2, 3)))
    pred_k = np.sum(eb_k[i, j] * r)
    pred_v = np.sum(eb_v[i, j] * r)
    if i != j:
        assert pred_k == pytest.approx(0.5 * np.sum(eb_k * D), rel=1e-12)

    def solve(e):
        k, ev = O.solve_k(W0, e, g, nev=1, k_tol=1e-13, eig_tol=1e-12, seed=11)
        return k[0], abs(np.vdot(ev[0], c)) ** 2
    h = 1e-3
    (kp, jp), (kn, jn) = solve(ei + h * D), solve(ei - h * D)
    assert (kp - kn) / (2 * h) == pytest.approx(pred_k, rel=1e-6)
    assert (jp - jn) / (2 * h) == pytest.approx(pred_v, rel=1e-4)


def test_sliceinv_pullback_fd():
    g = O.Grid(1.0, 1.0, 3, 2)
    A = np.zeros((3, 3) + g.shape)
    for i in range(3):
        for j in range(2):
            A[:, :, i, j] = [[2.0 + 0.1 * i, 0.05, 0.0], [0.05, 2.3, 0.0], [0.0, 0.0, 2.1 + 0.2 * j]]
    rng = np.random.default_rng(9)
    Ybar = rng.standard_normal(A.shape)
    dirA = rng.standard_normal(A.shape)
    f = lambda t: np.sum(Ybar * O.sliceinv_3x3(A + t * dirA))  # noqa: E731
    h = 1e-5
    fd = (f(h) - f(-h)) / (2 * h)
    Abar = O.sliceinv_3x3_pullback(O.sliceinv_3x3(A), Ybar)
    assert np.sum(Abar * dirA) == pytest.approx(fd, rel=1e-7)
