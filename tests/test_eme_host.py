"""Host S-matrix algebra of omx.eme (mirroring lib/EigenmodeExpansion/src/smatrix.jl):
the reference's interface/propagation checks (lib/EigenmodeExpansion/test/runtests.jl:
84-103) and the algebraic identities of the Redheffer star product."""
import numpy as np
import pytest

from omx import eme as X


def _rand_passive(rng, nl, nr):
    A = rng.standard_normal((nl + nr, nl + nr)) + 1j * rng.standard_normal((nl + nr, nl + nr))
    U, s, Vh = np.linalg.svd(A)
    return X.SMat((U * (0.9 * s / s.max())[None, :]) @ Vh, nl, nr)


def _through(n):
    Z, I = np.zeros((n, n)), np.eye(n)
    return X.SMat(np.block([[Z, I], [I, Z]]), n, n)


def test_star_identity_and_associativity():
    rng = np.random.default_rng(0)
    A, B, C = _rand_passive(rng, 3, 3), _rand_passive(rng, 3, 3), _rand_passive(rng, 3, 3)
    assert np.allclose(X.star(A, _through(3)).S, A.S, atol=1e-13)
    assert np.allclose(X.star(_through(3), A).S, A.S, atol=1e-13)
    assert np.allclose(X.star(X.star(A, B), C).S, X.star(A, X.star(B, C)).S, atol=1e-12)
    assert np.allclose(X.cascade([A, B, C]).S, X.star(X.star(A, B), C).S)
    with pytest.raises(ValueError):
        X.star(A, _rand_passive(rng, 2, 2))


def test_propagation_matrices_compose_and_are_lossless():
    k = np.array([2.1, 1.9, 1.4])
    P = X.propagation_smatrix(k, 0.7)
    assert np.allclose(np.abs(np.diag(P.transmission())), 1.0)
    assert np.allclose(P.reflection(), 0.0)
    assert np.allclose(X.star(P, X.propagation_smatrix(k, 0.5)).S, X.propagation_smatrix(k, 1.2).S, atol=1e-13)


def test_interface_of_identical_bases_is_transparent():
    n = 3
    S = X.interface_smatrix(np.eye(n, dtype=complex), np.eye(n, dtype=complex))
    assert np.allclose(S.transmission(), np.eye(n), atol=1e-8)
    assert np.abs(S.reflection()).max() < 1e-8


def test_interface_passivity_and_reciprocity():
    rng = np.random.default_rng(1)
    O_LR = np.eye(3) + 0.3 * (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))
    O_RL = np.eye(3) + 0.3 * (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))
    for method in ("invert", "clip", "subtract"):
        S = X.interface_smatrix(O_LR, O_RL, passivity=method, reciprocity=False)
        assert np.linalg.svd(S.S, compute_uv=False).max() <= 1 + 1e-12
    S = X.interface_smatrix(O_LR, O_RL)
    assert np.allclose(S.S, S.S.T)
    assert np.allclose(X.enforce_passivity([0.5, 2.0], "invert"), [0.5, 0.5])
    assert np.allclose(X.enforce_passivity([0.5, 2.0], "clip"), [0.5, 1.0])
    assert np.allclose(X.enforce_passivity([0.5, 1.5, 2.5], "subtract"), [0.5, 0.5, 0.0])
    with pytest.raises(ValueError):
        X.enforce_passivity([1.0], "bogus")


def test_reg_solve_is_a_pseudo_inverse():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    B = rng.standard_normal((4, 2))
    assert np.allclose(X.reg_solve(A, B, reg=0.0), np.linalg.solve(A, B), atol=1e-10)
    # rank-deficient: the regularised solve stays finite and close to the pseudo-inverse
    A[:, 3] = A[:, 2]
    assert np.allclose(X.reg_solve(A, B, reg=1e-12), np.linalg.pinv(A) @ B, atol=1e-4)


def test_power_coupling_reads_the_transmission_block():
    S = X.SMat(np.arange(16).reshape(4, 4) * (1 + 0j), 2, 2)
    assert X.power_coupling(S, in_mode=1, out_mode=2) == abs(S.S[3, 0]) ** 2


def test_assemble_pullback_matches_finite_differences():
    """The reverse pass of the host S-matrix algebra (interfaces with the regularised solve,
    propagation phases, Redheffer stars) against central differences of the same algebra,
    for complex directions of every overlap matrix and real directions of every k."""
    rng = np.random.default_rng(0)
    n = 3
    ks = [rng.uniform(1, 2, n) for _ in range(3)]
    ov = []
    for _ in range(2):
        A = np.eye(n) + 0.2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        ov.append((A, A.T + 0.1 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))))
    L = [0.7, 1.3, 0.4]

    def fwd(ov, ks):
        acc = X.propagation_smatrix(ks[0], L[0])
        for i in range(2):
            acc = X.star(X.star(acc, X.interface_smatrix(ov[i][0], ov[i][1], passivity="none")),
                         X.propagation_smatrix(ks[i + 1], L[i + 1]))
        return acc.S

    S0 = fwd(ov, ks)
    W = rng.standard_normal(S0.shape) + 1j * rng.standard_normal(S0.shape)

    def f(S):
        return np.real(np.sum(np.conj(W) * S)) + abs(S[n, 0]) ** 2
    Sbar = W.copy()
    Sbar[n, 0] += 2 * S0[n, 0]
    S, back = X.assemble_pullback(ov, ks, L, passivity="none")
    assert np.abs(S.S - S0).max() == 0
    ob, kb = back(Sbar)
    h = 1e-6
    for (i, which, r, c) in [(0, 0, 1, 2), (1, 1, 0, 1), (0, 1, 2, 2), (1, 0, 2, 0)]:
        for dz in (1.0, 1j):
            ovp = [(a.copy(), b.copy()) for a, b in ov]
            ovm = [(a.copy(), b.copy()) for a, b in ov]
            ovp[i][which][r, c] += h * dz
            ovm[i][which][r, c] -= h * dz
            fd = (f(fwd(ovp, ks)) - f(fwd(ovm, ks))) / (2 * h)
            assert np.real(np.conj(ob[i][which][r, c]) * dz) == pytest.approx(fd, rel=1e-6, abs=1e-9)
    for j in range(3):
        kp = [k.copy() for k in ks]
        km = [k.copy() for k in ks]
        kp[j][1] += h
        km[j][1] -= h
        fd = (f(fwd(ov, kp)) - f(fwd(ov, km))) / (2 * h)
        assert kb[j][1] == pytest.approx(fd, rel=1e-6)


def test_assemble_pullback_refuses_an_active_passivity_projection():
    n = 2
    A = 0.3 * np.eye(n, dtype=complex)   # overlaps far from unit: interface singular values > 1
    with pytest.raises(NotImplementedError):
        X.assemble_pullback([(A, 0.2 * A)], [np.ones(n), np.ones(n)], [1.0, 1.0], passivity="invert")
