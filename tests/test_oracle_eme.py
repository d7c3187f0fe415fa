"""Pins the oracle's EME mode fields and MEOW inner products (oracle.eme, a restatement of
lib/EigenmodeExpansion/src/modes.jl:95-170) by physics the formulas must satisfy:
E and H are a consistent Maxwell pair, the power-normalised self overlap is 1
(lib/EigenmodeExpansion/test/runtests.jl:81), distinct modes of one cross-section are
orthogonal in the unconjugated product, and the energy velocity P/W equals the group
velocity 1/n_g of group_index."""
import numpy as np
import pytest

import oracle as O
from oracle.eme import mode_fields, inner_product, build_mode, overlap_matrix

W0 = 1 / 1.55
G = O.Grid(6.0, 4.0, 32, 24)


@pytest.fixture(scope="module")
def modes():
    ei = O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], G)
    km, evs = O.solve_k(W0, ei, G, nev=3, k_tol=1e-12, eig_tol=1e-11, maxiter=300)
    return ei, km, evs


def test_fields_are_a_maxwell_pair(modes):
    """H = fft(tc(h)) with h recovered from E through the operator's back half:
    kx_ct(bfft(-i E)) / N = M h = ω² h."""
    ei, km, evs = modes
    for k, v in zip(km, evs):
        E, H = mode_fields(k, v, ei, G)
        mag, m, n = O.mag_mn(k, G)
        h = O.kx_ct(O.bfft2(-1j * E), m, n, mag) / (G.N * W0 ** 2)
        Hh = O.fft2(h[0][None] * m + h[1][None] * n)
        assert np.abs(Hh - H).max() <= 1e-9 * np.abs(H).max()


def test_energy_velocity_is_group_velocity(modes):
    """|P| / W = 1/n_g with P = ½ Σ (E* × H)·ẑ δA and W = ¼ Σ (E*·εE + |H|²) δA, the fields
    carrying mode_fields' factor ω (E = ω E_phys)."""
    ei, km, evs = modes
    eps = O.sliceinv_3x3(ei)
    dA = G.dV()
    for k, v in zip(km, evs):
        E, H = mode_fields(k, v, ei, G)
        P = abs(inner_product(E, H, E, H, dA, conjugate=True)) / W0
        W = 0.25 * dA * (np.real(np.einsum("axy,abxy,bxy->", np.conj(E), eps, E)) / W0 ** 2 + np.sum(np.abs(H) ** 2))
        ng = O.group_index(k, v, W0, ei, np.zeros_like(ei), G)
        assert P / W == pytest.approx(1.0 / ng, rel=1e-10)


def test_normalised_self_overlap_and_orthogonality(modes):
    ei, km, evs = modes
    dA = G.dV()
    built = [build_mode(k, v, ei, G) for k, v in zip(km, evs)]
    Om = overlap_matrix([(b[0], b[1]) for b in built], [(b[0], b[1]) for b in built], dA)
    assert np.allclose(np.diag(Om), 1.0, atol=1e-13)
    assert np.abs(Om - np.diag(np.diag(Om))).max() < 1e-8
    # the conjugated product of the same power-normalised modes has unit modulus
    Oc = overlap_matrix([(b[0], b[1]) for b in built], [(b[0], b[1]) for b in built], dA, conjugate=True)
    assert np.allclose(np.abs(np.diag(Oc)), 1.0, atol=1e-9)
    # build_mode: nrm² is the unnormalised self product
    E, H = mode_fields(km[0], evs[0], ei, G)
    assert built[0][2] ** 2 == pytest.approx(inner_product(E, H, E, H, dA))


# ---------------------------------------------------------------------------------------
# pull-backs of the EME chain's per-slice steps, pinned by finite differences of re-solves
# ---------------------------------------------------------------------------------------

from oracle.eme import mode_fields_pullback, normalization_pullback, overlap_pullback  # noqa: E402

GS = O.Grid(6.0, 4.0, 24, 16)
SLICE_P = ([4.2, 2.1, 1.0, 0.6], [4.0, 2.1, 1.2, 0.55])


def _slice(ei, conjugate):
    km, evs = O.solve_k(W0, ei, GS, nev=2, k_tol=1e-13, eig_tol=1e-11, maxiter=300, seed=3)
    built = [build_mode(k, v, ei, GS, conjugate) for k, v in zip(km, evs)]
    return km, evs, built


def _objective(b0, b1, conjugate):
    """A phase-invariant scalar of the two interface overlap matrices."""
    dA = GS.dV()
    OLR = overlap_matrix([(b[0], b[1]) for b in b0], [(b[0], b[1]) for b in b1], dA, conjugate)
    ORL = overlap_matrix([(b[0], b[1]) for b in b1], [(b[0], b[1]) for b in b0], dA, conjugate)
    wl, wr = np.array([[1.0, 0.7], [0.4, 1.3]]), np.array([[0.5, 1.1], [0.9, 0.2]])
    L = np.sum(wl * np.abs(OLR) ** 2) + np.sum(wr * np.abs(ORL) ** 2)
    return L, OLR, ORL, 2 * wl * OLR, 2 * wr * ORL


@pytest.mark.parametrize("conjugate", [False, True])
def test_slice_pullback_matches_fd(conjugate):
    """Overlap → normalisation → mode fields → solve_k pull-back of one slice of a two-slice
    chain equals central differences of the whole chain re-solved along symmetric ε⁻¹
    directions of that slice (diagonal and off-diagonal entries)."""
    eis = [O.gaussian_wg_epsi(p, GS) for p in SLICE_P]
    s0, s1 = _slice(eis[0], conjugate), _slice(eis[1], conjugate)
    L, OLR, ORL, ObLR, ObRL = _objective(s0[2], s1[2], conjugate)
    dA = GS.dV()
    m0 = [(b[0], b[1]) for b in s0[2]]
    m1 = [(b[0], b[1]) for b in s1[2]]
    c0a, c1a = overlap_pullback(m0, m1, dA, ObLR, conjugate)
    c1b, c0b = overlap_pullback(m1, m0, dA, ObRL, conjugate)
    grads = []
    for (km, evs, built), ei, ca, cb in ((s0, eis[0], c0a, c0b), (s1, eis[1], c1a, c1b)):
        kbar, evbar, direct = [], [], np.zeros_like(ei)
        for j, (E, H, nrm) in enumerate(built):
            Eb, Hb = ca[j][0] + cb[j][0], ca[j][1] + cb[j][1]
            Erb, Hrb = normalization_pullback(E, H, nrm, dA, Eb, Hb, conjugate)
            kb, eb, ed = mode_fields_pullback(km[j], evs[j], ei, GS, Erb, Hrb)
            kbar.append(kb)
            evbar.append(eb)
            direct += ed
        _, eib = O.solve_k_pullback(W0, ei, GS, km, evs, kbar, evbar, adj_rtol=1e-11)
        grads.append((direct, eib))
    rng = np.random.default_rng(31)
    for s, (a, b) in ((0, (0, 0)), (0, (0, 1)), (1, (1, 2))):
        if True:
            r = rng.standard_normal(GS.shape) * 0.02
            D = np.zeros_like(eis[s])
            D[a, b] = r
            D[b, a] = r
            direct, eib = grads[s]
            # the solve_k rule holds the symmetric-pair derivative in (a, b) and (b, a); the
            # field rule (the contraction's own) holds each entry's own derivative
            pred = np.sum(direct * D) + (np.sum(eib[a, b] * r))
            h = 1e-3

            def L_at(t):
                e = [x.copy() for x in eis]
                e[s] = e[s] + t * D
                q0 = _slice(e[0], conjugate) if s == 0 else s0
                q1 = _slice(e[1], conjugate) if s == 1 else s1
                return _objective(q0[2], q1[2], conjugate)[0]
            fd = (L_at(h) - L_at(-h)) / (2 * h)
            print(s, (a, b), pred, fd, flush=True)
            # the re-solves' own accuracy (eig_tol 1e-11) bounds the difference quotient to ~1e-7
            assert pred == pytest.approx(fd, rel=2e-5, abs=2e-7), (s, a, b, pred, fd)
