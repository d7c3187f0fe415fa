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
