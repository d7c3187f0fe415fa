"""Pins the oracle's group index / GVD (oracle.analysis.ng_gvd, a restatement of
lib/ModeAnalysis/src/group_index.jl:126-162) against finite differences of full re-solves
-- the reference's own check (lib/ModeAnalysis/test/runtests.jl:132-146, test/runtests.jl:
232-245), here at much tighter tolerances -- on its analytic dispersive Gaussian guide and
on a rotated, fully anisotropic variant, and against the closed form of a homogeneous
dispersive medium."""
import numpy as np
import pytest

import oracle as O

W0 = 1 / 1.55
G16 = O.Grid(6.0, 4.0, 16, 16)   # lib/ModeAnalysis/test/runtests.jl:102


def _rot():
    a, b = 0.4, 0.3
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    return Rz @ Ry


def aniso_dispersive(omega, grid):
    """Gaussian guide whose three principal permittivities disperse differently, rotated so
    that every tensor entry (xy, xz, yz included) is non-zero: (ε, ∂ε/∂ω, ∂²ε/∂ω²)."""
    w = float(omega)
    xs, ys = grid.x(), grid.y()
    G = np.exp(-(xs[:, None] ** 2 / 1.0 ** 2 + ys[None, :] ** 2 / 0.6 ** 2))
    tabs = [[(1.5 + 0.4 * w * w, 4.0 + 2.0 * w * w), (1.6 + 0.3 * w * w, 3.5 + 2.5 * w * w),
             (1.4 + 0.5 * w * w, 4.4 + 1.5 * w * w)],
            [(0.8 * w, 4.0 * w), (0.6 * w, 5.0 * w), (1.0 * w, 3.0 * w)],
            [(0.8, 4.0), (0.6, 5.0), (1.0, 3.0)]]
    R = _rot()
    out = []
    for tab in tabs:
        d = np.stack([bg + (c - bg) * G for (bg, c) in tab])
        out.append(np.einsum("ai,ixy,bi->abxy", R, d, R))
    return tuple(out)


def _solve(profile, omega, grid, nev):
    eps, de, dde = profile(omega, grid)
    ei = O.sliceinv_3x3(eps)
    km, evs = O.solve_k(omega, ei, grid, nev=nev, k_tol=1e-13, eig_tol=1e-12, maxiter=300)
    return km, evs, ei, de, dde


def _fd(f, x, h):
    return (f(x - 2 * h) - 8 * f(x - h) + 8 * f(x + h) - f(x + 2 * h)) / (12 * h)


@pytest.mark.parametrize("profile,band", [(O.gaussian_wg_eps_dispersive, 0), (aniso_dispersive, 0),
                                          (aniso_dispersive, 1)],
                         ids=["isotropic", "aniso-band0", "aniso-band1"])
def test_ng_gvd_matches_fd_of_resolves(profile, band):
    nev = 2  # a two-column block converges band 0 fast where bands 0 and 1 are close
    km, evs, ei, de, dde = _solve(profile, W0, G16, nev)
    ng, gvd = O.ng_gvd(W0, km[band], evs[band], ei, de, dde, G16)
    # n_g of ng_gvd is group_index's (group_index.jl:27 vs :145)
    assert ng == pytest.approx(O.group_index(km[band], evs[band], W0, ei, de, G16), rel=1e-13)

    def k_of(w):
        return _solve(profile, w, G16, nev)[0][band]

    def ng_of(w):
        km, evs, ei, de, _ = _solve(profile, w, G16, nev)
        return O.group_index(km[band], evs[band], w, ei, de, G16)

    h = 2e-3
    assert ng == pytest.approx(_fd(k_of, W0, h), rel=1e-9)
    assert gvd == pytest.approx(_fd(ng_of, W0, h), rel=1e-8)


def test_ng_gvd_homogeneous_closed_form():
    """Uniform ε(ω) = a + bω²: the fundamental mode is the G = 0 plane wave, k = ω n with
    n = sqrt(ε); n_g = n + ω n' and GVD = 2 n' + ω n''."""
    a, b = 2.1, 0.7
    g = O.Grid(1.0, 1.0, 8, 8)
    eps = a + b * W0 * W0
    n = np.sqrt(eps)
    e1, e2 = 2 * b * W0, 2 * b
    n1 = e1 / (2 * n)
    n2 = e2 / (2 * n) - e1 * e1 / (4 * n ** 3)
    I3 = np.eye(3)[:, :, None, None] * np.ones((1, 1, 8, 8))
    ei = I3 / eps
    # the G = 0 plane wave in either polarisation (and any mix) is an exact eigenvector
    assert np.allclose(O.apply_M(_plane_wave(g, 0.6, 0.8), W0 * n, ei, g), W0 ** 2 * _plane_wave(g, 0.6, 0.8))
    for c1, c2 in ((1.0, 0.0), (0.0, 1.0), (0.6, 0.8j)):
        ng, gvd = O.ng_gvd(W0, W0 * n, _plane_wave(g, c1, c2), ei, e1 * I3, e2 * I3, g)
        assert ng == pytest.approx(n + W0 * n1, rel=1e-12)
        assert gvd == pytest.approx(2 * n1 + W0 * n2, rel=1e-9)


def _plane_wave(g, c1, c2):
    ev = np.zeros(2 * g.N, dtype=complex)
    ev[0], ev[1] = c1, c2
    return ev / np.linalg.norm(ev)


def test_ng_gvd_E_is_the_field_of_ng_gvd_E():
    km, evs, ei, de, dde = _solve(O.gaussian_wg_eps_dispersive, W0, G16, 1)
    ng, gvd, E = O.ng_gvd(W0, km[0], evs[0], ei, de, dde, G16, return_E=True)
    assert (ng, gvd) == O.ng_gvd(W0, km[0], evs[0], ei, de, dde, G16)
    assert np.allclose(E, O.E_field(km[0], evs[0], ei, G16), rtol=0, atol=1e-12 * np.abs(E).max())
