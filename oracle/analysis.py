"""Group index and group-velocity dispersion of a solved mode (oracle; test infrastructure).

``ng_gvd`` restates ``ng_gvd`` / ``ng_gvd_E`` of ``lib/ModeAnalysis/src/group_index.jl``
(:55-95 and :126-162, identical arithmetic) for 2D grids and ``dk̂ = ẑ``: the group index
from the mode field and ``∂ε/∂ω``, and ``GVD = ∂n_g/∂ω`` from one adjoint solve
(``eig_adjt``) against the sensitivity of ``n_g`` to the eigenvector, closed with the
``k̄`` map of the basis (``∇ₖmag_mn``, ``src/maxwell.jl:670-685``).
"""
import numpy as np

from .maxwell import mag_mn, kx_tc, kx_ct, fft2, bfft2, ifft2, to_grid, to_flat, HelmholtzOracle
from .adjoint import eig_adjt, grad_k_mag_m_n, _cross


def _dot2(A, v):
    """``_dot(χ, v)`` (``src/linalg.jl:53-55``): ``out[i] = Σ_j χ[i, j] v[j]`` per pixel."""
    return np.einsum("ijxy,jxy->ixy", A, v)


def _dot3(A, B, C):
    """``_dot(A, B, C)`` (``src/linalg.jl:93-95``): per-pixel ``A B C``."""
    return np.einsum("ajxy,jkxy,kbxy->abxy", A, B, C)


def _outer(a, b):
    """``_outer(v1, v2)`` (``src/linalg.jl:300-304``): ``A[i, j] = v1[i] conj(v2[j])``."""
    return np.einsum("ixy,jxy->ijxy", a, np.conj(b))


def _tc(H, m, n):
    """``tc`` (``src/maxwell.jl:184-189``): ``H1 m + H2 n``."""
    return H[0][None] * m + H[1][None] * n


def _ct(h, m, n):
    """``ct`` (``src/maxwell.jl:191-196``): ``(h.m, h.n)``."""
    return np.stack([(h * m).sum(0), (h * n).sum(0)])


def _zx_ct(e, m, n):
    """``zx_ct`` (``src/maxwell.jl:231-240``): ``H[b] = -e_x mn[y, b] + e_y mn[x, b]``."""
    return np.stack([-e[0] * m[1] + e[1] * m[0], -e[0] * n[1] + e[1] * n[0]])


def ng_gvd(omega, k, ev, epsinv, deps, ddeps, grid, adj_rtol=1e-13, return_E=False):
    """``ng_gvd(ω, k, ev, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid)`` (``group_index.jl:126-162``);
    with ``return_E`` also the real-space ``E = ε⁻¹ fft(kx_tc(ev))`` of ``ng_gvd_E``
    (``:55-95``).  Returns ``(ng, gvd)`` or ``(ng, gvd, E)``."""
    w = float(omega)
    N = grid.N
    Ninv = 1.0 / N
    mag, m, n = mag_mn(k, grid)
    evg = to_grid(np.asarray(ev, dtype=np.complex128), grid)
    zhat = np.array([0.0, 0.0, 1.0])
    D = fft2(kx_tc(evg, m, n, mag))
    E = _dot2(epsinv, D)
    H = w * fft2(_tc(evg, m, n))
    inv_HMkH = 1.0 / (-np.real(np.vdot(to_flat(evg), to_flat(_zx_ct(ifft2(E), m, n)))))
    deps_E = _dot2(deps, E)
    epsi_deps_E = _dot2(epsinv, deps_E)
    Fi_epsi_deps_E = ifft2(epsi_deps_E)
    kx_Fi_epsi_deps_E = kx_ct(Fi_epsi_deps_E, m, n, mag)
    EdepsiE = np.real(np.vdot(to_flat(evg), to_flat(kx_Fi_epsi_deps_E)))
    ng = (w + EdepsiE / 2) * inv_HMkH
    dng_dEdepsiE = inv_HMkH / 2
    dng_dHMkH = -(w + EdepsiE / 2) * inv_HMkH ** 2
    # sensitivity of <ev|dM/dk|ev> expressed through E and H
    zb = zhat[:, None, None] * np.ones((1,) + E.shape[1:])
    Hbar = _cross(zb, E) * dng_dHMkH * Ninv / w
    Ebar = _cross(H, zb) * dng_dHMkH * Ninv / w
    Fi_ei_Ebar = bfft2(_dot2(epsinv, Ebar))
    Fi_Hbar = bfft2(Hbar)
    # adjoint solve with xbar = d ng / d ev
    op = HelmholtzOracle(k, epsinv, grid)
    xbar = (2 * dng_dEdepsiE) * to_flat(kx_Fi_epsi_deps_E) + to_flat(kx_ct(Fi_ei_Ebar, m, n, mag)) \
        + w * to_flat(_ct(Fi_Hbar, m, n))
    lam_v = eig_adjt(op.M, w ** 2, np.asarray(ev, dtype=np.complex128), 0.0, xbar, P=op.P, rtol=adj_rtol)
    lam = to_grid(lam_v, grid)
    lamd = fft2(kx_tc(lam, m, n, mag))
    lam_e = ifft2(_dot2(epsinv, lamd))
    e_t = ifft2(E)
    rev_evg = evg[::-1]
    Fi_ei_Ebar_x_evgT = np.einsum("ixy,jxy->ijxy", np.conj(Fi_ei_Ebar), rev_evg)
    X = (2 * dng_dEdepsiE) * Fi_epsi_deps_E - lam_e
    dng_dkx = (np.real(_outer(X, evg)) - np.real(_outer(e_t, lam)))[:, ::-1] + np.real(Fi_ei_Ebar_x_evgT)
    one_mone = np.array([1.0, -1.0])
    mn = np.stack([m, n], 1)  # (3, 2, Nx, Ny)
    dng_dmag = np.einsum("ijxy,ijxy,j->xy", dng_dkx, mn, one_mone)
    dng_dmn = np.einsum("ijxy,xy,j->ijxy", dng_dkx, mag, one_mone) + w * np.real(_outer(Fi_Hbar, evg))
    dng_dk = grad_k_mag_m_n(np.real(dng_dmag), np.real(dng_dmn[:, 0]), np.real(dng_dmn[:, 1]), mag, m, n)
    EE = _outer(E, E)
    herm_EE = 0.5 * np.real(EE + np.conj(np.swapaxes(EE, 0, 1)))
    inner = (2 * dng_dEdepsiE * Ninv) * deps_E + Ebar \
        - Ninv * (lamd + fft2(kx_tc((dng_dk * inv_HMkH / 2) * evg, m, n, mag)))
    gvd = (dng_dEdepsiE * Ninv) * np.sum(ddeps * herm_EE) + inv_HMkH * (w * dng_dk + 1.0) \
        - np.sum(deps * _dot3(epsinv, np.real(_outer(inner, D)), epsinv))
    if return_E:
        return float(np.real(ng)), float(np.real(gvd)), E
    return float(np.real(ng)), float(np.real(gvd))


def gaussian_wg_eps_dispersive(omega, grid):
    """Analytic dispersive Gaussian index bump of the reference's ModeAnalysis tests
    (``lib/ModeAnalysis/test/runtests.jl:20-41``): ``ε_core = 4 + 2ω²``,
    ``ε_bg = 1.5 + 0.4ω²``, widths ``(1.0, 0.6)``.  Returns ``(ε, ∂ε/∂ω, ∂²ε/∂ω²)``, each
    ``(3, 3, Nx, Ny)``."""
    w = float(omega)
    xs, ys = grid.x(), grid.y()
    G = np.exp(-(xs[:, None] ** 2 / 1.0 ** 2 + ys[None, :] ** 2 / 0.6 ** 2))
    vals = [(1.5 + 0.4 * w * w) + ((4.0 + 2.0 * w * w) - (1.5 + 0.4 * w * w)) * G,
            0.8 * w + (4.0 * w - 0.8 * w) * G,
            0.8 + (4.0 - 0.8) * G]
    out = []
    for v in vals:
        a = np.zeros((3, 3, grid.Nx, grid.Ny))
        for i in range(3):
            a[i, i] = v
        out.append(a)
    return tuple(out)
