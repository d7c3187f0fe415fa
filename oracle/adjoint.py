"""Adjoint (pull-back) of ``solve_k`` and its helpers, FP64 (oracle; test infrastructure)."""
import numpy as np
import scipy.sparse.linalg as spla

from .maxwell import (mag_mn, kx_tc, eid, fft2, bfft2, to_grid, to_flat, HelmholtzOracle,
                      HMkH, HMH)


def eig_adjt(A, alpha, x, alphabar, xbar, P=None, rtol=1e-13, maxiter=1000):
    """``eig_adjt`` (``grads/solve.jl:96-113``): solve ``(A - αI) λx = x̄ - x<x, x̄>``
    (GMRES with left preconditioner ``P``, ``my_linsolve`` ``:56-66``), re-orthogonalise
    against ``x`` and add ``ᾱ x``.  ``A`` is a dense matrix or a callable."""
    x = np.asarray(x)
    if xbar is None or not np.any(xbar):
        return np.zeros_like(x) if alphabar == 0 else alphabar * x
    b = xbar - x * np.vdot(x, xbar)
    n = x.shape[0]
    if callable(A):
        mv = lambda v: A(v) - alpha * v  # noqa: E731
    else:
        mv = lambda v: A @ v - alpha * v  # noqa: E731
    op = spla.LinearOperator((n, n), matvec=mv, dtype=np.complex128)
    Mop = None
    if P is not None:
        Mop = spla.LinearOperator((n, n), matvec=P, dtype=np.complex128)
    lam, info = spla.gmres(op, b, M=Mop, rtol=rtol, atol=0.0, restart=min(60, n), maxiter=maxiter)
    lam = lam - x * np.vdot(x, lam)
    return lam + alphabar * x


def epsinv_bar(d, lamd):
    """2D ``ε⁻¹_bar`` (``grads/solve.jl:490-515``): diagonal ``-Re(λd_a conj d_a)``;
    off-diagonal ``(i,j)`` and ``(j,i)`` both ``-Re(λd_j conj d_i + λd_i conj d_j)``.
    ``d``, ``lamd`` have shape ``(3, Nx, Ny)``; returns ``(3, 3, Nx, Ny)``."""
    out = np.zeros((3, 3) + d.shape[1:])
    for a in range(3):
        out[a, a] = -np.real(lamd[a] * np.conj(d[a]))
    for i, j in ((0, 1), (1, 2), (0, 2)):
        v = -np.real(lamd[j] * np.conj(d[i]) + lamd[i] * np.conj(d[j]))
        out[i, j] = v
        out[j, i] = v
    return out


def _cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def grad_k_mag_m_n(magbar, mbar, nbar, mag, m, n):
    """``∇ₖmag_m_n`` with ``dk̂ = ẑ`` (``maxwell.jl:560-566``): the ``k̄`` contribution of
    cotangents ``(māg, m̄, n̄)`` of the basis fields."""
    kpg_over_mag = _cross(m, n) / mag[None]
    dk = np.array([0.0, 0.0, 1.0])[:, None, None] * np.ones_like(m)
    kbar_mag = np.sum(magbar * kpg_over_mag[2] * mag)
    q = _cross(kpg_over_mag, dk)
    kbar_m = -np.sum(mbar * _cross(m, q))
    kbar_n = -np.sum(nbar * _cross(n, q))
    return kbar_mag + kbar_m + kbar_n


def solve_k_pullback(omega, epsinv, grid, kmags, evecs, kbar, evbar, adj_rtol=1e-13, return_parts=False):
    """Pull-back of ``rrule(solve_k)`` (``solve.jl:213-289``): maps ``(k̄mags, ēvecs)`` to
    ``(ω̄, ε̄⁻¹)``.  ``kbar`` entries may be ``None`` (zero); ``evbar`` may be ``None`` or a
    list with ``None`` entries."""
    nev = len(kmags)
    N = grid.N
    ei_bar = np.zeros_like(epsinv, dtype=float)
    w_bar = 0.0
    parts = []
    for i in range(nev):
        k = float(kmags[i])
        ev = np.asarray(evecs[i], dtype=np.complex128)
        kb = 0.0 if (kbar is None or kbar[i] is None) else float(kbar[i])
        eb = None if (evbar is None or evbar[i] is None) else np.asarray(evbar[i], dtype=np.complex128)
        op = HelmholtzOracle(k, epsinv, grid)
        mag, m, n = op.basis
        dw2dk = 2.0 * op.HMkH(ev)
        evg = to_grid(ev, grid)
        kh = 0.0
        if eb is not None and np.any(eb):
            lam = eig_adjt(op.M, omega ** 2, ev, 0.0, eb, P=op.P, rtol=adj_rtol)
            lam = lam - np.vdot(ev, lam) * ev
            lg = to_grid(lam, grid)
            d = fft2(kx_tc(evg / N, m, n, mag))
            lamd = fft2(kx_tc(lg, m, n, mag))
            ei_bar += epsinv_bar(d, lamd)
            lamd = lamd / N
            lam_e = bfft2(eid(epsinv, lamd))
            e_t = bfft2(eid(epsinv, d))
            kxbar_m = np.real(lam_e * np.conj(evg[1])[None] + e_t * np.conj(lg[1])[None])
            kxbar_n = -np.real(lam_e * np.conj(evg[0])[None] + e_t * np.conj(lg[0])[None])
            magbar = (n * kxbar_n).sum(0) + (m * kxbar_m).sum(0)
            kh = -grad_k_mag_m_n(magbar, kxbar_m * mag[None], kxbar_n * mag[None], mag, m, n)
        scale = (kb + kh) / dw2dk
        d = fft2(kx_tc(evg / N, m, n, mag))
        lamd = fft2(kx_tc(scale * evg, m, n, mag))
        ei_bar += epsinv_bar(d, lamd)
        w_bar += 2 * omega * scale
        parts.append(dict(kh=kh, dw2dk=dw2dk))
    if return_parts:
        return w_bar, ei_bar, parts
    return w_bar, ei_bar


def sliceinv_3x3(A):
    """Per-pixel 3x3 inverse (``linalg.jl:126-149``)."""
    Am = np.moveaxis(A, (0, 1), (-2, -1))
    return np.moveaxis(np.linalg.inv(Am), (-2, -1), (0, 1))


def sliceinv_3x3_pullback(Y, Ybar):
    """``Ā = -Yᵀ Ȳ Yᵀ`` per pixel (``grads/linalg.jl:40-47``), ``Y = A⁻¹``."""
    Ym = np.moveaxis(Y, (0, 1), (-2, -1))
    Yb = np.moveaxis(Ybar, (0, 1), (-2, -1))
    YT = np.swapaxes(Ym, -1, -2)
    return np.moveaxis(-(YT @ Yb @ YT), (-2, -1), (0, 1))


def group_index(k, ev, omega, epsinv, deps_dw, grid):
    """``group_index`` (``lib/ModeAnalysis/src/group_index.jl:21-29``):
    ``n_g = (ω + HMH(ev, ε⁻¹ ∂ωε ε⁻¹)/2) / HMkH(ev)``."""
    basis = mag_mn(k, grid)
    Ei = np.moveaxis(epsinv, (0, 1), (-2, -1))
    Dw = np.moveaxis(deps_dw, (0, 1), (-2, -1))
    A = np.moveaxis(Ei @ Dw @ Ei, (-2, -1), (0, 1))
    return (omega + HMH(ev, k, A, grid, basis) / 2) / HMkH(ev, k, epsinv, grid, basis)
