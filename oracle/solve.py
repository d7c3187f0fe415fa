"""``solve_ω²`` with DFTK-LOBPCG and the Newton ``solve_k`` (oracle; test infrastructure)."""
import numpy as np

from .maxwell import HelmholtzOracle, k_guess, canonicalize_phase
from .lobpcg import lobpcg
from .periodic import Grid3, HelmholtzOracle3, canonicalize_phase3


def _helmholtz(k, epsinv, grid):
    """The operator bundle of the grid's dimension (the reference dispatches on
    ``Grid{2}`` / ``Grid{3}``)."""
    return HelmholtzOracle3(k, epsinv, grid) if isinstance(grid, Grid3) else HelmholtzOracle(k, epsinv, grid)


def _canonicalize(ev, k, epsinv, grid):
    return (canonicalize_phase3 if isinstance(grid, Grid3) else canonicalize_phase)(ev, k, epsinv, grid)


def solve_omega2(k, epsinv, grid, H0, tol=1e-8, maxiter=100, n_conv_check=None, seed=0):
    """``solve_ω²(ms, DFTK_LOBPCG(); tol, maxiter)`` (``solvers/dftk.jl:32-37``) at
    ``k_vec = (0, 0, k)`` (``update_k!``, ``maxwell.jl:1098-1104``).

    ``H0`` is the ``(2N, nev)`` starting block (``ms.H⃗``).  Returns
    ``(omega2[nev], X (2N, nev), info)``; ``X`` is what the reference writes back to
    ``ms.H⃗``."""
    op = _helmholtz(k, epsinv, grid)
    res = lobpcg(op.M, np.array(H0, dtype=np.complex128), precon=op.P, tol=tol, maxiter=maxiter,
                 n_conv_check=n_conv_check, seed=seed)
    return res["lam"].copy(), res["X"].copy(), res


def solve_k_single(omega, epsinv, grid, state, eigind, nev, k_tol=1e-8, eig_tol=1e-8,
                   max_eigsolves=60, maxiter=100, n_conv_check=None):
    """One band of ``solve_k``: ``Roots.find_zero(_solve_Δω², k0, Newton(); atol=k_tol,
    maxevals=max_eigsolves)`` (``solve.jl:77-101``).

    ``state`` is a dict with ``k`` (last k used, the Newton start) and ``H`` (the
    warm-start block), both updated in place like ``ms.M̂.k⃗``/``ms.H⃗``.  Newton stops
    when ``|Δω²| <= k_tol`` (Roots' ``atol`` on the function value), returning the
    ``k`` where that value was computed together with its eigenvector.  Returns
    ``(k, evec, n_eigsolves, lobpcg_iters)``.
    """
    w2 = float(omega) ** 2
    x = float(state["k"])
    iters = []
    for nevals in range(1, max_eigsolves + 1):
        lam, X, info = solve_omega2(x, epsinv, grid, state["H"], tol=eig_tol, maxiter=maxiter,
                                    n_conv_check=n_conv_check, seed=nevals)
        iters.append(info["niter"])
        state["H"] = X
        state["k"] = x
        ev = X[:, eigind].copy()
        f = lam[eigind] - w2
        slope = 2.0 * _helmholtz(x, epsinv, grid).HMkH(ev)
        if abs(f) <= k_tol:
            return x, ev, nevals, iters
        x_new = x - f / slope
        if abs(x_new - x) <= 4 * np.finfo(float).eps * max(abs(x), 1.0):
            return x, ev, nevals, iters
        x = x_new
    raise RuntimeError(f"solve_k: band {eigind} did not converge in {max_eigsolves} eigensolves")


def solve_k(omega, epsinv, grid, nev=1, k_tol=1e-8, eig_tol=1e-8, max_eigsolves=60, maxiter=100,
            kguess=None, Hguess=None, seed=0, return_info=False, lock_band_only=False):
    """``solve_k(ω, ε⁻¹, grid, DFTK_LOBPCG(); nev, ...)`` (``solve.jl:182-196`` ->
    ``:104-115``): start at ``k_guess`` (or ``kguess``) with a random ``(2N, nev)`` block
    (``ModeSolver``, ``maxwell.jl:956-972``; columns replaced by ``Hguess``,
    ``solve.jl:201-210``), run one Newton solve per band, each warm-started from the
    previous band's ``k`` and block, and canonicalise every eigenvector's phase.

    ``lock_band_only=True`` stops each eigensolve once bands ``1..eigind`` are locked
    (DFTK's ``n_conv_check``) instead of all ``nev``; the converged pairs are the same,
    only the work differs."""
    rng = np.random.default_rng(seed)
    n2 = 2 * grid.N
    H = rng.standard_normal((n2, nev)) + 1j * rng.standard_normal((n2, nev))
    if Hguess is not None:
        cols = [Hguess] if np.ndim(Hguess) == 1 else [Hguess[:, j] for j in range(np.shape(Hguess)[1])]
        for j, h in enumerate(cols):
            if j >= nev:
                break
            if len(h) == n2:
                H[:, j] = h
    k0 = k_guess(omega, epsinv) if kguess is None else float(np.ravel([kguess])[0])
    state = dict(k=k0, H=H)
    kmags, evecs, infos = [], [], []
    for eigind in range(nev):
        k, ev, nsolves, iters = solve_k_single(omega, epsinv, grid, state, eigind, nev, k_tol, eig_tol,
                                               max_eigsolves, maxiter,
                                               n_conv_check=(eigind + 1) if lock_band_only else None)
        kmags.append(k)
        evecs.append(_canonicalize(ev, k, epsinv, grid))
        infos.append(dict(n_eigsolves=nsolves, lobpcg_iters=iters))
    if return_info:
        return np.array(kmags), evecs, infos
    return np.array(kmags), evecs
