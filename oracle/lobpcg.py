"""Block LOBPCG with Cholesky orthogonalisation and in-order locking, FP64 (oracle).

Restates the vendored DFTK solver in
``lib/MaxwellEigenmodes/src/solvers/dftk.jl:129-506`` for ``B = I`` (the only case the
Helmholtz path uses, ``dftk.jl:33``).  Test infrastructure only.
"""
import numpy as np

EPS64 = np.finfo(np.float64).eps


class LobpcgError(RuntimeError):
    pass


def _normest(M):
    """``normest`` (``dftk.jl:146``): max|diag| + Frobenius norm of the off-diagonal."""
    d = np.diag(M)
    off = M - np.diag(d)
    return np.max(np.abs(d)) + np.linalg.norm(off)


def _chol_upper(O):
    """Upper Cholesky factor ``U`` with ``O = U^H U``; raises LinAlgError if not PD."""
    L = np.linalg.cholesky(O)
    return L.conj().T


def ortho(X, tol=2 * EPS64):
    """Cholesky orthonormalisation of the columns of ``X`` in place
    (``ortho!``, ``dftk.jl:151-221``).  Returns ``(X, nchol, growth_factor)``."""
    growth = 1.0
    nchol = 0
    while True:
        O = X.conj().T @ X
        O = (O + O.conj().T) / 2
        try:
            R = _chol_upper(O)
            nchol += 1
            success = True
        except np.linalg.LinAlgError:
            alpha = 100.0
            nbad = 0
            nrmX2 = np.linalg.norm(X) ** 2
            while True:
                O = O + alpha * EPS64 * nrmX2 * np.eye(O.shape[0])
                alpha *= 10
                try:
                    R = _chol_upper(O)
                    nchol += 1
                    break
                except np.linalg.LinAlgError:
                    pass
                nbad += 1
                if nbad > 10:
                    raise LobpcgError("Cholesky shifting is failing badly")
            success = False
        invR = np.linalg.inv(R)
        X[...] = X @ invR
        norminvR = _normest(invR)
        growth *= norminvR
        condR = _normest(R) * norminvR
        if success and EPS64 * condR ** 2 < tol:
            break
        if nchol > 10:
            raise LobpcgError("Ortho(X) is failing badly")
    return X, nchol, growth


def _drop(X, rng, tol=2 * EPS64):
    """``drop!`` (``dftk.jl:224-234``): randomise columns whose norm is <= tol."""
    dropped = []
    for i in range(X.shape[1]):
        if np.linalg.norm(X[:, i]) <= tol:
            X[:, i] = rng.standard_normal(X.shape[0]) + (1j * rng.standard_normal(X.shape[0]) if np.iscomplexobj(X) else 0)
            dropped.append(i)
    return dropped


def ortho_against(X, Y, BY, rng, tol=2 * EPS64):
    """Orthonormalise ``X`` and make it orthogonal to ``Y`` (``ortho!(X, Y, BY)``,
    ``dftk.jl:238-285``), in place."""
    for i in range(X.shape[1]):
        X[:, i] /= np.linalg.norm(X[:, i])
    niter = 1
    while True:
        BYX = BY.conj().T @ X
        X -= Y @ BYX
        dropped = _drop(X, rng)
        if dropped:
            X[:, dropped] -= Y @ (BY.conj().T @ X[:, dropped])
        if np.linalg.norm(BYX) < tol and niter > 1:
            break
        _, _, growth = ortho(X, tol)
        if growth * EPS64 < tol:
            break
        if niter > 10:
            raise LobpcgError("Ortho(X,Y) is failing badly")
        niter += 1
    return X


def rayleigh_ritz(Y, AY, N):
    """``rayleigh_ritz`` (``dftk.jl:129-132``)."""
    G = Y.conj().T @ AY
    G = (G + G.conj().T) / 2
    w, V = np.linalg.eigh(G)
    return V[:, :N], w[:N]


def lobpcg(A, X0, precon=None, tol=1e-10, maxiter=100, miniter=1, n_conv_check=None,
           ortho_tol=2 * EPS64, seed=0):
    """``LOBPCG(A, X, I, precon, tol, maxiter)`` (``dftk.jl:304-506``).

    ``A`` and ``precon`` are callables acting on ``(n, m)`` blocks.  Returns a dict with
    ``lam``, ``X``, ``residual_norms``, ``residual_history``, ``n_matvec``, ``niter``.
    """
    rng = np.random.default_rng(seed)
    N, M = X0.shape
    if n_conv_check is None:
        n_conv_check = M
    resid_history = np.zeros((M, maxiter + 1))
    full_X = np.array(X0, dtype=np.complex128, copy=True)
    ortho(full_X, ortho_tol)
    n_matvec = M
    full_AX = A(full_X)
    P = np.zeros_like(full_X)
    AP = np.zeros_like(full_X)
    R = np.zeros_like(full_X)
    AR = np.zeros_like(full_X)
    nlocked = 0
    niter = 0
    # views onto the active part (columns nlocked..M-1)
    X, AX = full_X, full_AX
    lam = np.array([np.real(np.vdot(X[:, i], AX[:, i]) / np.vdot(X[:, i], X[:, i])) for i in range(M)])
    new_X, new_AX = X, AX
    cX = None
    Y = AY = None
    while True:
        if niter > 0:
            AR = A(R)
            n_matvec += R.shape[1]
            if niter > 1:
                Y = np.hstack([X, R, P])
                AY = np.hstack([AX, AR, AP])
            else:
                Y = np.hstack([X, R])
                AY = np.hstack([AX, AR])
            cX, lam = rayleigh_ritz(Y, AY, M - nlocked)
            new_X = Y @ cX
            new_AX = AY @ cX
        new_R = new_AX - new_X * lam[None, :]
        for i in range(X.shape[1]):
            resid_history[i + nlocked, niter] = np.linalg.norm(new_R[:, i])
        if precon is not None:
            new_R = precon(new_R)
        prev_nlocked = nlocked
        if niter >= miniter:
            for i in range(nlocked, M):
                if resid_history[i, niter] < tol:
                    nlocked += 1
                else:
                    break
        if nlocked >= n_conv_check:
            full_X[:, prev_nlocked:] = new_X
            full_AX[:, prev_nlocked:] = new_AX
            return _final(full_X, full_AX, resid_history, niter, n_matvec)
        newly_locked = nlocked - prev_nlocked
        if niter > 0:
            # cP = cX - e with e[Xn_indices[i], i] = 1 (dftk.jl:426-436), then restricted
            # to the columns Xn_indices and orthogonalised against cX.
            nact = M - prev_nlocked
            xn = list(range(newly_locked, nact))
            e = np.zeros((cX.shape[0], nact), dtype=cX.dtype)
            for i in range(len(xn)):
                e[xn[i], i] = 1.0
            cP = (cX - e)[:, xn].copy()
            ortho_against(cP, cX, cX, rng, ortho_tol)
            new_P = Y @ cP
            new_AP = AY @ cP
        full_X[:, prev_nlocked:] = new_X
        full_AX[:, prev_nlocked:] = new_AX
        for i in range(prev_nlocked, M):
            if abs(np.vdot(full_X[:, i], full_X[:, i]) - 1) >= np.sqrt(EPS64):
                raise LobpcgError("LOBPCG is badly failing to keep the vectors normalized")
        # restrict to the active window
        X = full_X[:, nlocked:]
        AX = full_AX[:, nlocked:]
        R = new_R[:, newly_locked:].copy()
        AR = AR[:, newly_locked:] if niter > 0 else np.zeros_like(R)
        P = P[:, newly_locked:]
        AP = AP[:, newly_locked:]
        lam = lam[newly_locked:]
        if niter > 0:
            P = new_P
            AP = new_AP
        Z = np.hstack([full_X, P]) if niter > 0 else full_X
        ortho_against(R, Z, Z, rng, ortho_tol)
        if niter >= maxiter:
            break
        niter += 1
    return _final(full_X, full_AX, resid_history, maxiter, n_matvec)


def _final(X, AX, resid_history, niter, n_matvec):
    """``final_retval`` (``dftk.jl:288-295``)."""
    lam = np.real(np.einsum("ij,ij->j", X.conj(), AX))
    res = AX - X * lam[None, :]
    return dict(lam=lam, X=X, residual_norms=np.linalg.norm(res, axis=0),
                residual_history=resid_history[:, :niter + 1], n_matvec=n_matvec, niter=niter)
