"""Eigenmode expansion over a chain of cross-sections on the GPU mode solver.

Mirrors ``lib/EigenmodeExpansion``: per cell, ``solve_k`` (warm-started from the previous
cell, ``eme.jl:58-76``), the power-normalised mode fields (``build_mode``,
``modes.jl:105-122``) and the MEOW overlaps between neighbouring cells
(``inner_product`` / ``overlap_matrix``, ``modes.jl:135-170``) run on the GPU, with the
fields kept on the device; the small ``nev x nev`` interface, propagation and
Redheffer-star algebra (``smatrix.jl:40-190``) and the cascade (``eme.jl:125-136``) run
on the host.
"""
import ctypes
import time

import numpy as np

from . import _lib, solve_k, OmxError

__all__ = ["build_modes", "overlap_matrix", "mode_fields", "SMat", "reg_solve", "enforce_passivity",
           "interface_smatrix", "propagation_smatrix", "star", "cascade", "EMEResult", "eme", "power_coupling",
           "assemble_pullback", "overlap_pullback", "mode_pullback", "power_coupling_objective", "eme_gradient"]


# ---------------------------------------------------------------------------------------
# GPU mode sets
# ---------------------------------------------------------------------------------------

def build_modes(kmags, evecs, epsinv, grid, solver, slot=0, conjugate=False, normalize=True):
    """``build_mode`` for every mode (``modes.jl:116-122``) into the device mode set
    ``slot`` (0..3) of the solver's handle.  ``evecs=None`` uses the eigenvectors of the
    handle's last ``solve_k`` (still on the device).  Returns the complex norms
    ``sqrt(inner_product(m, m))`` the fields were divided by."""
    h = solver.handle(grid, epsinv)
    km = np.ascontiguousarray(np.atleast_1d(kmags), dtype=np.float64)
    nev = len(km)
    E = None
    if evecs is not None:
        E = np.ascontiguousarray(np.stack([np.asarray(v, dtype=np.complex128) for v in evecs]))
        if E.shape != (nev, 2 * grid.N):
            raise OmxError(_lib.OMX_ERR_BAD_SHAPE, "one eigenvector of length 2N per k is needed")
    norms = np.empty(nev, dtype=np.complex128)
    _lib.check(_lib.lib().omx_build_modes(h.h, int(slot), nev, _lib.dptr(km), _lib.dptr(E), int(bool(conjugate)),
                                          int(bool(normalize)), _lib.dptr(norms)), h.h)
    return norms


def overlap_matrix(solver, slot_a, slot_b, na, nb, conjugate=False):
    """``overlap_matrix`` (``modes.jl:156-170``) of two device mode sets: ``(na, nb)``."""
    out = np.empty(na * nb, dtype=np.complex128)
    _lib.check(_lib.lib().omx_overlap_matrix(solver._h.h, int(slot_a), int(slot_b), int(bool(conjugate)),
                                             _lib.dptr(out)), solver._h.h)
    return out.reshape((na, nb), order="F")


def mode_fields(solver, slot, nev, grid):
    """The fields of a device mode set: lists of ``(3, Nx, Ny)`` complex ``E`` and ``H``."""
    n3 = 3 * grid.N
    E = np.empty(nev * n3, dtype=np.complex128)
    H = np.empty(nev * n3, dtype=np.complex128)
    _lib.check(_lib.lib().omx_get_mode_fields(solver._h.h, int(slot), _lib.dptr(E), _lib.dptr(H)), solver._h.h)
    f = lambda a, j: a[j * n3:(j + 1) * n3].reshape((3, grid.Nx, grid.Ny), order="F")  # noqa: E731
    return [f(E, j) for j in range(nev)], [f(H, j) for j in range(nev)]


# ---------------------------------------------------------------------------------------
# S-matrix algebra (host; nev x nev)
# ---------------------------------------------------------------------------------------

class SMat:
    """Two-sided multimode S-matrix ``[[R_LL, T_RL], [T_LR, R_RR]]`` (``smatrix.jl:15-35``)."""

    def __init__(self, S, nl, nr):
        self.S, self.nl, self.nr = np.asarray(S, dtype=np.complex128), int(nl), int(nr)

    def transmission(self):
        return self.S[self.nl:, :self.nl]

    def reflection(self):
        return self.S[:self.nl, :self.nl]


def reg_solve(A, B, reg=1e-9):
    """Tikhonov-regularised ``pinv(A) B`` (``smatrix.jl:47-52``)."""
    AhA = A.conj().T @ A
    n = AhA.shape[0]
    mu = reg * (np.real(np.trace(AhA)) / max(n, 1) + np.finfo(float).eps)
    return np.linalg.solve(AhA + mu * np.eye(n), A.conj().T @ B)


def enforce_passivity(s, method="invert"):
    """Singular-value passivity projection (``smatrix.jl:65-71``)."""
    s = np.asarray(s, dtype=float)
    if method == "none":
        return s
    if method == "clip":
        return np.minimum(s, 1.0)
    if method == "invert":
        return np.where(s > 1, 1.0 / np.maximum(s, 1e-300), s)
    if method == "subtract":
        return np.maximum(np.where(s > 1, 2 - s, s), 0.0)
    raise ValueError(f"unknown passivity method {method!r}")


def interface_smatrix(O_LR, O_RL, conjugate=False, reg=1e-9, reciprocity=True, passivity="invert"):
    """Mode-matching interface S-matrix (``smatrix.jl:96-126``) from ``O_LR`` =
    overlap(left, right) and ``O_RL`` = overlap(right, left)."""
    NL, NR = O_LR.shape
    ORL_t = O_RL.conj().T if conjugate else O_RL.T
    OLR_t = O_LR.conj().T if conjugate else O_LR.T
    IL, IR = np.eye(NL, dtype=complex), np.eye(NR, dtype=complex)
    T_LR = reg_solve(O_LR + ORL_t, 2.0 * IL, reg)
    T_RL = reg_solve(O_RL + OLR_t, 2.0 * IR, reg)
    R_LL = 0.5 * ((ORL_t @ T_LR - IL) + (IL - O_LR @ T_LR))
    R_RR = 0.5 * ((OLR_t @ T_RL - IR) + (IR - O_RL @ T_RL))
    S = np.block([[R_LL, T_RL], [T_LR, R_RR]])
    if passivity != "none":
        U, sv, Vh = np.linalg.svd(S)
        S = (U * enforce_passivity(sv, passivity)[None, :]) @ Vh
    if reciprocity:
        S = 0.5 * (S + S.T)
    return SMat(S, NL, NR)


def propagation_smatrix(kmags, length):
    """Diagonal propagation S-matrix, phases ``exp(2πi k L)`` (``smatrix.jl:137-144``)."""
    ph = np.exp(2j * np.pi * np.asarray(kmags, dtype=float) * float(length))
    n = len(ph)
    Z = np.zeros((n, n), dtype=complex)
    return SMat(np.block([[Z, np.diag(ph)], [np.diag(ph), Z]]), n, n)


def star(A, B):
    """Redheffer star product joining A's right ports to B's left ports (``smatrix.jl:156-172``)."""
    if A.nr != B.nl:
        raise ValueError(f"port count mismatch in cascade: {A.nr} vs {B.nl}")
    a, b = A.S, B.S
    A11, A12, A21, A22 = a[:A.nl, :A.nl], a[:A.nl, A.nl:], a[A.nl:, :A.nl], a[A.nl:, A.nl:]
    B11, B12, B21, B22 = b[:B.nl, :B.nl], b[:B.nl, B.nl:], b[B.nl:, :B.nl], b[B.nl:, B.nl:]
    Ic = np.eye(A.nr, dtype=complex)
    L1 = Ic - B11 @ A22
    L2 = Ic - A22 @ B11
    C11 = A11 + A12 @ np.linalg.solve(L1, B11 @ A21)
    C12 = A12 @ np.linalg.solve(L1, B12)
    C21 = B21 @ np.linalg.solve(L2, A21)
    C22 = B22 + B21 @ np.linalg.solve(L2, A22 @ B12)
    return SMat(np.block([[C11, C12], [C21, C22]]), A.nl, B.nr)


def cascade(smats):
    """Left-to-right star products of a list of S-matrices (``smatrix.jl:183-190``)."""
    if not smats:
        raise ValueError("cannot cascade an empty stack")
    acc = smats[0]
    for s in smats[1:]:
        acc = star(acc, s)
    return acc


# ---------------------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------------------

class EMEResult:
    """``EMEResult`` (``eme.jl:20-25``) with the per-cell propagation constants in place of
    the field-carrying ``Mode`` objects (the fields of the last two cells stay on the
    device), the interface overlaps, and per-cell solve statistics."""

    def __init__(self, S, kmags, lengths, omega, overlaps, stats):
        self.S, self.kmags, self.lengths, self.omega = S, kmags, lengths, omega
        self.overlaps, self.stats = overlaps, stats


def eme(epsinvs, lengths, omega, grid, solver, nev=2, conjugate=False, reg=1e-9, reciprocity=True,
        passivity="invert", warmstart=True, k_tol=1e-8, eig_tol=1e-8, max_eigsolves=60, maxiter=100,
        continue_chain=False, learned=None):
    """``eme(ε⁻¹s, ∂ε_∂ωs, lengths, ω, grid, solver; nev, conjugate, reg, reciprocity,
    passivity)`` (``eme.jl:90-104``) with ``warmstart`` (``eme.jl:58-76``): solve every
    cell, build its power-normalised modes on the GPU, couple neighbours through interface
    S-matrices and cascade ``prop₀ ⋆ iface₀₁ ⋆ prop₁ ⋆ …`` (``_assemble``, ``eme.jl:125-136``).
    ``epsinvs`` may be arrays or zero-argument callables producing them (e.g. GPU
    smoothing of the cell's cross-section).  ``continue_chain=True`` warm-starts the first
    cell from the solver's previous solve as well (a chain processed in pieces).
    ``learned`` (a ``warmstart.LearnedStart`` predicting at least ``nev`` modes on this grid,
    e.g. ``warmstart.load_coupler()``) starts the cold cells -- the first one, every cell when
    ``warmstart`` is off, and a cell whose warm-started solve did not converge, which is solved
    again cold -- from the network's guess instead of ``k_guess`` and random vectors; the
    converged modes do not depend on the start."""
    n = len(epsinvs)
    if n < 1 or len(lengths) != n:
        raise ValueError("need one length per cell and at least one cell")
    if learned is not None and learned.nev < nev:
        raise ValueError(f"the learned start predicts {learned.nev} modes, {nev} requested")

    def cold(ei):
        solver.options["warm_start"] = 0.0
        if learned is None:
            return solve_k(omega, ei, grid, solver, nev=nev, k_tol=k_tol, eig_tol=eig_tol,
                           max_eigsolves=max_eigsolves, maxiter=maxiter, return_info=True, fetch_evecs=False)
        from .warmstart import solve_k_learned
        return solve_k_learned(omega, ei, grid, solver, learned, nev=nev, k_tol=k_tol, eig_tol=eig_tol,
                               max_eigsolves=max_eigsolves, maxiter=maxiter, return_info=True, fetch_evecs=False)

    kms, overlaps, stats = [], [], []
    acc = None
    prev_slot = None
    for i in range(n):
        ei = epsinvs[i]() if callable(epsinvs[i]) else epsinvs[i]
        warm = warmstart and (i > 0 or continue_chain)
        t0 = time.perf_counter()
        if warm:
            solver.options["warm_start"] = 1.0
            try:
                km, _, info = solve_k(omega, ei, grid, solver, nev=nev, k_tol=k_tol, eig_tol=eig_tol,
                                      max_eigsolves=max_eigsolves, maxiter=maxiter, return_info=True,
                                      fetch_evecs=False)
            except OmxError as e:
                if e.status != _lib.OMX_ERR_NOT_CONVERGED:
                    raise
                info = None
            if info is None or not info["converged"]:
                km, _, info = cold(ei)   # the previous cell's modes were no start for this one
                info["cold_restart"] = True
        else:
            km, _, info = cold(ei)
        slot = i % 2
        build_modes(km, None, ei, grid, solver, slot=slot, conjugate=conjugate)
        info["seconds"] = time.perf_counter() - t0
        kms.append(km)
        stats.append(info)
        prop = propagation_smatrix(km, lengths[i])
        if prev_slot is None:
            acc = prop
        else:
            O_LR = overlap_matrix(solver, prev_slot, slot, len(kms[-2]), nev, conjugate)
            O_RL = overlap_matrix(solver, slot, prev_slot, nev, len(kms[-2]), conjugate)
            overlaps.append((O_LR, O_RL))
            iface = interface_smatrix(O_LR, O_RL, conjugate, reg, reciprocity, passivity)
            acc = star(star(acc, iface), prop)
        prev_slot = slot
    solver.options["warm_start"] = 0.0
    return EMEResult(acc, kms, list(lengths), omega, overlaps, stats)


def power_coupling(result, in_mode=1, out_mode=1):
    """``|S₂₁[out, in]|²`` (``eme.jl:140-149``), 1-based mode numbers as in the reference."""
    S = result.S if isinstance(result, EMEResult) else result
    return float(np.abs(S.transmission()[out_mode - 1, in_mode - 1]) ** 2)


# ---------------------------------------------------------------------------------------
# reverse mode of the S-matrix algebra (host)
# ---------------------------------------------------------------------------------------
# The reference differentiates the cascade with Zygote (eme.jl:80-88: every cell's solve_k
# rrule, then plain differentiable linear algebra).  Here the nev x nev algebra records its
# operations on a small tape and replays them backwards, with Zygote's convention for
# complex cotangents: a real scalar L of a complex array Z changes by Re Σ conj(Z̄) dZ.

class _Tape:
    """Reverse-mode recorder for complex matrix expressions (products, solves, sums,
    transposes, blocks, elementwise phases)."""

    def __init__(self):
        self.ops = []      # (output id, backward fn)
        self.vals = []
        self.grads = {}

    def var(self, value):
        self.vals.append(np.asarray(value, dtype=np.complex128))
        return len(self.vals) - 1

    def _op(self, value, back):
        i = self.var(value)
        self.ops.append((i, back))
        return i

    def v(self, i):
        return self.vals[i]

    def acc(self, i, g):
        self.grads[i] = self.grads[i] + g if i in self.grads else np.array(g, dtype=np.complex128)

    # operations
    def matmul(self, a, b):
        A, B = self.vals[a], self.vals[b]
        def back(g):
            self.acc(a, g @ B.conj().T)
            self.acc(b, A.conj().T @ g)
        return self._op(A @ B, back)

    def solve(self, a, b):
        A, B = self.vals[a], self.vals[b]
        X = np.linalg.solve(A, B)
        def back(g):
            gb = np.linalg.solve(A.conj().T, g)
            self.acc(b, gb)
            self.acc(a, -gb @ X.conj().T)
        return self._op(X, back)

    def add(self, a, b, sb=1.0):
        def back(g):
            self.acc(a, g)
            self.acc(b, np.conj(sb) * g)
        return self._op(self.vals[a] + sb * self.vals[b], back)

    def scale(self, a, s):
        return self._op(s * self.vals[a], lambda g: self.acc(a, np.conj(s) * g))

    def const_add(self, a, C):
        return self._op(self.vals[a] + C, lambda g: self.acc(a, g))

    def T(self, a):
        return self._op(self.vals[a].T, lambda g: self.acc(a, g.T))

    def H(self, a):
        return self._op(self.vals[a].conj().T, lambda g: self.acc(a, g.conj().T))

    def block(self, parts):
        """2 x 2 block matrix of ids [[a, b], [c, d]]."""
        rows = [[self.vals[p] for p in r] for r in parts]
        out = np.block(rows)
        r0, c0 = rows[0][0].shape

        def back(g):
            self.acc(parts[0][0], g[:r0, :c0])
            self.acc(parts[0][1], g[:r0, c0:])
            self.acc(parts[1][0], g[r0:, :c0])
            self.acc(parts[1][1], g[r0:, c0:])
        return self._op(out, back)

    def sub(self, a, r0, r1, c0, c1):
        A = self.vals[a]

        def back(g):
            z = np.zeros_like(A)
            z[r0:r1, c0:c1] = g
            self.acc(a, z)
        return self._op(A[r0:r1, c0:c1], back)

    def reg_solve(self, a, b, reg):
        """``reg_solve`` (``smatrix.jl:47-52``) on the tape: (AᴴA + μI)⁻¹ AᴴB with
        μ = reg (Re tr(AᴴA)/n + eps)."""
        A = self.vals[a]
        n = A.shape[1]
        Ah = self.H(a)
        AhA = self.matmul(Ah, a)
        mu = reg * (np.real(np.trace(self.vals[AhA])) / max(n, 1) + np.finfo(float).eps)
        # μ I as a function of A: dμ = (reg/n) 2 Re tr(Aᴴ dA)
        G = self.const_add(AhA, mu * np.eye(n))
        gi = len(self.vals) - 1

        def back_mu(g):
            self.acc(a, (2.0 * reg / max(n, 1)) * np.real(np.trace(g)) * A)
        self.ops.append((gi, back_mu))
        return self.solve(G, self.matmul(Ah, b))

    def backward(self, out, seed):
        self.grads = {out: np.asarray(seed, dtype=np.complex128)}
        for i, back in reversed(self.ops):
            if i in self.grads:
                back(self.grads[i])
        return self.grads

    def grad(self, i):
        return self.grads.get(i, np.zeros_like(self.vals[i]))


def _t_interface(tp, oLR, oRL, conjugate, reg, reciprocity, passivity):
    """``interface_smatrix`` on the tape; the passivity projection is differentiated only
    where it leaves the matrix unchanged (every singular value <= 1)."""
    NL, NR = tp.v(oLR).shape
    oRLt = tp.H(oRL) if conjugate else tp.T(oRL)
    oLRt = tp.H(oLR) if conjugate else tp.T(oLR)
    IL, IR = np.eye(NL, dtype=complex), np.eye(NR, dtype=complex)
    TLR = tp.reg_solve(tp.add(oLR, oRLt), tp.var(2.0 * IL), reg)
    TRL = tp.reg_solve(tp.add(oRL, oLRt), tp.var(2.0 * IR), reg)
    # R_LL = ½ (O_RLᵀ T_LR − O_LR T_LR), R_RR likewise (the identities cancel)
    RLL = tp.scale(tp.add(tp.matmul(oRLt, TLR), tp.matmul(oLR, TLR), -1.0), 0.5)
    RRR = tp.scale(tp.add(tp.matmul(oLRt, TRL), tp.matmul(oRL, TRL), -1.0), 0.5)
    S = tp.block([[RLL, TRL], [TLR, RRR]])
    if passivity != "none":
        sv = np.linalg.svd(tp.v(S), compute_uv=False)
        if sv.max() > 1.0:
            raise NotImplementedError("pull-back through an active passivity projection (singular value "
                                      f"{sv.max():.6g} > 1); use passivity='none'")
    if reciprocity:
        S = tp.scale(tp.add(S, tp.T(S)), 0.5)
    return S


def _t_propagation(tp, k, length):
    """``propagation_smatrix`` on the tape from a real k vector (its id)."""
    kv = np.real(tp.v(k))
    ph = np.exp(2j * np.pi * kv * float(length))
    n = len(kv)
    Z = np.zeros((n, n), dtype=complex)
    out = np.block([[Z, np.diag(ph)], [np.diag(ph), Z]])

    def back(g):
        gd = np.diag(g[:n, n:]) + np.diag(g[n:, :n])
        # dP = 2πi L ph dk: dL = Re Σ conj(ḡ) 2πi L ph dk
        tp.acc(k, np.real(np.conj(gd) * 2j * np.pi * float(length) * ph).astype(complex))
    return tp._op(out, back)


def _t_star(tp, A, nAl, B, nBl):
    """Redheffer ``star`` (``smatrix.jl:156-172``) on the tape; ids of the two S matrices
    with their left port counts."""
    a, b = tp.v(A), tp.v(B)
    nAr, nBr = a.shape[0] - nAl, b.shape[0] - nBl
    A11, A12 = tp.sub(A, 0, nAl, 0, nAl), tp.sub(A, 0, nAl, nAl, nAl + nAr)
    A21, A22 = tp.sub(A, nAl, nAl + nAr, 0, nAl), tp.sub(A, nAl, nAl + nAr, nAl, nAl + nAr)
    B11, B12 = tp.sub(B, 0, nBl, 0, nBl), tp.sub(B, 0, nBl, nBl, nBl + nBr)
    B21, B22 = tp.sub(B, nBl, nBl + nBr, 0, nBl), tp.sub(B, nBl, nBl + nBr, nBl, nBl + nBr)
    I = tp.var(np.eye(nAr, dtype=complex))
    L1 = tp.add(I, tp.matmul(B11, A22), -1.0)
    L2 = tp.add(I, tp.matmul(A22, B11), -1.0)
    C11 = tp.add(A11, tp.matmul(A12, tp.solve(L1, tp.matmul(B11, A21))))
    C12 = tp.matmul(A12, tp.solve(L1, B12))
    C21 = tp.matmul(B21, tp.solve(L2, A21))
    C22 = tp.add(B22, tp.matmul(B21, tp.solve(L2, tp.matmul(A22, B12))))
    return tp.block([[C11, C12], [C21, C22]])


def assemble_pullback(overlaps, kmags, lengths, conjugate=False, reg=1e-9, reciprocity=True, passivity="invert"):
    """``_assemble`` (``eme.jl:125-136``) recorded for its pull-back: from the interface
    overlaps ``[(O_LR, O_RL), ...]`` and the per-cell propagation constants, the device S
    matrix and a function mapping a cotangent ``Sbar`` of ``S.S`` to the cotangents of every
    overlap matrix and of every cell's k: ``back(Sbar) -> ([(Ō_LR, Ō_RL), ...], [k̄, ...])``."""
    tp = _Tape()
    kid = [tp.var(np.asarray(k, dtype=float)) for k in kmags]
    oid = [(tp.var(a), tp.var(b)) for a, b in overlaps]
    acc = _t_propagation(tp, kid[0], lengths[0])
    nl = len(kmags[0])
    for i in range(len(kmags) - 1):
        iface = _t_interface(tp, oid[i][0], oid[i][1], conjugate, reg, reciprocity, passivity)
        prop = _t_propagation(tp, kid[i + 1], lengths[i + 1])
        acc = _t_star(tp, acc, nl, iface, len(kmags[i]))
        acc = _t_star(tp, acc, nl, prop, len(kmags[i + 1]))

    def back(Sbar):
        tp.backward(acc, Sbar)
        return [(tp.grad(a), tp.grad(b)) for a, b in oid], [np.real(tp.grad(k)) for k in kid]
    return SMat(tp.v(acc), nl, len(kmags[-1])), back


def overlap_pullback(solver, slot_a, slot_b, obar, conjugate=False, which=3):
    """``omx_overlap_pullback``: add the field cotangents of ``O = overlap_matrix(set a,
    set b)`` for the cotangent ``obar`` to the device accumulators of set a's E (which & 1)
    and set b's H (which & 2)."""
    ob = np.asfortranarray(np.asarray(obar, dtype=np.complex128))
    _lib.check(_lib.lib().omx_overlap_pullback(solver._h.h, int(slot_a), int(slot_b), int(bool(conjugate)),
                                               _lib.dptr(ob.reshape(-1, order="F").copy()), int(which)), solver._h.h)


def mode_pullback(solver, slot, nev, grid):
    """``omx_mode_pullback``: ``(k̄[nev], [ēv_j], ε̄⁻¹ (3,3,Nx,Ny))`` of a mode set from its
    accumulated field cotangents (the direct ε⁻¹ dependence of E, not symmetrised)."""
    kb = np.empty(nev)
    eb = np.empty((nev, 2 * grid.N), dtype=np.complex128)
    ec = np.empty((grid.Ny, grid.Nx, 3, 3))
    _lib.check(_lib.lib().omx_mode_pullback(solver._h.h, int(slot), _lib.dptr(kb), _lib.dptr(eb), _lib.dptr(ec)),
               solver._h.h)
    return kb, [eb[j].copy() for j in range(nev)], np.ascontiguousarray(ec.transpose(3, 2, 1, 0))


def power_coupling_objective(in_mode=1, out_mode=1):
    """``power_coupling`` (``eme.jl:140-149``) as an objective for ``eme_gradient``:
    ``S -> (|T[out, in]|², S̄)``."""
    def f(S):
        i, o = in_mode - 1, out_mode - 1
        t = S.S[S.nl + o, i]
        Sbar = np.zeros_like(S.S)
        Sbar[S.nl + o, i] = 2.0 * t
        return float(abs(t) ** 2), Sbar
    return f


def eme_gradient(epsinvs, lengths, omega, grid, solver, objective, nev=2, conjugate=False, reg=1e-9,
                 reciprocity=True, passivity="invert", warmstart=True, k_tol=1e-8, eig_tol=1e-8, max_eigsolves=60,
                 maxiter=100, return_parts=False):
    """Value and gradient of a scalar of the EME device S matrix with respect to every cell's
    ε⁻¹ and ω: what ``Zygote.gradient`` gives through ``eme(ε⁻¹s, ∂ε_∂ωs, lengths, ω, grid,
    solver)`` (``eme.jl:80-104``).  ``objective(S) -> (value, S̄)``.

    Forward: the cell chain of ``eme`` (solve, mode sets, overlaps), keeping each cell's k and
    eigenvectors.  Backward: the host S-matrix algebra's reverse pass gives the cotangents of
    the overlap matrices and of the propagation constants; per cell, the GPU rebuilds its mode
    set and its neighbours', pulls the overlap cotangents back to field cotangents
    (``omx_overlap_pullback``) and those to (k̄, ēv) and the direct ε̄⁻¹ of E
    (``omx_mode_pullback``), then runs the ``solve_k`` pull-back (``omx_solve_k_adjoint``).
    Returns ``(value, [ε̄⁻¹ per cell], ω̄)``; each cell's ε̄⁻¹ is the sum of the ``solve_k``
    rule's (mirrored off-diagonals) and the field rule's (each entry its own derivative), as
    the two rules add up in the reference; ``return_parts=True`` appends the two parts per
    cell, ``[(solve_k part, field part), ...]``."""
    from . import _adjoint
    n = len(epsinvs)
    eis = [e() if callable(e) else np.asarray(e, dtype=np.float64) for e in epsinvs]
    kms, evs, overlaps = [], [], []
    for i in range(n):
        solver.options["warm_start"] = 1.0 if (warmstart and i > 0) else 0.0
        km, ev = solve_k(omega, eis[i], grid, solver, nev=nev, k_tol=k_tol, eig_tol=eig_tol,
                         max_eigsolves=max_eigsolves, maxiter=maxiter)
        slot = i % 2
        build_modes(km, ev, eis[i], grid, solver, slot=slot, conjugate=conjugate)
        if i > 0:
            overlaps.append((overlap_matrix(solver, 1 - slot, slot, nev, nev, conjugate),
                             overlap_matrix(solver, slot, 1 - slot, nev, nev, conjugate)))
        kms.append(km)
        evs.append(ev)
    solver.options["warm_start"] = 0.0
    S, back = assemble_pullback(overlaps, kms, lengths, conjugate, reg, reciprocity, passivity)
    value, Sbar = objective(S)
    obar, kbar_prop = back(Sbar)
    grads, parts, wbar = [], [], 0.0
    for s in range(n):
        # neighbours first (their fields use their own ε⁻¹), the cell itself last
        if s > 0:
            build_modes(kms[s - 1], evs[s - 1], eis[s - 1], grid, solver, slot=1, conjugate=conjugate)
        if s < n - 1:
            build_modes(kms[s + 1], evs[s + 1], eis[s + 1], grid, solver, slot=2, conjugate=conjugate)
        build_modes(kms[s], evs[s], eis[s], grid, solver, slot=0, conjugate=conjugate)
        if s > 0:
            oLR, oRL = obar[s - 1]          # O_LR = overlap(cell s-1, cell s), O_RL = overlap(s, s-1)
            overlap_pullback(solver, 1, 0, oLR, conjugate, which=2)
            overlap_pullback(solver, 0, 1, oRL, conjugate, which=1)
        if s < n - 1:
            oLR, oRL = obar[s]              # O_LR = overlap(s, s+1), O_RL = overlap(s+1, s)
            overlap_pullback(solver, 0, 2, oLR, conjugate, which=1)
            overlap_pullback(solver, 2, 0, oRL, conjugate, which=2)
        kb, evb, direct = mode_pullback(solver, 0, nev, grid)
        h = solver.handle(grid, eis[s])
        wb, eib, _ = _adjoint(h, grid, omega, nev, kms[s], evs[s], (list(kb + kbar_prop[s]), evb), period=False)
        grads.append(eib + direct)
        parts.append((eib, direct))
        wbar += wb
    if return_parts:
        return value, grads, wbar, parts
    return value, grads, wbar
