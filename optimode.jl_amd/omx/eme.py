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
           "interface_smatrix", "propagation_smatrix", "star", "cascade", "EMEResult", "eme", "power_coupling"]


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
        continue_chain=False):
    """``eme(ε⁻¹s, ∂ε_∂ωs, lengths, ω, grid, solver; nev, conjugate, reg, reciprocity,
    passivity)`` (``eme.jl:90-104``) with ``warmstart`` (``eme.jl:58-76``): solve every
    cell, build its power-normalised modes on the GPU, couple neighbours through interface
    S-matrices and cascade ``prop₀ ⋆ iface₀₁ ⋆ prop₁ ⋆ …`` (``_assemble``, ``eme.jl:125-136``).
    ``epsinvs`` may be arrays or zero-argument callables producing them (e.g. GPU
    smoothing of the cell's cross-section).  ``continue_chain=True`` warm-starts the first
    cell from the solver's previous solve as well (a chain processed in pieces)."""
    n = len(epsinvs)
    if n < 1 or len(lengths) != n:
        raise ValueError("need one length per cell and at least one cell")
    kms, overlaps, stats = [], [], []
    acc = None
    prev_slot = None
    for i in range(n):
        ei = epsinvs[i]() if callable(epsinvs[i]) else epsinvs[i]
        solver.options["warm_start"] = 1.0 if (warmstart and (i > 0 or continue_chain)) else 0.0
        t0 = time.perf_counter()
        km, _, info = solve_k(omega, ei, grid, solver, nev=nev, k_tol=k_tol, eig_tol=eig_tol,
                              max_eigsolves=max_eigsolves, maxiter=maxiter, return_info=True, fetch_evecs=False)
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
