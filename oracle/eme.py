"""EME per-slice mode fields and MEOW inner products (oracle; test infrastructure).

Restates ``lib/EigenmodeExpansion/src/modes.jl:95-170`` (``mode_fields``, ``build_mode``,
``inner_product``, ``overlap_matrix``) for 2D grids.
"""
import numpy as np

from .maxwell import mag_mn, kx_tc, fft2, to_grid


def mode_fields(k, ev, epsinv, grid):
    """``mode_fields`` (``modes.jl:105-113``): ``E = i ε⁻¹ fft(kx_tc(H))``,
    ``H = fft(tc(H))``, each ``(3, Nx, Ny)`` complex."""
    mag, m, n = mag_mn(k, grid)
    Ht = to_grid(np.asarray(ev, dtype=np.complex128), grid)
    D = fft2(kx_tc(Ht, m, n, mag))
    E = 1j * np.einsum("abxy,bxy->axy", epsinv, D)
    H = fft2(Ht[0][None] * m + Ht[1][None] * n)
    return E, H


def inner_product(E1, H1, E2, H2, dA, conjugate=False):
    """``inner_product`` (``modes.jl:135-148``): ``½ Σ (E1x H2y − E1y H2x) δA``, with
    ``E1`` conjugated when ``conjugate``."""
    e1x, e1y = (np.conj(E1[0]), np.conj(E1[1])) if conjugate else (E1[0], E1[1])
    return 0.5 * np.sum(e1x * H2[1] - e1y * H2[0]) * dA


def build_mode(k, ev, epsinv, grid, conjugate=False):
    """``build_mode`` (``modes.jl:116-122``): fields scaled by ``1/sqrt(⟨m, m⟩)`` so that
    the self inner product is 1.  Returns ``(E, H, nrm)``."""
    E, H = mode_fields(k, ev, epsinv, grid)
    dA = grid.dV()
    nrm = np.sqrt(complex(inner_product(E, H, E, H, dA, conjugate)))
    return E / nrm, H / nrm, nrm


def overlap_matrix(modes1, modes2, dA, conjugate=False):
    """``overlap_matrix`` (``modes.jl:156-170``): ``M[i, j] = inner_product(m1_i, m2_j)``;
    ``modes`` are lists of ``(E, H)``."""
    return np.array([[inner_product(a[0], a[1], b[0], b[1], dA, conjugate) for b in modes2] for a in modes1])
