"""EME per-slice mode fields and MEOW inner products (oracle; test infrastructure).

Restates ``lib/EigenmodeExpansion/src/modes.jl:95-170`` (``mode_fields``, ``build_mode``,
``inner_product``, ``overlap_matrix``) for 2D grids.
"""
import numpy as np

from .maxwell import mag_mn, kx_tc, kx_ct, fft2, bfft2, to_grid, to_flat


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


# ---------------------------------------------------------------------------------------
# pull-backs (cotangent convention of the reference's Zygote rules: a real scalar L of a
# complex array Z changes by dL = Re Σ conj(Z̄) dZ)
# ---------------------------------------------------------------------------------------

def mode_fields_pullback(k, ev, epsinv, grid, Ebar, Hbar):
    """Pull-back of ``mode_fields`` (``modes.jl:105-113``): cotangents ``(Ē, H̄)`` of
    ``E = i ε⁻¹ fft(kx_tc(H))`` and ``H = fft(tc(H))`` to ``(k̄, ēv, ε̄⁻¹)``.  ``ēv`` is the
    adjoint of both linear maps applied to the cotangents; ``k̄`` comes from the basis'
    dependence on k (``d(m|k+G|)/dk = n × ẑ``, ``d|k+G|/dk = k/|k+G|``, n independent of k);
    ``ε̄⁻¹[a, b] = Re(conj(Ē_a) i D_b)`` pixel by pixel, the rule of the contraction
    ``ε⁻¹_dot`` itself (``maxwell.jl:108-110``), not symmetrised."""
    mag, m, n = mag_mn(k, grid)
    Ht = to_grid(np.asarray(ev, dtype=np.complex128), grid)
    D = fft2(kx_tc(Ht, m, n, mag))
    gE = bfft2(-1j * np.einsum("baxy,bxy->axy", epsinv, Ebar))   # (i ε⁻¹ F)^H Ē
    gH = bfft2(Hbar)                                               # F^H H̄
    hb = kx_ct(gE, m, n, mag) + np.stack([(m * gH).sum(0), (n * gH).sum(0)])
    nxz = np.stack([n[1], -n[0], np.zeros_like(n[0])])             # n × ẑ
    dkx = Ht[1][None] * nxz - Ht[0][None] * n * (k / mag)[None]    # d kx_tc / dk
    dtc = Ht[0][None] * (nxz - m * (k / mag)[None]) / mag[None]    # d tc / dk
    kbar = float(np.real(np.sum(np.conj(gE) * dkx) + np.sum(np.conj(gH) * dtc)))
    eibar = np.real(np.conj(Ebar)[:, None] * 1j * D[None, :])
    return kbar, to_flat(hb), eibar


def normalization_pullback(E, H, nrm, dA, Ebar, Hbar, conjugate=False):
    """Pull-back of ``build_mode``'s power normalisation (``modes.jl:116-122``):
    ``(E, H) = (E_r, H_r) / sqrt(s)``, ``s = inner_product(E_r, H_r, E_r, H_r)``, from the
    cotangents of the normalised fields to those of the raw fields.  ``E, H`` are the
    normalised fields and ``nrm = sqrt(s)``."""
    s = nrm * nrm
    c = 0.5 / s * (np.sum(np.conj(Ebar) * E) + np.sum(np.conj(Hbar) * H))
    Er_bar = Ebar / np.conj(nrm)
    Hr_bar = Hbar / np.conj(nrm)
    Er, Hr = E * nrm, H * nrm
    h = 0.5 * dA
    if conjugate:
        Er_bar[0] += -c * h * Hr[1]
        Er_bar[1] += c * h * Hr[0]
        Hr_bar[1] += -np.conj(c) * h * Er[0]
        Hr_bar[0] += np.conj(c) * h * Er[1]
    else:
        Er_bar[0] += -np.conj(c) * h * np.conj(Hr[1])
        Er_bar[1] += np.conj(c) * h * np.conj(Hr[0])
        Hr_bar[1] += -np.conj(c) * h * np.conj(Er[0])
        Hr_bar[0] += np.conj(c) * h * np.conj(Er[1])
    return Er_bar, Hr_bar


def overlap_pullback(modes1, modes2, dA, Obar, conjugate=False):
    """Pull-back of ``overlap_matrix`` (``modes.jl:156-170``): the cotangent ``Ō`` of
    ``O[i, j] = inner_product(m1_i, m2_j)`` to cotangents of the fields (lists of
    ``(Ē, H̄)`` for ``modes1`` and ``modes2``, zero in the components the product does
    not read)."""
    h = 0.5 * dA
    out1 = [(np.zeros_like(a[0]), np.zeros_like(a[1])) for a in modes1]
    out2 = [(np.zeros_like(b[0]), np.zeros_like(b[1])) for b in modes2]
    for i, a in enumerate(modes1):
        for j, b in enumerate(modes2):
            o = Obar[i, j]
            if conjugate:
                out1[i][0][0] += h * np.conj(o) * b[1][1]
                out1[i][0][1] += -h * np.conj(o) * b[1][0]
                out2[j][1][1] += h * o * a[0][0]
                out2[j][1][0] += -h * o * a[0][1]
            else:
                out1[i][0][0] += h * o * np.conj(b[1][1])
                out1[i][0][1] += -h * o * np.conj(b[1][0])
                out2[j][1][1] += h * o * np.conj(a[0][0])
                out2[j][1][0] += -h * o * np.conj(a[0][1])
    return out1, out2
