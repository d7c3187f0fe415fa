"""Grid, transverse basis and the Helmholtz operator, FP64 (oracle; test infrastructure).

Array conventions follow the reference's column-major layouts, expressed as NumPy
arrays with the same index order:

* ``epsinv``: shape ``(3, 3, Nx, Ny)``, ``epsinv[a, b, ix, iy]``
* field in the transverse basis: shape ``(2, Nx, Ny)``; the flat eigenvector of length
  ``2N`` is its Fortran-order ravel, i.e. flat index ``c + 2*(ix + Nx*iy)``
* Cartesian fields ``d``, ``e``: shape ``(3, Nx, Ny)``

The forward transform is NumPy's ``fft2`` (FFTW ``fft``: exponent ``-2*pi*i``), the
backward one the unnormalised inverse (FFTW ``bfft``) -- the same pair the reference
plans at ``lib/MaxwellEigenmodes/src/maxwell.jl:940-941``.
"""
import os

import numpy as np
import scipy.fft as _sfft

_WORKERS = int(os.environ.get("ORACLE_FFT_WORKERS", str(os.cpu_count() or 1)))
_SMALL = 3 * 96 * 96  # below this many elements one thread is faster than starting several


def _workers(a):
    return _WORKERS if np.size(a) >= _SMALL else 1

__all__ = [
    "Grid", "my_fftfreq", "gvec", "k_guess", "mag_mn", "kx_tc", "kx_ct", "zx_tc",
    "eid", "fft2", "bfft2", "ifft2", "apply_M", "apply_P", "HMkH", "HMH",
    "to_grid", "to_flat", "epsinv_ave", "E_field", "canonicalize_phase",
    "gaussian_wg_epsi", "HelmholtzOracle",
]


class Grid:
    """2D periodic cell ``Dx x Dy`` (um) on ``Nx x Ny`` pixels.

    Restates ``Grid``/``x``/``y``/``δV`` of
    ``lib/DielectricSmoothing/src/grid.jl:46-71`` and ``:108-127`` (pixel centres at
    ``-D/2 + (i-1)*D/N``).
    """

    def __init__(self, Dx, Dy, Nx, Ny):
        self.Dx, self.Dy, self.Nx, self.Ny = float(Dx), float(Dy), int(Nx), int(Ny)

    @property
    def N(self):
        return self.Nx * self.Ny

    @property
    def shape(self):
        return (self.Nx, self.Ny)

    def x(self):
        return -self.Dx / 2 + np.arange(self.Nx) * (self.Dx / self.Nx)

    def y(self):
        return -self.Dy / 2 + np.arange(self.Ny) * (self.Dy / self.Ny)

    def dV(self):
        return self.Dx * self.Dy / (self.Nx * self.Ny)

    def __repr__(self):
        return f"Grid({self.Dx}, {self.Dy}, {self.Nx}, {self.Ny})"


def my_fftfreq(n, fs):
    """``my_fftfreq`` (``lib/DielectricSmoothing/src/grid.jl:241-243``)."""
    if n % 2 == 0:
        idx = np.concatenate([np.arange(0, n // 2), np.arange(-(n // 2), 0)])
    else:
        idx = np.concatenate([np.arange(0, (n - 1) // 2 + 1), np.arange(-((n - 1) // 2), 0)])
    return idx * fs / n


def gvec(grid):
    """Reciprocal lattice in FFT order, cycles/um, no 2*pi
    (``lib/DielectricSmoothing/src/grid.jl:255-258``).  Returns ``(gx[Nx], gy[Ny])``."""
    return (my_fftfreq(grid.Nx, grid.Nx / grid.Dx), my_fftfreq(grid.Ny, grid.Ny / grid.Dy))


def k_guess(omega, epsinv):
    """``k_guess`` (``lib/MaxwellEigenmodes/src/MaxwellEigenmodes.jl:63``)."""
    return float(omega) * np.sqrt(1.0 / min(epsinv[a, a].min() for a in range(3)))


def mag_mn(k, grid):
    """Transverse basis at ``k_vec = (0, 0, k)``: ``w = k_vec - G``; ``mag = |w|``;
    ``n = normalize(z x w)`` (``y`` at the gimbal point ``w_x = w_y = 0``);
    ``m = normalize(n x w)``.  Restates ``mag_mn!``
    (``lib/MaxwellEigenmodes/src/maxwell.jl:646-660``) and ``mag_m_n``
    (``:491-505``).  Returns ``mag (Nx,Ny)``, ``m (3,Nx,Ny)``, ``n (3,Nx,Ny)``."""
    gx, gy = gvec(grid)
    wx = -gx[:, None] * np.ones((1, grid.Ny))
    wy = -gy[None, :] * np.ones((grid.Nx, 1))
    wz = np.full((grid.Nx, grid.Ny), float(k))
    mag = np.sqrt(wx * wx + wy * wy + wz * wz)
    # z x w = (-w_y, w_x, 0)
    nx_, ny_ = -wy, wx
    nn = np.sqrt(nx_ * nx_ + ny_ * ny_)
    gimbal = (wx * wx + wy * wy) <= 0.0
    nn_safe = np.where(gimbal, 1.0, nn)
    nx_ = np.where(gimbal, 0.0, nx_ / nn_safe)
    ny_ = np.where(gimbal, 1.0, ny_ / nn_safe)
    nz_ = np.zeros_like(nx_)
    # n x w
    mx_ = ny_ * wz - nz_ * wy
    my_ = nz_ * wx - nx_ * wz
    mz_ = nx_ * wy - ny_ * wx
    mm = np.sqrt(mx_ * mx_ + my_ * my_ + mz_ * mz_)
    m = np.stack([mx_ / mm, my_ / mm, mz_ / mm])
    n = np.stack([nx_, ny_, nz_])
    return mag, m, n


def to_grid(Hflat, grid, ncomp=2):
    """Flat column-major vector -> ``(ncomp, Nx, Ny)``."""
    return np.asarray(Hflat).reshape((ncomp, grid.Nx, grid.Ny), order="F")


def to_flat(Hg):
    """``(ncomp, Nx, Ny)`` -> flat column-major vector."""
    return np.asarray(Hg).reshape(-1, order="F")


def fft2(a):
    return _sfft.fft2(a, axes=(-2, -1), workers=_workers(a))


def bfft2(a):
    """Unnormalised inverse transform (FFTW ``bfft``)."""
    return _sfft.ifft2(a, axes=(-2, -1), norm="forward", workers=_workers(a))


def ifft2(a):
    return _sfft.ifft2(a, axes=(-2, -1), workers=_workers(a))


def kx_tc(H, m, n, mag):
    """``d_a = (H2 m_a - H1 n_a) mag`` (``maxwell.jl:385-394``, ``kx_tc`` ``:198-204``)."""
    return (H[1][None] * m - H[0][None] * n) * mag[None]


def kx_ct(e, m, n, mag):
    """``H1 = -(e.n) mag``, ``H2 = (e.m) mag`` without the ``1/N`` (``maxwell.jl:209-216``;
    the mutating form ``kx_ct!`` ``:404-411`` also multiplies by ``Ninv``)."""
    en = (e * n).sum(0)
    em = (e * m).sum(0)
    return np.stack([-en * mag, em * mag])


def zx_tc(H, m, n):
    """``(-(H1 m_y + H2 n_y), H1 m_x + H2 n_x, 0)`` (``maxwell.jl:220-227``)."""
    d0 = -(H[0] * m[1] + H[1] * n[1])
    d1 = H[0] * m[0] + H[1] * n[0]
    return np.stack([d0, d1, np.zeros_like(d0)])


def eid(epsinv, d):
    """``e_a = sum_b epsinv[b, a] d_b`` (``eid!``, ``maxwell.jl:413-419``; equal to the
    ``[a, b]`` form of ``ε⁻¹_dot`` for the symmetric tensors this path supports)."""
    return np.einsum("baxy,bxy->axy", epsinv, d)


def epsinv_ave(epsinv):
    """``ε_ave = 3 / tr(ε⁻¹)`` per pixel (``maxwell.jl:945``)."""
    return 3.0 / (epsinv[0, 0] + epsinv[1, 1] + epsinv[2, 2])


def _block_grid(Hflat, grid):
    """(2N, m) block -> (m, 2, Nx, Ny)."""
    m = Hflat.shape[1]
    return np.asarray(Hflat).T.reshape((m, grid.Ny, grid.Nx, 2)).transpose(0, 3, 2, 1)


def _block_flat(Hg):
    """(m, 2, Nx, Ny) -> (2N, m)."""
    m = Hg.shape[0]
    return np.ascontiguousarray(Hg.transpose(0, 3, 2, 1)).reshape(m, -1).T


def apply_M(Hflat, k, epsinv, grid, basis=None):
    """Helmholtz operator ``_M!`` (``maxwell.jl:471-481``): ``kx_tc -> fft -> eid ->
    bfft -> kx_ct * (1/N)``.  ``Hflat`` may be a vector (2N) or a matrix (2N, m)."""
    Hflat = np.asarray(Hflat)
    mag, m, n = basis if basis is not None else mag_mn(k, grid)
    if Hflat.ndim == 2:
        H = _block_grid(Hflat, grid)
        d = (H[:, 1, None] * m[None] - H[:, 0, None] * n[None]) * mag[None, None]
        d = fft2(d)
        e = bfft2(np.einsum("baxy,jbxy->jaxy", epsinv, d))
        en = (e * n[None]).sum(1)
        em = (e * m[None]).sum(1)
        out = np.stack([-en * mag[None], em * mag[None]], 1) / grid.N
        return _block_flat(out)
    H = to_grid(Hflat, grid)
    d = fft2(kx_tc(H, m, n, mag))
    e = bfft2(eid(epsinv, d))
    return to_flat(kx_ct(e, m, n, mag) / grid.N)


def apply_P(Hflat, k, epsinv, grid, basis=None):
    """Preconditioner ``_P!`` (``maxwell.jl:459-469``): ``e = (H1 n - H2 m)/mag``
    (``kxinv_tc!`` ``:430-437``) -> ``bfft`` -> ``d = ε_ave e`` (``ed_approx!``
    ``:448-455``) -> ``fft`` -> ``H1 = (d.n)/(mag N)``, ``H2 = -(d.m)/(mag N)``
    (``kxinv_ct!`` ``:439-446``)."""
    Hflat = np.asarray(Hflat)
    mag, m, n = basis if basis is not None else mag_mn(k, grid)
    inv_mag = 1.0 / mag
    if Hflat.ndim == 2:
        H = _block_grid(Hflat, grid)
        e = (H[:, 0, None] * n[None] - H[:, 1, None] * m[None]) * inv_mag[None, None]
        d = fft2(epsinv_ave(epsinv)[None, None] * bfft2(e))
        scale = inv_mag / grid.N
        out = np.stack([(d * n[None]).sum(1) * scale[None], -(d * m[None]).sum(1) * scale[None]], 1)
        return _block_flat(out)
    H = to_grid(Hflat, grid)
    e = (H[0][None] * n - H[1][None] * m) * inv_mag[None]
    e = bfft2(e)
    d = fft2(epsinv_ave(epsinv)[None] * e)
    scale = inv_mag / grid.N
    out = np.stack([(d * n).sum(0) * scale, -(d * m).sum(0) * scale])
    return to_flat(out)


def HMkH(Hflat, k, epsinv, grid, basis=None):
    """``-Re <H, kx_ct(ifft(ε⁻¹ . fft(zx_tc H)))>`` with normalised ``ifft``
    (``maxwell.jl:267-275``).  ``∂ω²/∂k = 2 HMkH`` for a unit eigenvector
    (``solve.jl:82``)."""
    mag, m, n = basis if basis is not None else mag_mn(k, grid)
    H = to_grid(Hflat, grid)
    d = fft2(zx_tc(H, m, n))
    e = ifft2(np.einsum("abxy,bxy->axy", epsinv, d))
    return -np.real(np.vdot(to_flat(H), to_flat(kx_ct(e, m, n, mag))))


def HMH(Hflat, k, A, grid, basis=None):
    """``Re <H, kx_ct(ifft(A . fft(kx_tc H)))>`` with normalised ``ifft``
    (``maxwell.jl:288-291``); ``= ω²`` for ``A = ε⁻¹`` and a unit eigenvector."""
    mag, m, n = basis if basis is not None else mag_mn(k, grid)
    H = to_grid(Hflat, grid)
    d = fft2(kx_tc(H, m, n, mag))
    e = ifft2(np.einsum("abxy,bxy->axy", A, d))
    return np.real(np.vdot(to_flat(H), to_flat(kx_ct(e, m, n, mag))))


def E_field(k, Hflat, epsinv, grid):
    """Unnormalised, un-canonicalised ``E = ε⁻¹ . fft(kx_tc(H))`` (``fields.jl:194-210``
    with ``canonicalize=false, normalized=false``).  Shape ``(3, Nx, Ny)``."""
    mag, m, n = mag_mn(k, grid)
    return np.einsum("abxy,bxy->axy", epsinv, fft2(kx_tc(to_grid(Hflat, grid), m, n, mag)))


def canonicalize_phase(Hflat, k, epsinv, grid):
    """``evec * cis(-angle(val_magmax(E)))`` (``fields.jl:258-260``); ``val_magmax``
    (``fields.jl:244``) picks the first maximum of ``|E|^2`` in column-major order over
    ``(3, Nx, Ny)``."""
    E = E_field(k, Hflat, epsinv, grid)
    flat = E.reshape(-1, order="F")
    idx = int(np.argmax(np.abs(flat) ** 2))
    return np.asarray(Hflat) * np.exp(-1j * np.angle(flat[idx]))


def gaussian_wg_epsi(p, grid):
    """Closed-form isotropic Gaussian index bump
    (``lib/MaxwellEigenmodes/test/runtests.jl:29-41``): ``ε = ε_bg + (ε_core - ε_bg)
    exp(-(x²/wx² + y²/wy²))``, returned as ``ε⁻¹`` of shape ``(3,3,Nx,Ny)``."""
    eps_core, eps_bg, wx, wy = p
    xs, ys = grid.x(), grid.y()
    eps = eps_bg + (eps_core - eps_bg) * np.exp(-(xs[:, None] ** 2 / wx ** 2 + ys[None, :] ** 2 / wy ** 2))
    out = np.zeros((3, 3, grid.Nx, grid.Ny))
    for a in range(3):
        out[a, a] = 1.0 / eps
    return out


class HelmholtzOracle:
    """Operator bundle at fixed ``(k, ε⁻¹, grid)`` with a cached basis
    (``HelmholtzMap``, ``maxwell.jl:922-949``)."""

    def __init__(self, k, epsinv, grid):
        self.k, self.epsinv, self.grid = float(k), epsinv, grid
        self.basis = mag_mn(k, grid)

    def M(self, X):
        return apply_M(X, self.k, self.epsinv, self.grid, self.basis)

    def P(self, X):
        return apply_P(X, self.k, self.epsinv, self.grid, self.basis)

    def HMkH(self, H):
        return HMkH(H, self.k, self.epsinv, self.grid, self.basis)
