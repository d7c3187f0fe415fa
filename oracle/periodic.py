"""3D cells periodic along z: grid, transverse basis, operator, the period-weighted
``HMₖH`` and the period pull-back of ``solve_k`` (oracle; test infrastructure).

On a ``Grid{3}`` the reference runs the same functions as in 2D with the reciprocal
lattice extended by ``g_z = fftfreq(Nz, Nz/Δz)``, so that the shifted wavevector of plane
wave ``j`` is ``w_j = k_vec - g_j = (-g_x, -g_y, k - g_z)`` (``mag_m_n``,
``lib/MaxwellEigenmodes/src/maxwell.jl:491-505``; ``g⃗``,
``lib/DielectricSmoothing/src/grid.jl:255-258``) and the transforms run over the three
spatial axes.  The period ``Λ = Δz`` enters only through ``g_z = m/Λ``
(``lib/MaxwellEigenmodes/src/grads/period.jl:1-27``).

Layouts follow the reference's column-major arrays: ``epsinv[a, b, ix, iy, iz]`` of shape
``(3, 3, Nx, Ny, Nz)``; the flat eigenvector index is ``c + 2*(ix + Nx*(iy + Ny*iz))``.
"""
import numpy as np
import scipy.fft as _sfft

from .maxwell import _workers, my_fftfreq

__all__ = [
    "Grid3", "gvec3", "gz_field", "period_weight", "grid_with_period", "mag_mn3", "to_grid3", "to_flat3",
    "apply_M3", "apply_P3", "HMkH3", "HMkH_weighted", "dw2_dL", "E_field3", "canonicalize_phase3",
    "HelmholtzOracle3", "solve_k_periodic", "solve_k_periodic_kbar_pullback", "bragg_wg_epsi",
    "epsinv_bar3", "wbar_z", "solve_k_periodic_pullback",
]


class Grid3:
    """3D cell ``Dx x Dy x Dz`` (um) on ``Nx x Ny x Nz`` pixels (``Grid{3}``,
    ``lib/DielectricSmoothing/src/grid.jl:46-71``); ``Dz`` is the period along z."""

    def __init__(self, Dx, Dy, Dz, Nx, Ny, Nz):
        self.Dx, self.Dy, self.Dz = float(Dx), float(Dy), float(Dz)
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)

    @property
    def N(self):
        return self.Nx * self.Ny * self.Nz

    @property
    def shape(self):
        return (self.Nx, self.Ny, self.Nz)

    def x(self):
        return -self.Dx / 2 + np.arange(self.Nx) * (self.Dx / self.Nx)

    def y(self):
        return -self.Dy / 2 + np.arange(self.Ny) * (self.Dy / self.Ny)

    def z(self):
        return -self.Dz / 2 + np.arange(self.Nz) * (self.Dz / self.Nz)

    def __repr__(self):
        return f"Grid3({self.Dx}, {self.Dy}, {self.Dz}, {self.Nx}, {self.Ny}, {self.Nz})"


def gvec3(grid):
    """Reciprocal lattice in FFT order (cycles/um): ``(gx[Nx], gy[Ny], gz[Nz])``."""
    return (my_fftfreq(grid.Nx, grid.Nx / grid.Dx), my_fftfreq(grid.Ny, grid.Ny / grid.Dy),
            my_fftfreq(grid.Nz, grid.Nz / grid.Dz))


def gz_field(grid):
    """``gz_field`` (``grads/period.jl:48-55``): ``g_z`` broadcast to ``(Nx, Ny, Nz)``."""
    gz = my_fftfreq(grid.Nz, grid.Nz / grid.Dz)
    return np.broadcast_to(gz[None, None, :], grid.shape).copy()


def period_weight(grid):
    """``period_weight`` (``grads/period.jl:38-41``): ``g_z / Λ`` with ``Λ = Δz``."""
    return gz_field(grid) / grid.Dz


def grid_with_period(grid, L):
    """``_grid_with_period`` (``grads/period.jl:123-125``): the same pixels, period ``L``."""
    return Grid3(grid.Dx, grid.Dy, L, grid.Nx, grid.Ny, grid.Nz)


def mag_mn3(k, grid):
    """Basis at ``k_vec = (0, 0, k)``: ``w = k_vec - G``; ``mag = |w|``;
    ``n = normalize(z x w)`` (``y`` where ``w_x = w_y = 0``); ``m = normalize(n x w)``
    (``maxwell.jl:491-505``).  Shapes ``(Nx,Ny,Nz)``, ``(3,Nx,Ny,Nz)``, ``(3,Nx,Ny,Nz)``."""
    gx, gy, gz = gvec3(grid)
    sh = grid.shape
    wx = np.broadcast_to(-gx[:, None, None], sh)
    wy = np.broadcast_to(-gy[None, :, None], sh)
    wz = np.broadcast_to(float(k) - gz[None, None, :], sh)
    mag = np.sqrt(wx * wx + wy * wy + wz * wz)
    gimbal = (wx * wx + wy * wy) <= 0.0
    rho = np.where(gimbal, 1.0, np.sqrt(wx * wx + wy * wy))
    nx_ = np.where(gimbal, 0.0, -wy / rho)
    ny_ = np.where(gimbal, 1.0, wx / rho)
    nz_ = np.zeros(sh)
    mx_ = ny_ * wz - nz_ * wy
    my_ = nz_ * wx - nx_ * wz
    mz_ = nx_ * wy - ny_ * wx
    mm = np.sqrt(mx_ * mx_ + my_ * my_ + mz_ * mz_)
    return mag, np.stack([mx_ / mm, my_ / mm, mz_ / mm]), np.stack([nx_, ny_, nz_])


def to_grid3(Hflat, grid, ncomp=2):
    return np.asarray(Hflat).reshape((ncomp,) + grid.shape, order="F")


def to_flat3(Hg):
    return np.asarray(Hg).reshape(-1, order="F")


def _fft3(a):
    return _sfft.fftn(a, axes=(-3, -2, -1), workers=_workers(a))


def _bfft3(a):
    return _sfft.ifftn(a, axes=(-3, -2, -1), norm="forward", workers=_workers(a))


def _ifft3(a):
    return _sfft.ifftn(a, axes=(-3, -2, -1), workers=_workers(a))


def _cols(Hflat):
    H = np.asarray(Hflat)
    return (H[:, None], True) if H.ndim == 1 else (H, False)


def apply_M3(Hflat, k, epsinv, grid, basis=None):
    """``_M!`` on a 3D cell (``maxwell.jl:471-481``): ``kx_tc -> fft -> eid -> bfft ->
    kx_ct / N`` with 3D transforms.  Vector (2N) or block (2N, m)."""
    mag, m, n = basis if basis is not None else mag_mn3(k, grid)
    H, single = _cols(Hflat)
    out = np.empty_like(H, dtype=np.complex128)
    for j in range(H.shape[1]):
        h = to_grid3(H[:, j], grid)
        d = _fft3((h[1][None] * m - h[0][None] * n) * mag[None])
        e = _bfft3(np.einsum("baxyz,bxyz->axyz", epsinv, d))
        en, em = (e * n).sum(0), (e * m).sum(0)
        out[:, j] = to_flat3(np.stack([-en * mag, em * mag]) / grid.N)
    return out[:, 0] if single else out


def apply_P3(Hflat, k, epsinv, grid, basis=None):
    """``_P!`` on a 3D cell (``maxwell.jl:459-469``): ``(H1 n - H2 m)/mag -> bfft ->
    ε_ave -> fft -> ((d.n), -(d.m)) / (mag N)``."""
    mag, m, n = basis if basis is not None else mag_mn3(k, grid)
    eave = 3.0 / (epsinv[0, 0] + epsinv[1, 1] + epsinv[2, 2])
    H, single = _cols(Hflat)
    out = np.empty_like(H, dtype=np.complex128)
    for j in range(H.shape[1]):
        h = to_grid3(H[:, j], grid)
        e = _bfft3((h[0][None] * n - h[1][None] * m) / mag[None])
        d = _fft3(eave[None] * e)
        s = 1.0 / (mag * grid.N)
        out[:, j] = to_flat3(np.stack([(d * n).sum(0) * s, -(d * m).sum(0) * s]))
    return out[:, 0] if single else out


def HMkH_weighted(Hflat, k, epsinv, grid, wt, basis=None):
    """``HMₖH_weighted`` (``grads/period.jl:69-74``): ``-Re <H, kx_ct(ifft(ε⁻¹ .
    fft(wt . zx_tc H)))>`` with the normalised ``ifft``; ``wt`` is a real field on the
    plane waves, ``(Nx, Ny, Nz)``."""
    mag, m, n = basis if basis is not None else mag_mn3(k, grid)
    h = to_grid3(Hflat, grid)
    zx = np.stack([-(h[0] * m[1] + h[1] * n[1]), h[0] * m[0] + h[1] * n[0], np.zeros(grid.shape)])
    d = _fft3(zx * np.asarray(wt)[None])
    e = _ifft3(np.einsum("abxyz,bxyz->axyz", epsinv, d))
    en, em = (e * n).sum(0), (e * m).sum(0)
    kxct = np.stack([-en * mag, em * mag])
    return -np.real(np.vdot(to_flat3(h), to_flat3(kxct)))


def HMkH3(Hflat, k, epsinv, grid, basis=None):
    """``HMₖH`` on a 3D cell (``maxwell.jl:267-275``): the weighted form with ``wt = 1``
    (``grads/period.jl:64-66``)."""
    return HMkH_weighted(Hflat, k, epsinv, grid, np.ones(grid.shape), basis)


def dw2_dL(Hflat, k, epsinv, grid, basis=None):
    """``∂ω²_∂Λ`` (``grads/period.jl:88-90``): ``2 HMₖH_weighted(H, g_z/Λ)``."""
    return 2.0 * HMkH_weighted(Hflat, k, epsinv, grid, period_weight(grid), basis)


def E_field3(k, Hflat, epsinv, grid):
    """``E = ε⁻¹ . fft(kx_tc(H))`` on a 3D cell (``fields.jl:194-210``), ``(3,Nx,Ny,Nz)``."""
    mag, m, n = mag_mn3(k, grid)
    h = to_grid3(Hflat, grid)
    return np.einsum("abxyz,bxyz->axyz", epsinv, _fft3((h[1][None] * m - h[0][None] * n) * mag[None]))


def canonicalize_phase3(Hflat, k, epsinv, grid):
    """``canonicalize_phase`` (``fields.jl:258-260``) on a 3D cell: the first maximum of
    ``|E|^2`` in column-major order over ``(3, Nx, Ny, Nz)``."""
    flat = E_field3(k, Hflat, epsinv, grid).reshape(-1, order="F")
    idx = int(np.argmax(np.abs(flat) ** 2))
    return np.asarray(Hflat) * np.exp(-1j * np.angle(flat[idx]))


class HelmholtzOracle3:
    """Operator bundle at fixed ``(k, ε⁻¹, grid)`` on a 3D cell."""

    def __init__(self, k, epsinv, grid):
        self.k, self.epsinv, self.grid = float(k), epsinv, grid
        self.basis = mag_mn3(k, grid)

    def M(self, X):
        return apply_M3(X, self.k, self.epsinv, self.grid, self.basis)

    def P(self, X):
        return apply_P3(X, self.k, self.epsinv, self.grid, self.basis)

    def HMkH(self, H):
        return HMkH3(H, self.k, self.epsinv, self.grid, self.basis)


def solve_k_periodic(omega, epsinv, L, grid, **kw):
    """``solve_k_periodic`` (``grads/period.jl:144-148``): ``solve_k`` on the grid with
    period ``L``."""
    from .solve import solve_k
    return solve_k(omega, epsinv, grid_with_period(grid, L), **kw)


def solve_k_periodic_kbar_pullback(omega, epsinv, L, grid, kmags, evecs, kbar):
    """``_solve_k_period_grads`` (``grads/period.jl:182-234``) for cotangents on ``kmags``
    only (``ēvecs`` zero, so ``k̄ₕ = Λ̄_H = 0``): per band ``ω̄ += 2ω k̄/∂ω²∂k`` and
    ``Λ̄ += k̄ (-∂ω²∂Λ / ∂ω²∂k)``.  Returns ``(ω̄, Λ̄)``."""
    g = grid_with_period(grid, L)
    wbar, Lbar = 0.0, 0.0
    for kb, k, ev in zip(kbar, kmags, evecs):
        if kb is None or kb == 0:
            continue
        basis = mag_mn3(k, g)
        dk = 2.0 * HMkH3(ev, k, epsinv, g, basis)
        dL = dw2_dL(ev, k, epsinv, g, basis)
        wbar += 2.0 * omega * kb / dk
        Lbar += kb * (-dL / dk)
    return wbar, Lbar


def bragg_wg_epsi(grid, eps_core=4.0, eps_bg=2.1, w0=1.0, h0=0.6, mod=0.25):
    """Isotropic Gaussian core whose width is modulated along z, ``w(z) = w0 (1 + mod
    cos(2 pi z / Λ))`` (a Bragg grating), returned as ``ε⁻¹`` of shape
    ``(3,3,Nx,Ny,Nz)``."""
    xs, ys, zs = grid.x(), grid.y(), grid.z()
    w = w0 * (1.0 + mod * np.cos(2 * np.pi * zs / grid.Dz))
    eps = eps_bg + (eps_core - eps_bg) * np.exp(-(xs[:, None, None] ** 2 / w[None, None, :] ** 2
                                                  + ys[None, :, None] ** 2 / h0 ** 2))
    out = np.zeros((3, 3) + grid.shape)
    for a in range(3):
        out[a, a] = 1.0 / eps
    return out


def epsinv_bar3(d, lamd):
    """3D ``ε⁻¹_bar`` (``grads/solve.jl:464-487``): diagonal ``-Re(λd_a conj d_a)``; the upper
    off-diagonal entries ``(0,1)``, ``(1,2)``, ``(0,2)`` hold ``-Re(λd_j conj d_i + λd_i conj d_j)``
    and the lower ones stay zero (unlike the mirrored 2D form).  ``d``, ``lamd``:
    ``(3, Nx, Ny, Nz)``; returns ``(3, 3, Nx, Ny, Nz)``."""
    out = np.zeros((3, 3) + d.shape[1:])
    for a in range(3):
        out[a, a] = -np.real(lamd[a] * np.conj(d[a]))
    for i, j in ((0, 1), (1, 2), (0, 2)):
        out[i, j] = -np.real(lamd[j] * np.conj(d[i]) + lamd[i] * np.conj(d[j]))
    return out


def wbar_z(magbar, mbar, nbar, mag, m, n):
    """``_wbar_z`` (``grads/period.jl:97-119``): the z-component of the per-plane-wave
    cotangent of ``w = k_vec - g`` from those of ``(mag, m, n)``:
    ``w̄ = māg ŵ - (m̄.ŵ/mag) m - (n̄.ŵ/mag) n`` with ``ŵ = m x n``."""
    wx = m[1] * n[2] - m[2] * n[1]
    wy = m[2] * n[0] - m[0] * n[2]
    wz = m[0] * n[1] - m[1] * n[0]
    mw = (mbar[0] * wx + mbar[1] * wy + mbar[2] * wz) / mag
    nw = (nbar[0] * wx + nbar[1] * wy + nbar[2] * wz) / mag
    return magbar * wz - mw * m[2] - nw * n[2]


def solve_k_periodic_pullback(omega, epsinv, L, grid, kmags, evecs, kbar, evbar, adj_rtol=1e-11):
    """``_solve_k_period_grads`` (``grads/period.jl:182-234``): ``(k̄mags, ēvecs)`` ->
    ``(ω̄, ε̄⁻¹, Λ̄)``.  Eigenvector channel: ``λ`` from ``eig_adjt`` (re-orthogonalised), ``ε̄⁻¹``
    from ``d = fft(kx_tc ev)/N`` and ``λd = fft(kx_tc λ)``, and the per-plane-wave ``w̄_z`` from
    ``λẽ = bfft(ε⁻¹ λd/N)``, ``ẽ = bfft(ε⁻¹ d)``: ``k̄ₕ = -Σ w̄_z``, ``Λ̄_H = -Σ w̄_z g_z/Λ``.
    Eigenvalue channel with ``k̄tot = k̄ + k̄ₕ``: ``λ = (k̄tot/∂ω²∂k) ev``, ``ω̄ += 2ω k̄tot/∂ω²∂k``,
    ``Λ̄ += Λ̄_H + k̄tot (-∂ω²∂Λ/∂ω²∂k)``.  ``kbar``/``evbar`` entries may be ``None``.
    ``adj_rtol``: GMRES tolerance of ``eig_adjt``; 1e-11 is reached in FP64 (1e-13 stagnates at
    the iteration cap with results within 4e-9 of those at 1e-11)."""
    from .adjoint import eig_adjt
    g = grid_with_period(grid, L)
    N = g.N
    wt = period_weight(g)
    ei_bar = np.zeros_like(epsinv, dtype=float)
    w_bar, L_bar = 0.0, 0.0
    for i in range(len(kmags)):
        k = float(kmags[i])
        ev = np.asarray(evecs[i], dtype=np.complex128)
        kb = 0.0 if (kbar is None or kbar[i] is None) else float(kbar[i])
        eb = None if (evbar is None or evbar[i] is None) else np.asarray(evbar[i], dtype=np.complex128)
        op = HelmholtzOracle3(k, epsinv, g)
        mag, m, n = op.basis
        dk = 2.0 * op.HMkH(ev)
        dL = 2.0 * HMkH_weighted(ev, k, epsinv, g, wt, op.basis)
        evg = to_grid3(ev, g)
        kh, LH = 0.0, 0.0

        def kxtc(h):
            return (h[1][None] * m - h[0][None] * n) * mag[None]

        def eid3(x):
            return np.einsum("baxyz,bxyz->axyz", epsinv, x)

        if eb is not None and np.any(eb):
            lam = eig_adjt(op.M, omega ** 2, ev, 0.0, eb, P=op.P, rtol=adj_rtol)
            lam = lam - np.vdot(ev, lam) * ev
            lg = to_grid3(lam, g)
            d = _fft3(kxtc(evg / N))
            lamd = _fft3(kxtc(lg))
            ei_bar += epsinv_bar3(d, lamd)
            lamd = lamd / N
            lam_e = _bfft3(eid3(lamd))
            e_t = _bfft3(eid3(d))
            kxbar_m = np.real(lam_e * np.conj(evg[1])[None] + e_t * np.conj(lg[1])[None])
            kxbar_n = -np.real(lam_e * np.conj(evg[0])[None] + e_t * np.conj(lg[0])[None])
            magbar = (n * kxbar_n).sum(0) + (m * kxbar_m).sum(0)
            wz = wbar_z(magbar, kxbar_m * mag[None], kxbar_n * mag[None], mag, m, n)
            kh = -np.sum(wz)
            LH = -np.sum(wz * wt)
        ktot = kb + kh
        d = _fft3(kxtc(evg / N))
        lamd = _fft3(kxtc((ktot / dk) * evg))
        ei_bar += epsinv_bar3(d, lamd)
        w_bar += 2.0 * omega * ktot / dk
        L_bar += LH + ktot * (-dL / dk)
    return w_bar, ei_bar, L_bar
