"""Host-side input preparation: polygon cross-sections -> Kottke-smoothed ε⁻¹.

Produces the ``(3, 3, Nx, Ny)`` inverse-permittivity input of the eigensolver for the
benchmark geometries (rectangular cores, etched ribs with slanted sidewalls, coupler
cross-sections).  The averaging follows Kottke, Farjadpour & Johnson, Phys. Rev. E 77,
036611 (2008), the scheme of ``lib/DielectricSmoothing/src/kottke.jl:44-95``
(``τ_trans``/``τ⁻¹_trans``/``avg_param``): pixels cut by one interface get
``S τ⁻¹(f τ(Sᵀε₁S) + (1-f) τ(Sᵀε₂S)) Sᵀ`` with ``S`` the interface frame.  The fill
fraction ``f`` and the normal come from supersampling the pixel (the reference locates the
surface with GeometryPrimitives' ``surfpt_nearby``/``volfrac``, which is not available
here), so values at interface pixels differ from the reference's at the 1e-3 level while
bulk pixels are exact.
"""
import numpy as np

__all__ = ["Polygon", "Rect", "kottke_average", "smooth_epsinv", "rect_waveguide", "tfln_rib", "ridge_waveguide",
           "coupler_slice", "diag_eps"]


def diag_eps(*d):
    if len(d) == 1:
        d = (d[0],) * 3
    return np.diag(np.asarray(d, dtype=float))


class Polygon:
    """Convex polygon (vertices counter-clockwise) filled with tensor ``eps``."""

    def __init__(self, vertices, eps):
        self.v = np.asarray(vertices, dtype=float)
        self.eps = np.asarray(eps, dtype=float) if np.ndim(eps) == 2 else diag_eps(float(eps))

    def contains(self, x, y):
        inside = np.ones(np.broadcast(x, y).shape, dtype=bool)
        n = len(self.v)
        for i in range(n):
            x0, y0 = self.v[i]
            x1, y1 = self.v[(i + 1) % n]
            inside &= (x1 - x0) * (y - y0) - (y1 - y0) * (x - x0) >= 0
        return inside


def Rect(cx, cy, wx, wy, eps):
    return Polygon([(cx - wx / 2, cy - wy / 2), (cx + wx / 2, cy - wy / 2), (cx + wx / 2, cy + wy / 2),
                    (cx - wx / 2, cy + wy / 2)], eps)


def _tau(e):
    """τ transform (kottke.jl: τ_trans), batched over leading axes."""
    i = 1.0 / e[..., 0, 0]
    t = np.empty_like(e)
    t[..., 0, 0] = -i
    t[..., 1, 0] = e[..., 1, 0] * i
    t[..., 2, 0] = e[..., 2, 0] * i
    t[..., 0, 1] = e[..., 0, 1] * i
    t[..., 1, 1] = e[..., 1, 1] - e[..., 1, 0] * e[..., 0, 1] * i
    t[..., 2, 1] = e[..., 2, 1] - e[..., 2, 0] * e[..., 0, 1] * i
    t[..., 0, 2] = e[..., 0, 2] * i
    t[..., 1, 2] = e[..., 1, 2] - e[..., 1, 0] * e[..., 0, 2] * i
    t[..., 2, 2] = e[..., 2, 2] - e[..., 2, 0] * e[..., 0, 2] * i
    return t


def _tau_inv(t):
    """Inverse τ transform (kottke.jl: τ⁻¹_trans)."""
    i = 1.0 / t[..., 0, 0]
    e = np.empty_like(t)
    e[..., 0, 0] = -i
    e[..., 1, 0] = -t[..., 1, 0] * i
    e[..., 2, 0] = -t[..., 2, 0] * i
    e[..., 0, 1] = -t[..., 0, 1] * i
    e[..., 1, 1] = t[..., 1, 1] - t[..., 1, 0] * t[..., 0, 1] * i
    e[..., 2, 1] = t[..., 2, 1] - t[..., 2, 0] * t[..., 0, 1] * i
    e[..., 0, 2] = -t[..., 0, 2] * i
    e[..., 1, 2] = t[..., 1, 2] - t[..., 1, 0] * t[..., 0, 2] * i
    e[..., 2, 2] = t[..., 2, 2] - t[..., 2, 0] * t[..., 0, 2] * i
    return e


def kottke_average(eps1, eps2, normal, f):
    """Kottke average of ``eps1`` (fill fraction ``f``) and ``eps2`` across a plane interface
    with in-plane unit normal ``normal`` = (n_x, n_y) (lib/DielectricSmoothing/src/kottke.jl:90-95, ``avg_param``):
    ``S τ⁻¹(f τ(Sᵀε₁S) + (1-f) τ(Sᵀε₂S)) Sᵀ`` with ``S = [n, ẑ×n, ẑ]``."""
    n = np.array([normal[0], normal[1], 0.0])
    S = np.stack([n, np.array([-n[1], n[0], 0.0]), np.array([0.0, 0.0, 1.0])], 1)
    tau = f * _tau(S.T @ np.asarray(eps1, float) @ S) + (1 - f) * _tau(S.T @ np.asarray(eps2, float) @ S)
    return S @ _tau_inv(tau) @ S.T


def smooth_epsinv(shapes, eps_background, grid, nsub=8, backend="host", device=0):
    """Kottke-smoothed ε⁻¹ of ``shapes`` (earlier shapes take precedence where they
    overlap) over ``eps_background`` on ``grid`` (any object with Dx, Dy, Nx, Ny).
    Returns ``(3, 3, Nx, Ny)`` float64.  ``backend="gpu"`` runs the same algorithm in the
    HIP library (``omx_smooth_epsinv``)."""
    if backend == "gpu":
        return _smooth_epsinv_gpu(shapes, eps_background, grid, nsub, device)
    if backend != "host":
        raise ValueError("backend must be 'host' or 'gpu'")
    Nx, Ny = grid.Nx, grid.Ny
    dx, dy = grid.Dx / Nx, grid.Dy / Ny
    xc = -grid.Dx / 2 + np.arange(Nx) * dx
    yc = -grid.Dy / 2 + np.arange(Ny) * dy
    mats = [np.asarray(eps_background, float) if np.ndim(eps_background) == 2 else diag_eps(float(eps_background))]
    mats += [s.eps for s in shapes]
    nm = len(mats)
    # sub-sample offsets within a pixel centred at (xc, yc), spanning [-d/2, d/2)
    off = (np.arange(nsub) + 0.5) / nsub - 0.5
    sx = (xc[:, None] + off[None, :] * dx)   # (Nx, nsub)
    sy = (yc[:, None] + off[None, :] * dy)   # (Ny, nsub)
    X = sx[:, None, :, None] * np.ones((1, Ny, 1, nsub))
    Y = sy[None, :, None, :] * np.ones((Nx, 1, nsub, 1))
    mat = np.zeros(X.shape, dtype=np.int32)
    for idx in range(len(shapes) - 1, -1, -1):
        mat[shapes[idx].contains(X, Y)] = idx + 1
    counts = np.stack([(mat == m).sum(axis=(2, 3)) for m in range(nm)], -1)  # (Nx, Ny, nm)
    ns = nsub * nsub
    present = counts > 0
    npres = present.sum(-1)
    eps = np.zeros((Nx, Ny, 3, 3))
    # uniform pixels
    dom = np.argmax(counts, -1)
    for m in range(nm):
        eps[dom == m] = mats[m]
    # two-material pixels: Kottke average with the first-moment normal
    two = np.argwhere(npres >= 2)
    # sub-sample offsets in units of d / (2 nsub): odd integers summing to zero
    wo = 2 * np.arange(nsub) + 1 - nsub
    for (ix, iy) in two:
        c = counts[ix, iy]
        order = np.argsort(-c, kind="stable")
        m1, m2 = order[0], order[1]
        ind = mat[ix, iy] == m1
        f = c[m1] / ns
        # first moment of the material-1 indicator about its mean, in exact integers (the
        # mean term drops out because the offsets sum to zero); -moment points from
        # material 1 into material 2
        gx = float((ind * wo[:, None]).sum()) * dx
        gy = float((ind * wo[None, :]).sum()) * dy
        nrm = np.sqrt(gx * gx + gy * gy)
        if nrm == 0 or npres[ix, iy] > 2:
            ei = f * np.linalg.inv(mats[m1]) + (1 - f) * np.linalg.inv(mats[m2])
            eps[ix, iy] = np.linalg.inv(ei)
            continue
        eps[ix, iy] = kottke_average(mats[m1], mats[m2], (-gx / nrm, -gy / nrm), f)
    eps = 0.5 * (eps + np.swapaxes(eps, -1, -2))
    einv = np.linalg.inv(eps)
    einv = 0.5 * (einv + np.swapaxes(einv, -1, -2))
    return np.ascontiguousarray(np.moveaxis(einv, (-2, -1), (0, 1)))


def _smooth_epsinv_gpu(shapes, eps_background, grid, nsub, device):
    import ctypes
    from . import _lib
    bg = np.asarray(eps_background, float) if np.ndim(eps_background) == 2 else diag_eps(float(eps_background))
    nverts = (ctypes.c_int * max(1, len(shapes)))(*[len(sh.v) for sh in shapes])
    verts = np.ascontiguousarray(np.concatenate([sh.v.reshape(-1) for sh in shapes]) if shapes else np.zeros(2))
    eps_s = np.ascontiguousarray(np.concatenate([sh.eps.reshape(-1, order="F") for sh in shapes])
                                 if shapes else np.zeros(9))
    bgf = np.ascontiguousarray(bg.reshape(-1, order="F"))
    out = np.empty(9 * grid.Nx * grid.Ny)
    L = _lib.lib()
    st = L.omx_smooth_epsinv(int(device), grid.Nx, grid.Ny, float(grid.Dx), float(grid.Dy), len(shapes), nverts,
                             _lib.dptr(verts), _lib.dptr(eps_s), _lib.dptr(bgf), int(nsub), _lib.dptr(out))
    _lib.check(st, None)
    # column-major (3, 3, Nx, Ny) -> numpy (3, 3, Nx, Ny)
    return np.ascontiguousarray(out.reshape(grid.Ny, grid.Nx, 3, 3).transpose(3, 2, 1, 0))


# ---------------------------------------------------------------------------------------
# benchmark geometries
# ---------------------------------------------------------------------------------------

EPS_SIO2 = 2.085
EPS_SI3N4 = 3.96
EPS_LN = (4.89, 4.57, 4.89)


def rect_waveguide(grid, w=1.6, h=0.7, eps_core=EPS_SI3N4, eps_clad=EPS_SIO2, nsub=8, backend="host"):
    """Centred rectangular core ``w x h`` in a uniform cladding (Si₃N₄ in SiO₂)."""
    return smooth_epsinv([Rect(0.0, 0.0, w, h, eps_core)], eps_clad, grid, nsub, backend)


def _rib_shapes(top_w, film, etch, sidewall_deg, eps_film, eps_sub, y0=0.0, x0=0.0):
    slab = film - etch
    b = etch / np.tan(np.radians(sidewall_deg))
    rib = Polygon([(x0 - top_w / 2 - b, y0 + slab), (x0 + top_w / 2 + b, y0 + slab), (x0 + top_w / 2, y0 + film),
                   (x0 - top_w / 2, y0 + film)], diag_eps(*eps_film) if np.ndim(eps_film) else eps_film)
    return rib, slab


def tfln_rib(grid, top_w=1.2, film=0.6, etch=0.3, sidewall_deg=60.0, eps_ln=EPS_LN, eps_sub=EPS_SIO2,
             eps_clad=1.0, nsub=8, backend="host"):
    """x-cut thin-film LiNbO₃ rib (diag ε), SiO₂ below, air above; film bottom at y = 0."""
    big = 10 * (grid.Dx + grid.Dy)
    rib, slab = _rib_shapes(top_w, film, etch, sidewall_deg, eps_ln, eps_sub)
    slabp = Polygon([(-big, 0.0), (big, 0.0), (big, slab), (-big, slab)], diag_eps(*eps_ln))
    sub = Polygon([(-big, -big), (big, -big), (big, 0.0), (-big, 0.0)], eps_sub)
    return smooth_epsinv([rib, slabp, sub], eps_clad, grid, nsub, backend)


def ridge_waveguide(grid, w, h, sidewall_deg, eps_core, eps_sub=EPS_SIO2, eps_clad=1.0, nsub=8, backend="host"):
    """Fully etched trapezoidal ridge of top width ``w`` and height ``h`` on SiO₂."""
    big = 10 * (grid.Dx + grid.Dy)
    b = h / np.tan(np.radians(sidewall_deg))
    core = Polygon([(-w / 2 - b, 0.0), (w / 2 + b, 0.0), (w / 2, h), (-w / 2, h)],
                   eps_core if np.ndim(eps_core) == 2 else diag_eps(float(eps_core)))
    sub = Polygon([(-big, -big), (big, -big), (big, 0.0), (-big, 0.0)], eps_sub)
    return smooth_epsinv([core, sub], eps_clad, grid, nsub, backend)


def coupler_slice(grid, gap, w1, w2, h=0.8, eps_core=EPS_SI3N4, eps_clad=EPS_SIO2, nsub=8, backend="host"):
    """Two rectangular Si₃N₄ cores of widths ``w1``, ``w2`` separated by ``gap``."""
    x1 = -(gap / 2 + w1 / 2)
    x2 = gap / 2 + w2 / 2
    return smooth_epsinv([Rect(x1, 0.0, w1, h, eps_core), Rect(x2, 0.0, w2, h, eps_core)], eps_clad, grid, nsub,
                         backend)


def bragg_grating(grid, eps_core=4.0, eps_bg=2.1, w0=1.0, h0=0.4, mod=0.5):
    """ε⁻¹ ``(3,3,Nx,Ny,Nz)`` of one period of a Bragg-grating waveguide on a ``Grid3``: an
    isotropic Gaussian index core ``ε = ε_bg + (ε_core - ε_bg) exp(-(x²/w(z)² + y²/h0²))``
    whose width is modulated along the period, ``w(z) = w0 (1 + mod cos(2π z/Δz))``, on
    the pixel centres ``-D/2 + i D/N`` of each axis (grid.jl:108-127)."""
    xs = -grid.Dx / 2 + np.arange(grid.Nx) * (grid.Dx / grid.Nx)
    ys = -grid.Dy / 2 + np.arange(grid.Ny) * (grid.Dy / grid.Ny)
    zs = -grid.Dz / 2 + np.arange(grid.Nz) * (grid.Dz / grid.Nz)
    w = w0 * (1.0 + mod * np.cos(2 * np.pi * zs / grid.Dz))
    eps = eps_bg + (eps_core - eps_bg) * np.exp(-(xs[:, None, None] ** 2 / w[None, None, :] ** 2
                                                  + ys[None, :, None] ** 2 / h0 ** 2))
    out = np.zeros((3, 3, grid.Nx, grid.Ny, grid.Nz))
    for a in range(3):
        out[a, a] = 1.0 / eps
    return out
