"""omx: MI355X (gfx950) plane-wave Maxwell eigensolver, Python host interface.

This mirrors the MaxwellEigenmodes interface of OptiMode.jl for the HIP backend,
with the same names, argument meanings, layouts and error behaviour:

* ``Grid(Dx, Dy, Nx, Ny)``                 -- lib/DielectricSmoothing/src/grid.jl:46-71
* ``HIPSolver(device=0, ...)``             -- a new ``AbstractEigensolver`` beside
  ``GPUSolver`` (lib/MaxwellEigenmodes/src/solvers/gpu.jl:37-47)
* ``solve_omega2(k, epsinv, grid, solver; nev, maxiter, tol)``  -- src/solve.jl:53-63
* ``solve_k(omega, epsinv, grid, solver; nev, max_eigsolves, maxiter, k_tol, eig_tol,
  kguess, Hguess)``                        -- src/solve.jl:182-196
* ``rrule_solve_k(...)`` -> ``((kmags, evecs), pullback)`` -- src/solve.jl:213-289
* ``HelmholtzMap(k, epsinv, grid, solver)`` with ``*``/``precondition``/``HMkH``
                                           -- src/maxwell.jl:834-949, :256-275

``epsinv`` is a ``(3, 3, Nx, Ny)`` float64 array indexed ``[a, b, ix, iy]`` (the Julia
index order); eigenvectors are complex128 vectors of length ``2*Nx*Ny`` in the
reference's flat order ``c + 2*(ix + Nx*iy)``.  Every computation runs in libomx.so on
the GPU; nothing here falls back to the CPU.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import OmxError, OmxInfo, OMX_PREC_FP32, OMX_PREC_FP64  # noqa: F401

__all__ = ["Grid", "Grid3", "HIPSolver", "HelmholtzMap", "solve_omega2", "solve_k", "solve_k_batch", "rrule_solve_k",
           "group_index", "ng_gvd", "k_guess", "solve_k_periodic", "period_derivatives", "rrule_solve_k_periodic",
           "sliceinv_3x3", "rrule_sliceinv_3x3", "OmxError", "library_path", "exported_symbols", "kernel_stats"]


def library_path():
    return _lib.LIB_PATH


def exported_symbols():
    return [name for name, _, _ in _lib.SIGNATURES]


class Grid:
    """2D cell ``Dx x Dy`` (um) on ``Nx x Ny`` pixels (``Grid{2}``)."""

    def __init__(self, Dx, Dy, Nx, Ny):
        self.Dx, self.Dy, self.Nx, self.Ny = float(Dx), float(Dy), int(Nx), int(Ny)

    @property
    def N(self):
        return self.Nx * self.Ny

    def __repr__(self):
        return f"Grid({self.Dx}, {self.Dy}, {self.Nx}, {self.Ny})"


class Grid3:
    """3D cell ``Dx x Dy x Dz`` (um) on ``Nx x Ny x Nz`` pixels, periodic along z with period
    ``Dz`` (``Grid{3}``, lib/DielectricSmoothing/src/grid.jl:46-71)."""

    def __init__(self, Dx, Dy, Dz, Nx, Ny, Nz):
        self.Dx, self.Dy, self.Dz = float(Dx), float(Dy), float(Dz)
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)

    @property
    def N(self):
        return self.Nx * self.Ny * self.Nz

    def key(self):
        return (self.Nx, self.Ny, self.Nz, self.Dx, self.Dy, self.Dz)

    def __repr__(self):
        return f"Grid3({self.Dx}, {self.Dy}, {self.Dz}, {self.Nx}, {self.Ny}, {self.Nz})"


def _grid_key(grid):
    return grid.key() if isinstance(grid, Grid3) else (grid.Nx, grid.Ny, 1, grid.Dx, grid.Dy, 0.0)


def k_guess(omega, epsinv):
    """``k_guess`` (src/MaxwellEigenmodes.jl:63)."""
    return float(omega) * np.sqrt(1.0 / min(np.min(epsinv[a, a]) for a in range(3)))


def _epsinv_c(epsinv, grid):
    e = np.asarray(epsinv, dtype=np.float64)
    shape = (3, 3, grid.Nx, grid.Ny) + ((grid.Nz,) if isinstance(grid, Grid3) else ())
    if e.shape != shape:
        raise OmxError(_lib.OMX_ERR_BAD_SHAPE, f"epsinv must have shape {shape}, got {e.shape}")
    # column-major flat order: a + 3b + 9(ix + Nx iy [+ Nx Ny iz])
    return np.ascontiguousarray(e.transpose(tuple(range(e.ndim - 1, -1, -1))))


class _Handle:
    """One ``omx_solver`` handle: a grid and an eps^-1 field resident on the GPU."""

    def __init__(self, device, grid, epsinv_c):
        L = _lib.lib()
        h = ctypes.c_void_p()
        if isinstance(grid, Grid3):
            st = L.omx_create_3d(ctypes.byref(h), int(device), grid.Nx, grid.Ny, grid.Nz, grid.Dx, grid.Dy, grid.Dz,
                                 _lib.dptr(epsinv_c))
        else:
            st = L.omx_create(ctypes.byref(h), int(device), grid.Nx, grid.Ny, grid.Dx, grid.Dy, _lib.dptr(epsinv_c))
        if st != _lib.OMX_OK:
            raise OmxError(st, L.omx_last_error(None).decode() or L.omx_status_string(st).decode())
        self.h = h
        self.grid = grid
        self.eps_src = None  # the eps^-1 array (caller layout) now on the device

    def set_epsinv(self, epsinv_c):
        _lib.check(_lib.lib().omx_set_epsinv(self.h, _lib.dptr(epsinv_c)), self.h)

    def close(self):
        if self.h:
            _lib.lib().omx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HIPSolver:
    """HIP/gfx950 eigensolver backend.

    ``precision`` is the policy of the iterative phase: ``"mixed"`` (default: FP32
    operator applications, then an FP64 refinement phase) or ``"fp64"``.  Tolerance
    knobs of the FP32 phase are relative to ω²: ``fp32_tol`` (``solve_k``'s LOBPCG locking
    threshold, 2e-3: the FP64 refinement finishes the vectors) and ``fp32_switch``
    (|Δω²|/ω² at which Newton moves to FP64).  ``reg_ops=False`` keeps
    the FP32 operator on the LDS Stockham FFT even where the register-FFT kernels apply
    (power-of-two grids 64 to 512).  ``plain=True`` runs the reference's algorithm on the same
    kernels, as the baseline the optimised path is measured against: FP64 throughout, an
    nev-column block for every band, a random start at ``k_guess`` per band and every
    eigensolve to ``eig_tol`` (``solve.jl:89-113``).  ``warm_start=True`` makes ``solve_k``
    calls without ``kguess``/``Hguess`` start every band from the previous solve's k and
    eigenvectors, which stay on the device (the EME slice chain, ``eme.jl:58-76``).
    """

    def __init__(self, device=0, precision="mixed", fp32_tol=None, fp32_switch=None, maxiter_fp64=None, seed=None,
                 profile=False, reg_ops=None, plain=False, warm_start=False):
        if precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        self.device = int(device)
        self.precision = precision
        self.options = {}
        if fp32_tol is not None:
            self.options["fp32_tol"] = float(fp32_tol)
        if fp32_switch is not None:
            self.options["fp32_switch"] = float(fp32_switch)
        if maxiter_fp64 is not None:
            self.options["maxiter_fp64"] = float(maxiter_fp64)
        if seed is not None:
            self.options["seed"] = float(seed)
        if reg_ops is not None:
            self.options["reg_ops"] = 1.0 if reg_ops else 0.0
        self.options["plain"] = 1.0 if plain else 0.0
        self.options["warm_start"] = 1.0 if warm_start else 0.0
        self.options["band_guess"] = 0.0  # set per call by solve_k(..., band_guess=True)
        self.options["guess_lowres"] = 0.0  # set per call by solve_k(..., guess_hw=hw)
        self.options["fp64_only"] = 1.0 if precision == "fp64" else 0.0
        self.profile = bool(profile)
        self._h = None

    def handle(self, grid, epsinv):
        h = self._h
        same_grid = h is not None and _grid_key(h.grid) == _grid_key(grid)
        if not same_grid:
            ec = _epsinv_c(epsinv, grid)
            if h is not None:
                h.close()
            h = self._h = _Handle(self.device, grid, ec)
            h.eps_src = np.array(epsinv, dtype=np.float64, copy=True)
        elif not (np.shape(epsinv) == h.eps_src.shape and np.array_equal(epsinv, h.eps_src)):
            h.set_epsinv(_epsinv_c(epsinv, grid))
            h.eps_src = np.array(epsinv, dtype=np.float64, copy=True)
        L = _lib.lib()
        for k, v in self.options.items():
            _lib.check(L.omx_set_option(h.h, k.encode(), v), h.h)
        if self.profile and not getattr(h, "profiling", False):
            _lib.check(L.omx_set_option(h.h, b"profile", 1.0), h.h)
            h.profiling = True
        return h

    def __repr__(self):
        return f"HIPSolver(device={self.device}, precision={self.precision!r})"


def kernel_stats(solver, family):
    """``(total_ms, launches, algorithmic_bytes)`` of a kernel family timed live on the
    handle's stream (requires ``profile=True`` on the solver)."""
    h = solver._h
    if h is None:
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, "solver has no handle yet")
    ms, n, b = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    _lib.check(_lib.lib().omx_get_kernel_stats(h.h, family.encode(), ctypes.byref(ms), ctypes.byref(n),
                                               ctypes.byref(b)), h.h)
    return ms.value, n.value, b.value


def _cols(H, n2):
    H = np.asarray(H)
    if H.ndim == 1:
        H = H[:, None]
    if H.shape[0] != n2:
        raise OmxError(_lib.OMX_ERR_BAD_SHAPE, f"vectors must have length {n2}")
    # contiguous per vector: (m, 2N)
    return np.ascontiguousarray(H.T, dtype=np.complex128)


class HelmholtzMap:
    """Matrix-free ``M(k)`` on the GPU (``HelmholtzMap``, src/maxwell.jl:834-949)."""

    def __init__(self, k, epsinv, grid, solver=None, precision=OMX_PREC_FP64):
        self.k = float(k)
        self.grid = grid
        self.solver = solver if solver is not None else HIPSolver()
        self.h = self.solver.handle(grid, epsinv)
        self.precision = precision

    def _run(self, fn, H):
        n2 = 2 * self.grid.N
        X = _cols(H, n2)
        Y = np.empty_like(X)
        _lib.check(fn(self.h.h, self.k, _lib.dptr(X), _lib.dptr(Y), X.shape[0], self.precision), self.h.h)
        return Y.T if np.ndim(H) == 2 else Y[0]

    def __mul__(self, H):
        return self._run(_lib.lib().omx_apply_helmholtz, H)

    def precondition(self, H):
        return self._run(_lib.lib().omx_apply_preconditioner, H)

    def HMH(self, H, A=None):
        """``HMH`` (src/maxwell.jl:288-291): ``Re <H, kx_ct(ifft(A . fft(kx_tc H)))>`` per column
        in FP64; ``A`` a symmetric (3,3,Nx,Ny) field, default ε⁻¹ (then ``<H, M H>``)."""
        n2 = 2 * self.grid.N
        X = _cols(H, n2)
        out = np.empty(X.shape[0])
        Ac = None if A is None else _epsinv_c(A, self.grid)
        _lib.check(_lib.lib().omx_hmh(self.h.h, self.k, _lib.dptr(Ac), _lib.dptr(X), X.shape[0], _lib.dptr(out)),
                   self.h.h)
        return out if np.ndim(H) == 2 else float(out[0])

    def HMkH(self, H):
        n2 = 2 * self.grid.N
        X = _cols(H, n2)
        out = np.empty(X.shape[0])
        _lib.check(_lib.lib().omx_hmkh(self.h.h, self.k, _lib.dptr(X), X.shape[0], _lib.dptr(out), self.precision),
                   self.h.h)
        return out if np.ndim(H) == 2 else float(out[0])


def solve_omega2(k, epsinv, grid, solver, nev=1, maxiter=100, tol=1e-8, evecs_guess=None, return_info=False):
    """``solve_ω²(k, ε⁻¹, grid, solver; nev, maxiter, tol, evecs_guess)`` -> ``(ω², evecs)``."""
    h = solver.handle(grid, epsinv)
    n2 = 2 * grid.N
    H = np.zeros((nev, n2), dtype=np.complex128)
    use_guess = evecs_guess is not None
    if use_guess:
        G = np.asarray(evecs_guess)
        G = G[:, None] if G.ndim == 1 else G
        if isinstance(evecs_guess, (list, tuple)):
            G = np.stack([np.asarray(v) for v in evecs_guess], 1)
        H[:G.shape[1]] = G.T[:nev]
    w2 = np.empty(nev)
    info = OmxInfo()
    st = _lib.lib().omx_solve_omega2(h.h, float(k), int(nev), float(tol), int(maxiter),
                                     _lib.dptr(H) if use_guess else None, _lib.dptr(w2), ctypes.byref(info))
    _lib.check(st, h.h)
    if not use_guess:
        _lib.check(_lib.lib().omx_get_evecs(h.h, _lib.dptr(H)), h.h)
    evecs = [H[j].copy() for j in range(nev)]
    if return_info:
        return w2, evecs, info.as_dict()
    return w2, evecs


def solve_k(omega, epsinv, grid, solver, nev=1, max_eigsolves=60, maxiter=100, k_tol=1e-8, eig_tol=1e-8,
            kguess=None, Hguess=None, return_info=False, fetch_evecs=True, band_guess=False, guess_hw=None):
    """``solve_k(ω, ε⁻¹, grid, solver; nev, max_eigsolves, maxiter, k_tol, eig_tol, kguess,
    Hguess)`` -> ``(kmags, evecs)``: unit-norm, phase-canonicalised eigenvectors.
    ``band_guess=True`` (the learned warm start): ``kguess`` holds one k per band and
    ``Hguess`` the start block; the shared first eigensolve is skipped.  ``guess_hw=hw``:
    the rows of ``Hguess`` hold only the ``hw x hw`` lowest plane waves of each start vector
    (``2 hw^2`` entries, index ``c + 2(gx + hw gy)`` in FFT order), scattered on the GPU."""
    h = solver.handle(grid, epsinv)
    n2 = 2 * grid.N
    km = np.empty(nev)
    kg = None
    if kguess is not None:
        kv = np.ravel(np.asarray(kguess, dtype=np.float64))
        if band_guess:
            if kv.size != nev or Hguess is None:
                raise OmxError(_lib.OMX_ERR_INVALID_ARG, "band_guess needs one kguess per band and Hguess")
            kg = np.ascontiguousarray(kv)
        else:
            kg = kv[:1].copy()
    elif band_guess:
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, "band_guess needs one kguess per band and Hguess")
    Hg = None
    ng = 0
    if guess_hw is not None:
        hw = int(guess_hw)
        Hg = np.ascontiguousarray(np.atleast_2d(Hguess)[:nev], dtype=np.complex128)
        if Hg.shape[1] != 2 * hw * hw:
            raise OmxError(_lib.OMX_ERR_BAD_SHAPE, f"guess_hw={hw} needs rows of {2 * hw * hw} entries")
        ng = Hg.shape[0]
    elif isinstance(Hguess, np.ndarray) and Hguess.ndim == 2 and Hguess.shape[1] == n2 and Hguess.shape[0] != n2:
        # rows are the start vectors (as omx.warmstart builds them): no per-column copies
        Hg = np.ascontiguousarray(Hguess[:nev], dtype=np.complex128)
        ng = Hg.shape[0]
    elif Hguess is not None:
        cols = [np.asarray(Hguess)] if np.ndim(Hguess) == 1 else (
            [np.asarray(v) for v in Hguess] if isinstance(Hguess, (list, tuple)) else
            [np.asarray(Hguess)[:, j] for j in range(np.shape(Hguess)[1])])
        cols = [c for c in cols if len(c) == n2][:nev]
        if cols:
            Hg = np.ascontiguousarray(np.stack(cols), dtype=np.complex128)
            ng = len(cols)
    ev = np.empty((nev, n2), dtype=np.complex128) if fetch_evecs else None
    info = OmxInfo()
    L = _lib.lib()
    if band_guess:
        _lib.check(L.omx_set_option(h.h, b"band_guess", 1.0), h.h)
    if guess_hw is not None:
        _lib.check(L.omx_set_option(h.h, b"guess_lowres", float(guess_hw)), h.h)
    try:
        st = L.omx_solve_k(h.h, float(omega), int(nev), float(k_tol), float(eig_tol), int(max_eigsolves),
                           int(maxiter), _lib.dptr(kg), _lib.dptr(Hg), ng, _lib.dptr(km), _lib.dptr(ev),
                           ctypes.byref(info))
    finally:
        if band_guess:
            L.omx_set_option(h.h, b"band_guess", 0.0)
        if guess_hw is not None:
            L.omx_set_option(h.h, b"guess_lowres", 0.0)
    _lib.check(st, h.h)
    evecs = [ev[j].copy() for j in range(nev)] if fetch_evecs else None
    if return_info:
        return km, evecs, info.as_dict()
    return km, evecs


def solve_k_batch(problems, precision="mixed", device=0, fetch_evecs=False, max_group=0):
    """Many independent ``solve_k`` problems in one call (``omx_solve_k_batch``): each problem
    is a dict with ``grid``, ``epsinv``, ``omega``, ``nev`` and optionally ``k_tol``,
    ``eig_tol`` (1e-8), ``max_eigsolves`` (60), ``maxiter`` (100).  Problems may differ in
    grid, ε⁻¹, ω and nev.  Problems on one grid run as launch groups (``max_group`` problems at
    most when positive, else as many as fit half the free device memory): every band of
    every problem of a group is an instance of the same kernel launches.  Returns one dict
    per problem: ``kmags``, ``evecs`` (when ``fetch_evecs``), ``info`` and ``status``
    (0 = OMX_OK)."""
    n = len(problems)
    arr = (_lib.OmxProblem * max(1, n))()
    keep = []
    outs = []
    for j, p in enumerate(problems):
        g = p["grid"]
        ec = _epsinv_c(p["epsinv"], g)
        km = np.zeros(int(p["nev"]))
        ev = np.zeros((int(p["nev"]), 2 * g.N), dtype=np.complex128) if fetch_evecs else None
        keep += [ec, km, ev]
        q = arr[j]
        q.nx, q.ny, q.Dx, q.Dy = g.Nx, g.Ny, float(g.Dx), float(g.Dy)
        q.epsinv = _lib.dptr(ec)
        q.omega = float(p["omega"])
        q.nev = int(p["nev"])
        q.max_eigsolves = int(p.get("max_eigsolves", 60))
        q.maxiter = int(p.get("maxiter", 100))
        q.k_tol = float(p.get("k_tol", 1e-8))
        q.eig_tol = float(p.get("eig_tol", 1e-8))
        q.kmags_out = _lib.dptr(km)
        q.evecs_out = _lib.dptr(ev)
        outs.append((km, ev))
    L = _lib.lib()
    st = L.omx_solve_k_batch(int(device), n, arr, int(max_group), 1 if precision == "fp64" else 0)
    if st in (_lib.OMX_ERR_NO_DEVICE, _lib.OMX_ERR_INVALID_ARG) and all(arr[j].status == _lib.OMX_ERR_INTERNAL
                                                                          for j in range(n)):
        _lib.check(st, None)
    res = []
    for j, (km, ev) in enumerate(outs):
        res.append({"kmags": km, "evecs": None if ev is None else [ev[i].copy() for i in range(len(km))],
                    "info": arr[j].info.as_dict(), "status": int(arr[j].status)})
    return res


def group_index(k, ev, omega, epsinv, deps_domega, grid, solver):
    """``group_index(k, ev, ω, ε⁻¹, ∂ε_∂ω, grid)`` (lib/ModeAnalysis/src/group_index.jl:21-29):
    ``n_g = (ω + HMH(ev, ε⁻¹ ∂ωε ε⁻¹)/2) / HMₖH(ev)`` for one mode or a list of modes (``k``
    and ``ev`` matching); ``deps_domega=None`` for a non-dispersive medium."""
    h = solver.handle(grid, epsinv)
    single = np.ndim(k) == 0
    ks = np.atleast_1d(np.asarray(k, dtype=np.float64))
    evs = [ev] if single else list(ev)
    if len(evs) != len(ks):
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, "one eigenvector per k is needed")
    E = _cols(np.stack(evs, 1), 2 * grid.N)
    D = None if deps_domega is None else _epsinv_c(deps_domega, grid)
    ng = np.empty(len(ks))
    _lib.check(_lib.lib().omx_group_index(h.h, float(omega), len(ks), _lib.dptr(ks), _lib.dptr(E), _lib.dptr(D),
                                          _lib.dptr(ng)), h.h)
    return float(ng[0]) if single else ng


def ng_gvd(omega, k, ev, epsinv, deps_domega, d2eps_domega2, grid, solver, return_E=False):
    """``ng_gvd(ω, k, ev, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid)`` (lib/ModeAnalysis/src/group_index.jl:126-162):
    ``(n_g, GVD)`` of one mode, or arrays of both for a list of modes (``k`` and ``ev``
    matching).  ``return_E=True`` adds the real-space ``E = ε⁻¹ fft(kx_tc(ev))`` of each mode,
    shape ``(3, Nx, Ny)``, as ``ng_gvd_E`` (``:55-95``) returns it."""
    h = solver.handle(grid, epsinv)
    single = np.ndim(k) == 0
    ks = np.atleast_1d(np.asarray(k, dtype=np.float64))
    evs = [ev] if single else list(ev)
    if len(evs) != len(ks):
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, "one eigenvector per k is needed")
    E = _cols(np.stack(evs, 1), 2 * grid.N)
    De = _epsinv_c(deps_domega, grid)
    DDe = _epsinv_c(d2eps_domega2, grid)
    ng = np.empty(len(ks))
    gvd = np.empty(len(ks))
    Ef = np.empty(3 * grid.N * len(ks), dtype=np.complex128) if return_E else None
    _lib.check(_lib.lib().omx_ng_gvd(h.h, float(omega), len(ks), _lib.dptr(ks), _lib.dptr(E), _lib.dptr(De),
                                     _lib.dptr(DDe), _lib.dptr(ng), _lib.dptr(gvd), _lib.dptr(Ef)), h.h)
    if return_E:
        Es = [Ef[3 * grid.N * j:3 * grid.N * (j + 1)].reshape((3, grid.Nx, grid.Ny), order="F") for j in range(len(ks))]
        return (float(ng[0]), float(gvd[0]), Es[0]) if single else (ng, gvd, Es)
    return (float(ng[0]), float(gvd[0])) if single else (ng, gvd)


def rrule_solve_k(omega, epsinv, grid, solver, nev=1, **kw):
    """``rrule(solve_k, ω, ε⁻¹, grid, solver; ...)`` -> ``((kmags, evecs), pullback)``.

    ``pullback((kbar, evbar))`` returns ``(None, ω̄, ε̄⁻¹, None, None)``; ``kbar`` and
    ``evbar`` may be ``None`` or contain ``None`` entries (zero cotangents)."""
    kmags, evecs = solve_k(omega, epsinv, grid, solver, nev=nev, **kw)
    h = solver.handle(grid, epsinv)

    def pullback(cot):
        wbar, eibar, _ = _adjoint(h, grid, omega, nev, kmags, evecs, cot, period=False)
        return (None, wbar, eibar, None, None)

    return (kmags, evecs), pullback


def _adjoint(h, grid, omega, nev, kmags, evecs, cot, period):
    """Shared body of the solve_k / solve_k_periodic pull-backs: cotangents ``(kbar, evbar)``
    (``None`` entries are zero) -> ``(ω̄, ε̄⁻¹, Λ̄ or None)`` through the C-ABI."""
    kbar, evbar = cot
    kb = np.zeros(nev)
    if kbar is not None:
        for j, v in enumerate(kbar):
            kb[j] = 0.0 if v is None else float(v)
    eb = np.zeros((nev, 2 * grid.N), dtype=np.complex128)
    mask = np.zeros(nev, dtype=np.uint8)
    if evbar is not None:
        for j, v in enumerate(evbar):
            if v is not None:
                eb[j] = np.asarray(v, dtype=np.complex128)
                mask[j] = 1
    E = np.ascontiguousarray(np.stack(evecs), dtype=np.complex128)
    wbar, Lbar = np.zeros(1), np.zeros(1)
    # column-major (3,3,Nx,Ny[,Nz]): C order reversed
    dims = (grid.Nx, grid.Ny) + ((grid.Nz,) if isinstance(grid, Grid3) else ())
    eibar_c = np.zeros(tuple(reversed(dims)) + (3, 3))
    args = [h.h, float(omega), int(nev), _lib.dptr(np.ascontiguousarray(kmags, dtype=np.float64)), _lib.dptr(E),
            _lib.dptr(kb), _lib.dptr(eb), mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _lib.dptr(wbar),
            _lib.dptr(eibar_c)]
    if period:
        st = _lib.lib().omx_solve_k_periodic_adjoint(*args, _lib.dptr(Lbar))
    else:
        st = _lib.lib().omx_solve_k_adjoint(*args)
    _lib.check(st, h.h)
    eibar = np.ascontiguousarray(eibar_c.transpose(tuple(range(eibar_c.ndim - 1, -1, -1))))
    return float(wbar[0]), eibar, (float(Lbar[0]) if period else None)


def _grid_with_period(grid, L):
    """``_grid_with_period`` (lib/MaxwellEigenmodes/src/grads/period.jl:123-125)."""
    return Grid3(grid.Dx, grid.Dy, float(L), grid.Nx, grid.Ny, grid.Nz)


def solve_k_periodic(omega, epsinv, L, grid, solver, nev=1, **kw):
    """``solve_k_periodic(ω, ε⁻¹, Λ, grid, solver; nev, ...)``
    (lib/MaxwellEigenmodes/src/grads/period.jl:144-148): ``solve_k`` on the 3D cell with its
    period set to ``Λ`` (ε⁻¹ held fixed in index space)."""
    if not isinstance(grid, Grid3):
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, "solve_k_periodic needs a Grid3")
    return solve_k(omega, epsinv, _grid_with_period(grid, L), solver, nev=nev, **kw)


def period_derivatives(kmags, evecs, epsinv, grid, solver):
    """Per mode: ``(∂ω²/∂k, ∂ω²/∂Λ)`` = ``(2 HMₖH(ev), ∂ω²_∂Λ(ev))`` on a 3D cell
    (lib/MaxwellEigenmodes/src/grads/period.jl:88-90), FP64 on the GPU."""
    if not isinstance(grid, Grid3):
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, "period derivatives need a Grid3")
    h = solver.handle(grid, epsinv)
    ks = np.atleast_1d(np.asarray(kmags, dtype=np.float64))
    E = _cols(np.stack(list(evecs), 1), 2 * grid.N)
    dk, dL = np.empty(len(ks)), np.empty(len(ks))
    _lib.check(_lib.lib().omx_period_derivatives(h.h, len(ks), _lib.dptr(ks), _lib.dptr(E), _lib.dptr(dk),
                                                 _lib.dptr(dL)), h.h)
    return dk, dL


def rrule_solve_k_periodic(omega, epsinv, L, grid, solver, nev=1, **kw):
    """``rrule(solve_k_periodic, ω, ε⁻¹, Λ, grid, solver; nev)`` (period.jl:236-249) ->
    ``((kmags, evecs), pullback)``.  ``pullback((kbar, evbar))`` returns
    ``(None, ω̄, ε̄⁻¹, Λ̄, None, None)`` (``_solve_k_period_grads``, period.jl:163-234): both
    channels, the eigenvector one through ``eig_adjt`` and the plane-wave basis, on the GPU
    in FP64; ε̄⁻¹ is ``(3,3,Nx,Ny,Nz)`` with the off-diagonal sums above the diagonal."""
    g = _grid_with_period(grid, L)
    kmags, evecs = solve_k(omega, epsinv, g, solver, nev=nev, **kw)
    h = solver.handle(g, epsinv)

    def pullback(cot):
        wbar, eibar, Lbar = _adjoint(h, g, omega, nev, kmags, evecs, cot, period=True)
        return (None, wbar, eibar, Lbar, None, None)

    return (kmags, evecs), pullback


def _tensor_field(A):
    a = np.asarray(A, dtype=np.float64)
    if a.ndim < 2 or a.shape[:2] != (3, 3):
        raise OmxError(_lib.OMX_ERR_BAD_SHAPE, f"expected a (3,3,...) tensor field, got {a.shape}")
    # column-major flat order: pixel p at 9p + a + 3b
    return np.ascontiguousarray(a.transpose(tuple(range(a.ndim - 1, -1, -1)))), a.shape


def _tensor_back(flat, shape):
    return np.ascontiguousarray(flat.transpose(tuple(range(flat.ndim - 1, -1, -1)))).reshape(shape)


def sliceinv_3x3(A, device=0):
    """``sliceinv_3x3(A)`` (lib/MaxwellEigenmodes/src/linalg.jl:126-149): the per-pixel inverse of
    a ``(3,3,Nx,Ny[,Nz])`` tensor field, FP64 on the GPU."""
    a, shape = _tensor_field(A)
    y = np.empty_like(a)
    _lib.check(_lib.lib().omx_sliceinv_3x3(int(device), a.size // 9, _lib.dptr(a), _lib.dptr(y)))
    return _tensor_back(y, shape)


def rrule_sliceinv_3x3(A, device=0):
    """``rrule(sliceinv_3x3, A)`` (src/grads/linalg.jl:40-47) -> ``(Y, pullback)`` with
    ``pullback(Ȳ) = Ā = -Yᵀ Ȳ Yᵀ`` per pixel, on the GPU."""
    Y = sliceinv_3x3(A, device)

    def pullback(Ybar):
        y, shape = _tensor_field(Y)
        b, _ = _tensor_field(Ybar)
        if b.shape != y.shape:
            raise OmxError(_lib.OMX_ERR_BAD_SHAPE, "the cotangent must have the shape of the inverse")
        out = np.empty_like(y)
        _lib.check(_lib.lib().omx_sliceinv_3x3_pullback(int(device), y.size // 9, _lib.dptr(y), _lib.dptr(b),
                                                        _lib.dptr(out)))
        return _tensor_back(out, shape)

    return Y, pullback
