"""Learned warm start of ``solve_k``: a small convolutional network (kernels in
``csrc/warmnet.hip``, trained on the GPU) predicts each band's effective index and a
low-resolution start block from the smoothed ε⁻¹ map plus ω.  The plain start stays the
fallback, and the converged result does not depend on the start (the same k_tol / eig_tol
stopping rules apply).  No counterpart in the reference, which starts every solve from
random vectors at ``k_guess`` (``src/MaxwellEigenmodes.jl:63``).

Representation:
- input: 7 maps of ``hw x hw`` -- the 6 unique ε⁻¹ components averaged over pixel blocks
  and the normalised ω, axes ``[channel][ix][iy]``;
- start vectors: per band, the ``hw x hw`` lowest plane waves of the eigenvector's two
  components (FFT order), transformed with an unnormalised forward FFT into a smooth
  localised map (4 channels: re/im of the two components);
- effective index ``n = k/ω`` per band, normalised as ``(n - N_SHIFT) / N_SCALE``.

Training data come from the solver itself (``generate``): TFLN ribs with random top width,
film, etch depth, sidewall angle and ω around the 256x256 rib workload.
"""
import ctypes
import os
import time

import numpy as np

from . import _lib
from . import Grid, OmxError

N_SHIFT, N_SCALE = 1.8, 0.2
OMEGA0, OMEGA_SCALE = 0.65, 0.05
DIAG0, DIAG_SCALE, OFF_SCALE = 0.55, 0.25, 0.05
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


class WarmNet:
    """The network on the GPU (``omx_warmnet_*``): FP32 forward, loss gradient and Adam."""

    def __init__(self, hw=64, nev=8, seed=0, device=0, bf16=False):
        self.hw, self.nev = int(hw), int(nev)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().omx_warmnet_create(int(device), self.hw, self.nev, int(seed), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int64()
        _lib.check(_lib.lib().omx_warmnet_num_params(self.h, ctypes.byref(n)))
        self.nparams = int(n.value)
        self.bf16 = False
        self.set_bf16(bf16)

    def set_bf16(self, on):
        """Convolutions in bf16 on the matrix cores (FP32 accumulation) or FP32."""
        _lib.check(_lib.lib().omx_warmnet_set_precision(self.h, int(bool(on))))
        self.bf16 = bool(on)

    def close(self):
        if self.h:
            _lib.lib().omx_warmnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        _lib.check(_lib.lib().omx_warmnet_get_params(self.h, _lib.fptr(out)))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise OmxError(_lib.OMX_ERR_BAD_SHAPE, f"expected {self.nparams} parameters, got {p.size}")
        _lib.check(_lib.lib().omx_warmnet_set_params(self.h, _lib.fptr(p)))

    def _x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 4 or x.shape[1:] != (7, self.hw, self.hw):
            raise OmxError(_lib.OMX_ERR_BAD_SHAPE, f"expected (B, 7, {self.hw}, {self.hw}) inputs, got {x.shape}")
        return x

    def forward(self, x):
        """``(fields (B, 4nev, hw, hw), k (B, nev))``."""
        x = self._x(x)
        B = x.shape[0]
        y = np.empty((B, 4 * self.nev, self.hw, self.hw), dtype=np.float32)
        k = np.empty((B, self.nev), dtype=np.float32)
        _lib.check(_lib.lib().omx_warmnet_forward(self.h, B, _lib.fptr(x), _lib.fptr(y), _lib.fptr(k)))
        return y, k

    def loss_grad(self, x, tf, tk, wk=1.0):
        x = self._x(x)
        B = x.shape[0]
        tf = np.ascontiguousarray(tf, dtype=np.float32)
        tk = np.ascontiguousarray(tk, dtype=np.float32)
        if tf.shape != (B, 4 * self.nev, self.hw, self.hw) or tk.shape != (B, self.nev):
            raise OmxError(_lib.OMX_ERR_BAD_SHAPE, "target shapes do not match the batch")
        loss = ctypes.c_double()
        g = np.empty(self.nparams, dtype=np.float32)
        _lib.check(_lib.lib().omx_warmnet_loss_grad(self.h, B, _lib.fptr(x), _lib.fptr(tf), _lib.fptr(tk), float(wk),
                                                    ctypes.byref(loss), _lib.fptr(g)))
        return loss.value, g

    def train(self, x, tf, tk, wk=1.0, epochs=100, batch=16, lr=1e-3, seed=0, resume=False):
        """Adam over the whole set on the device; returns the per-epoch mean loss.  With
        ``resume`` the optimiser continues the previous call's moments and step count."""
        x = self._x(x)
        n = x.shape[0]
        tf = np.ascontiguousarray(tf, dtype=np.float32)
        tk = np.ascontiguousarray(tk, dtype=np.float32)
        if tf.shape != (n, 4 * self.nev, self.hw, self.hw) or tk.shape != (n, self.nev):
            raise OmxError(_lib.OMX_ERR_BAD_SHAPE, "target shapes do not match the data set")
        hist = np.empty(int(epochs))
        fn = _lib.lib().omx_warmnet_train_resume if resume else _lib.lib().omx_warmnet_train
        _lib.check(fn(self.h, n, _lib.fptr(x), _lib.fptr(tf), _lib.fptr(tk), float(wk), int(epochs), int(batch),
                      float(lr), int(seed), _lib.dptr(hist)))
        return hist

    def forward_all(self, x, chunk=64):
        """``forward`` over a data set in chunks."""
        ys, ks = [], []
        for i in range(0, x.shape[0], chunk):
            y, k = self.forward(x[i:i + chunk])
            ys.append(y)
            ks.append(k)
        return np.concatenate(ys), np.concatenate(ks)


def _complex_maps(f, nev):
    """(n, 4 nev, hw, hw) re/im channels -> (n, nev, 2, hw, hw) complex band maps."""
    n, _, hw, _ = f.shape
    g = f.reshape(n, nev, 2, 2, hw, hw)
    return g[:, :, :, 0] + 1j * g[:, :, :, 1]


def _channel_maps(c):
    n, nev, _, hw, _ = c.shape
    out = np.empty((n, nev, 2, 2, hw, hw), dtype=np.float32)
    out[:, :, :, 0] = c.real
    out[:, :, :, 1] = c.imag
    return out.reshape(n, 4 * nev, hw, hw)


def align_targets(y, tf, nev):
    """The targets rotated within each sample's band span to best match the outputs: per
    sample, T W with the unitary W minimising sum_k ||y_k - (T W)_k||^2 (orthogonal
    Procrustes: W = U V^H from the SVD T^H y = U S V^H).  The field maps are a linear image
    of the band vectors, so T W are the maps of the rotated vectors, which span the same
    subspace -- all that the shared first eigensolve takes from the start block."""
    T, Y = _complex_maps(tf, nev), _complex_maps(y, nev)
    n, _, _, hw, _ = T.shape
    Tf, Yf = T.reshape(n, nev, -1), Y.reshape(n, nev, -1)
    M = np.matmul(Tf.conj(), Yf.transpose(0, 2, 1))    # (n, nev, nev): <T_j, y_k>
    U, _, Vh = np.linalg.svd(M)
    W = U @ Vh
    return _channel_maps(np.matmul(W.transpose(0, 2, 1), Tf).reshape(n, nev, 2, hw, hw))


def train_subspace(net, x, tf, tk, wk=1.0, epochs=600, realign=10, batch=16, lr=2e-3, seed=1, log=None):
    """Training on a loss invariant to rotations of the predicted block within each sample's
    band span: every ``realign`` epochs the targets are rotated by ``align_targets`` to the
    current outputs, and Adam continues on them (its state carried across the pieces).  The
    per-band effective-index outputs keep their band order."""
    hist = []
    for e in range(0, int(epochs), int(realign)):
        y, _ = net.forward_all(x)
        t_al = align_targets(y, tf, net.nev)
        h = net.train(x, t_al, tk, wk=wk, epochs=min(int(realign), int(epochs) - e), batch=batch, lr=lr,
                      seed=seed + e, resume=e > 0)
        hist.extend(h.tolist())
        if log and (e // int(realign)) % 10 == 0:
            log(f"subspace training: epoch {e + len(h)}, loss {h[-1]:.4f}")
    return np.array(hist)


def _blocks(grid, hw):
    if grid.Nx % hw or grid.Ny % hw:
        raise OmxError(_lib.OMX_ERR_BAD_SHAPE, f"grid {grid.Nx}x{grid.Ny} is not a multiple of the {hw} map")
    return grid.Nx // hw, grid.Ny // hw


def features(epsinv, grid, omega, hw=64):
    """``(7, hw, hw)`` float32 network input: block means of ε⁻¹ xx, yy, zz, xy, xz, yz and ω."""
    fx, fy = _blocks(grid, hw)
    e = np.asarray(epsinv, dtype=np.float64)
    a, b = np.array([0, 1, 2, 0, 0, 1]), np.array([0, 1, 2, 1, 2, 2])
    m = e[a, b].reshape(6, hw, fx, hw, fy).mean(axis=(2, 4))
    out = np.empty((7, hw, hw), dtype=np.float32)
    out[:3] = (m[:3] - DIAG0) / DIAG_SCALE
    out[3:6] = m[3:] / OFF_SCALE
    out[6] = (float(omega) - OMEGA0) / OMEGA_SCALE
    return out


def _sel(n, hw):
    return np.r_[0:hw // 2, n - hw // 2:n]


def field_targets(evecs, grid, hw=64):
    """``(4 nev, hw, hw)`` float32: per band, the ``hw x hw`` lowest plane waves of both
    components (FFT order) mapped by an unnormalised forward FFT, as re/im channels."""
    sx, sy = _sel(grid.Nx, hw), _sel(grid.Ny, hw)
    out = np.empty((4 * len(evecs), hw, hw), dtype=np.float32)
    for j, ev in enumerate(evecs):
        A = np.asarray(ev).reshape((2, grid.Nx, grid.Ny), order="F")
        u = np.fft.fft2(A[:, sx][:, :, sy], axes=(1, 2))
        out[4 * j + 0], out[4 * j + 1] = u[0].real, u[0].imag
        out[4 * j + 2], out[4 * j + 3] = u[1].real, u[1].imag
    return out


def _flat_index(grid, hw):
    """Flat positions (2, hw, hw) of the ``hw x hw`` lowest plane waves of both components
    in a vector of 2N entries (index c + 2(ix + Nx iy))."""
    sx, sy = _sel(grid.Nx, hw), _sel(grid.Ny, hw)
    return np.arange(2)[:, None, None] + 2 * (sx[None, :, None] + grid.Nx * sy[None, None, :])


def start_vectors(fields, grid, hw=64):
    """Inverse of ``field_targets``: unit-norm vectors of 2N complex (zero above the ``hw``
    lowest plane waves), one per band, as the rows of a ``(nev, 2N)`` array."""
    nev = fields.shape[0] // 4
    f = np.asarray(fields, dtype=np.float64).reshape(nev, 2, 2, hw, hw)
    C = np.fft.ifft2(f[:, :, 0] + 1j * f[:, :, 1], axes=(2, 3))  # (nev, 2, hw, hw)
    nrm = np.sqrt(np.sum(np.abs(C) ** 2, axis=(1, 2, 3)))
    C /= np.where(nrm > 0, nrm, 1.0)[:, None, None, None]
    out = np.zeros((nev, 2 * grid.N), dtype=np.complex128)
    out[:, _flat_index(grid, hw).ravel()] = C.reshape(nev, -1)
    return out


def lowres_vectors(fields, hw=64):
    """The start vectors of ``start_vectors`` by their ``hw x hw`` lowest plane waves only:
    rows of ``2 hw^2`` complex, entry ``c + 2(gx + hw gy)`` (FFT order), unit norm -- the
    ``guess_hw`` form of ``solve_k``, scattered into the start block on the GPU."""
    nev = fields.shape[0] // 4
    f = np.asarray(fields, dtype=np.float64).reshape(nev, 2, 2, hw, hw)
    C = np.fft.ifft2(f[:, :, 0] + 1j * f[:, :, 1], axes=(2, 3))  # (nev, 2, gx, gy)
    nrm = np.sqrt(np.sum(np.abs(C) ** 2, axis=(1, 2, 3)))
    C /= np.where(nrm > 0, nrm, 1.0)[:, None, None, None]
    return np.ascontiguousarray(C.transpose(0, 3, 2, 1)).reshape(nev, -1)


def neff_targets(kmags, omega):
    return ((np.asarray(kmags, dtype=np.float64) / float(omega) - N_SHIFT) / N_SCALE).astype(np.float32)


class LearnedStart:
    """A trained network plus its map size: ``guess(ε⁻¹, grid, ω)`` -> ``(kguess, Hguess)``."""

    def __init__(self, net):
        self.net = net

    @property
    def hw(self):
        return self.net.hw

    @property
    def nev(self):
        return self.net.nev

    def guess(self, epsinv, grid, omega, lowres=False):
        """``(kguess per band, start vectors)``: full vectors ``(nev, 2N)``, or with ``lowres``
        their lowest plane waves ``(nev, 2 hw^2)`` for ``solve_k(..., guess_hw=hw)``."""
        y, k = self.net.forward(features(epsinv, grid, omega, self.hw)[None])
        kg = float(omega) * (N_SHIFT + N_SCALE * k[0].astype(np.float64))
        return kg, (lowres_vectors(y[0], self.hw) if lowres else start_vectors(y[0], grid, self.hw))

    def save(self, path):
        np.savez(path, params=self.net.params, hw=self.hw, nev=self.nev, bf16=int(self.net.bf16))

    @classmethod
    def load(cls, path, device=0):
        """The network of a file, its convolutions in the precision it was trained with."""
        with np.load(path, allow_pickle=False) as z:
            net = WarmNet(int(z["hw"]), int(z["nev"]), device=device, bf16=bool(int(z["bf16"])) if "bf16" in z else False)
            net.params = z["params"]
        return cls(net)


def device_guess(start, epsinv, grid, omega, solver):
    """The guess as ``omx_solve_k_learned`` builds it on the GPU from the handle's ε⁻¹:
    ``(k outputs (nev,), low-plane-wave vectors (nev, 2 hw^2))``."""
    h = solver.handle(grid, epsinv)
    k = np.empty(start.nev, dtype=np.float32)
    low = np.empty((start.nev, 2 * start.hw * start.hw), dtype=np.complex128)
    _lib.check(_lib.lib().omx_warmnet_guess(start.net.h, h.h, float(omega), _lib.fptr(k), _lib.dptr(low)), h.h)
    return k, low


def solve_k_learned(omega, epsinv, grid, solver, start, nev=None, mode="seed", k0_scale=1.0, return_info=False,
                    on_device=True, **kw):
    """``solve_k`` started from the network's guesses; the stopping rules and the converged
    result are those of ``solve_k``.  ``mode="seed"``: the shared first eigensolve starts at
    ``k0_scale`` times the predicted k of band 1 from the predicted block (instead of k_guess
    and random vectors; 1.0 measured best with the rotation-invariant network, ahead of 1.015,
    1.03 and 1.05); ``mode="band"``: every band starts at its own predicted k
    with the predicted block and the shared first eigensolve is skipped (``band_guess``)."""
    from . import solve_k
    nev = start.nev if nev is None else int(nev)
    if nev > start.nev:
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, f"the network predicts {start.nev} bands, {nev} requested")
    if mode not in ("seed", "band"):
        raise OmxError(_lib.OMX_ERR_INVALID_ARG, "mode must be 'seed' or 'band'")
    if mode == "seed" and on_device:
        # features, forward and start block on the GPU (omx_solve_k_learned)
        from . import OmxInfo
        h = solver.handle(grid, epsinv)
        km = np.empty(nev)
        fetch = kw.pop("fetch_evecs", True)
        ev = np.empty((nev, 2 * grid.N), dtype=np.complex128) if fetch else None
        info = OmxInfo()
        _lib.check(_lib.lib().omx_solve_k_learned(
            h.h, start.net.h, float(omega), nev, float(k0_scale), float(kw.pop("k_tol", 1e-8)),
            float(kw.pop("eig_tol", 1e-8)), int(kw.pop("max_eigsolves", 60)), int(kw.pop("maxiter", 100)),
            _lib.dptr(km), _lib.dptr(ev), ctypes.byref(info)), h.h)
        if kw:
            raise OmxError(_lib.OMX_ERR_INVALID_ARG, f"unexpected arguments {sorted(kw)}")
        evecs = [ev[j].copy() for j in range(nev)] if fetch else None
        return (km, evecs, info.as_dict()) if return_info else (km, evecs)
    kg, H = start.guess(epsinv, grid, omega, lowres=True)
    kg = np.maximum(kg[:nev], 1e-3)
    if mode == "band":
        return solve_k(omega, epsinv, grid, solver, nev=nev, kguess=kg, Hguess=H[:nev], band_guess=True,
                       guess_hw=start.hw, return_info=return_info, **kw)
    return solve_k(omega, epsinv, grid, solver, nev=nev, kguess=k0_scale * kg[0], Hguess=H[:nev],
                   guess_hw=start.hw, return_info=return_info, **kw)


# --- training data from the solver ------------------------------------------------------

def random_rib(rng):
    """One draw of the rib distribution: TFLN rib parameters and ω around the 256x256 rib
    (top width 1.2 um, film 0.6 um, etch 0.3 um, 60 degree sidewalls, ω = 1/1.55)."""
    film = rng.uniform(0.5, 0.7)
    return dict(top_w=rng.uniform(0.8, 1.6), film=film, etch=film * rng.uniform(0.3, 0.7),
                sidewall_deg=rng.uniform(55.0, 75.0), omega=rng.uniform(0.58, 0.72))


def rib_epsinv(grid, prm, backend="gpu"):
    from . import geometry as G
    return G.tfln_rib(grid, top_w=prm["top_w"], film=prm["film"], etch=prm["etch"], sidewall_deg=prm["sidewall_deg"],
                      backend=backend)


def random_coupler(rng):
    """One draw of the coupler distribution: cross-sections of two Si3N4 cores in SiO2 around
    the EME chain's adiabatic coupler (gap 0.3 -> 1.5 um, widths 0.8 <-> 1.2 um, 0.8 um
    thick, ω = 1/1.55) on its 8 x 4 um 512 x 256 grid."""
    return dict(gap=rng.uniform(0.2, 1.7), w1=rng.uniform(0.7, 1.3), w2=rng.uniform(0.7, 1.3), h=rng.uniform(0.7, 0.9),
                omega=rng.uniform(0.60, 0.69))


def coupler_epsinv(grid, prm, backend="gpu"):
    from . import geometry as G
    return G.coupler_slice(grid, prm["gap"], prm["w1"], prm["w2"], h=prm["h"], backend=backend)


# structure families: (parameter draw, ε⁻¹ of a draw)
FAMILIES = {"rib": (random_rib, rib_epsinv), "coupler": (random_coupler, coupler_epsinv)}


def generate(seeds, grid, nev, solver, hw=64, k_tol=1e-10, eig_tol=1e-8, log=None, family="rib"):
    """Inputs and targets for the draws of ``seeds`` of a structure family: ``(X, TF, TK,
    infos)``."""
    from . import solve_k
    draw, build = FAMILIES[family]
    X = np.empty((len(seeds), 7, hw, hw), dtype=np.float32)
    TF = np.empty((len(seeds), 4 * nev, hw, hw), dtype=np.float32)
    TK = np.empty((len(seeds), nev), dtype=np.float32)
    infos = []
    for i, s in enumerate(seeds):
        prm = draw(np.random.default_rng(int(s)))
        ei = build(grid, prm)
        km, ev, info = solve_k(prm["omega"], ei, grid, solver, nev=nev, k_tol=k_tol, eig_tol=eig_tol,
                               return_info=True)
        X[i] = features(ei, grid, prm["omega"], hw)
        TF[i] = field_targets(ev, grid, hw)
        TK[i] = neff_targets(km, prm["omega"])
        infos.append(info)
        if log and (i + 1) % 32 == 0:
            log(f"generated {i + 1}/{len(seeds)}")
    return X, TF, TK, infos


def compare(seeds, grid, nev, start, solver, k_tol=1e-10, eig_tol=1e-8, mode="seed", family="rib"):
    """Held-out draws solved from the plain start and from the learned start: per draw the
    two k vectors, LOBPCG iterations and wall times."""
    from . import solve_k
    draw, build = FAMILIES[family]
    rows = []
    for s in seeds:
        prm = draw(np.random.default_rng(int(s)))
        ei = build(grid, prm)
        w = prm["omega"]
        t0 = time.perf_counter()
        kp, _, ip = solve_k(w, ei, grid, solver, nev=nev, k_tol=k_tol, eig_tol=eig_tol, return_info=True,
                            fetch_evecs=False)
        t1 = time.perf_counter()
        kl, _, il = solve_k_learned(w, ei, grid, solver, start, nev=nev, mode=mode, k_tol=k_tol, eig_tol=eig_tol,
                                    return_info=True, fetch_evecs=False)
        t2 = time.perf_counter()
        rows.append(dict(seed=int(s), k_plain=kp, k_learned=kl, it_plain=ip["lobpcg_iters_fp32"] + ip["lobpcg_iters_fp64"],
                         it_learned=il["lobpcg_iters_fp32"] + il["lobpcg_iters_fp64"], newton_plain=ip["n_newton"],
                         newton_learned=il["n_newton"], ms_plain=1e3 * (t1 - t0), ms_learned=1e3 * (t2 - t1),
                         conv_plain=ip["converged"], conv_learned=il["converged"]))
    return rows


def train_default(n=1024, epochs=600, out=None, log=print):
    """The shipped network (``data/warmnet_tfln256.npz``): 1024 rib draws (seeds 1000..2023)
    solved at 256x256 with 8 bands, 600 epochs of Adam (lr 2e-3, minibatch 16, wk 1) on the
    GPU with the rotation-invariant field loss (``train_subspace``, targets re-aligned every
    10 epochs) and the convolutions in bf16 on the matrix cores (FP32 accumulation, FP32
    master weights and weight gradient)."""
    from . import HIPSolver
    grid = Grid(8.0, 4.0, 256, 256)
    nev, hw = 8, 64
    t = time.perf_counter()
    X, TF, TK, _ = generate(range(1000, 1000 + n), grid, nev, HIPSolver(), hw=hw, log=log)
    log(f"generated {n} draws in {time.perf_counter() - t:.1f} s")
    net = WarmNet(hw, nev, seed=0, bf16=True)
    t = time.perf_counter()
    hist = train_subspace(net, X, TF, TK, wk=1.0, epochs=epochs, realign=10, batch=16, lr=2e-3, seed=1, log=log)
    log(f"trained {epochs} epochs in {time.perf_counter() - t:.1f} s: loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    st = LearnedStart(net)
    if out:
        st.save(out)
    return st, hist


COUPLER_GRID = (8.0, 4.0, 512, 256)
COUPLER_NEV = 16
COUPLER_FILE = "warmnet_coupler512x256.npz"


def train_coupler(n=320, epochs=500, out=None, log=print):
    """The EME chain's network (``data/warmnet_coupler512x256.npz``): ``n`` coupler draws
    (seeds 5000..) solved at 512 x 256 with 16 modes, then Adam (lr 2e-3, minibatch 16) on
    the GPU with the rotation-invariant field loss and bf16 convolutions.  It starts the chain's cold cells
    (``eme.eme(..., learned=...)``)."""
    from . import HIPSolver
    grid = Grid(*COUPLER_GRID)
    nev, hw = COUPLER_NEV, 64
    t = time.perf_counter()
    X, TF, TK, _ = generate(range(5000, 5000 + n), grid, nev, HIPSolver(), hw=hw, log=log, family="coupler")
    log(f"generated {n} coupler draws in {time.perf_counter() - t:.1f} s")
    net = WarmNet(hw, nev, seed=0, bf16=True)
    t = time.perf_counter()
    hist = train_subspace(net, X, TF, TK, wk=1.0, epochs=epochs, realign=10, batch=16, lr=2e-3, seed=1, log=log)
    log(f"trained {epochs} epochs in {time.perf_counter() - t:.1f} s: loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    st = LearnedStart(net)
    if out:
        st.save(out)
    return st, hist


def load_coupler(device=0):
    """The shipped coupler network, or None when its file is absent."""
    path = os.path.join(DATA_DIR, COUPLER_FILE)
    return LearnedStart.load(path, device) if os.path.exists(path) else None


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="train a warm-start network on solver-converged draws of a family")
    ap.add_argument("--family", default="rib", choices=sorted(FAMILIES))
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--epochs", type=int, default=None)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.family == "rib":
        train_default(a.n or 1024, a.epochs or 600, a.out or os.path.join(DATA_DIR, "warmnet_tfln256.npz"))
    else:
        train_coupler(a.n or 320, a.epochs or 500, a.out or os.path.join(DATA_DIR, COUPLER_FILE))
