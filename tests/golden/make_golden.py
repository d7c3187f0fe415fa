"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no golden values for this path (its tests are invariants), so the
fixtures are produced by the oracle, which tests/test_oracle.py pins against closed forms,
an independent dense assembly and finite differences.  Usage:
    python tests/golden/make_golden.py [small|c1|c2|all] [--cache DIR]
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "optimode.jl_amd"))

import oracle as O  # noqa: E402
from omx import geometry as G  # noqa: E402

W0 = 1 / 1.55


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


FP_SEED = 4242      # random projection vectors of the eigenvector fingerprints
FP_NPROJ = 64
FP_HW = 16          # lowest hw x hw plane waves kept per mode
ADJ_SEED = 777      # cotangents and contraction directions of the adjoint fixture
ADJ_NDIR = 3
ADJ_BLOCK = 16      # block-sum downsampling of the cotangent fields


def fingerprint_vectors(n2, seed=FP_SEED, nproj=FP_NPROJ):
    """Seeded complex Gaussian projection vectors (nproj, 2N): E|r|^2 = 2 per entry."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nproj, n2)) + 1j * rng.standard_normal((nproj, n2))


def low_planewaves(ev, nx, ny, hw=FP_HW):
    """The hw x hw lowest plane waves of an eigenvector (2, Nx, Ny) column-major, FFT order."""
    h = np.asarray(ev).reshape((2, nx, ny), order="F")
    ix = np.r_[0:hw // 2, nx - hw // 2:nx]
    iy = np.r_[0:hw // 2, ny - hw // 2:ny]
    return h[:, ix][:, :, iy]


def adjoint_directions(shape, seed=ADJ_SEED, ndir=ADJ_NDIR):
    """Seeded symmetric (3,3,Nx,Ny) contraction directions."""
    rng = np.random.default_rng(seed + 1)
    out = []
    for _ in range(ndir):
        r = rng.standard_normal(shape)
        out.append((r + np.transpose(r, (1, 0, 2, 3))) / 2)
    return out


def adjoint_cotangents(nev, n2, seed=ADJ_SEED):
    """k̄ (nev values) and the vector c of the eigenvector cotangent of band 0,
    ēv_0 = 2 c <c, ev_0> (the cotangent of |<ev_0, c>|^2)."""
    rng = np.random.default_rng(seed)
    kbar = rng.standard_normal(nev)
    c = rng.standard_normal(n2) + 1j * rng.standard_normal(n2)
    return kbar, c


def block_sums(f, b=ADJ_BLOCK):
    """Six unique components (xx, yy, zz, xy, xz, yz) of a (3,3,Nx,Ny) field summed over
    b x b pixel blocks: (6, Nx/b, Ny/b)."""
    comps = [f[0, 0], f[1, 1], f[2, 2], f[0, 1], f[0, 2], f[1, 2]]
    nx, ny = f.shape[2], f.shape[3]
    return np.stack([c.reshape(nx // b, b, ny // b, b).sum((1, 3)) for c in comps])


def adjoint_fixture(grid, ei, km, evs, adj_rtol=1e-13):
    """Pull-backs of the oracle at the fixture's solution: channel 'k' (k̄ on every band) and
    channel 'v' (ēv on band 0), each as ω̄, contractions of ε̄⁻¹ and of ε̄ = sliceinv_3x3
    pull-back (the eps cotangent, grads/linalg.jl:40-47) with seeded symmetric directions,
    and block sums of both fields."""
    nev = len(km)
    kbar, c = adjoint_cotangents(nev, 2 * grid.N)
    dirs = adjoint_directions(ei.shape)
    out = {}
    evbar = [2 * c * np.vdot(c, evs[0])] + [None] * (nev - 1)
    for ch, (kb, eb) in (("k", (list(kbar), None)), ("v", ([None] * nev, evbar))):
        t = time.time()
        wbar, eibar = O.solve_k_pullback(W0, ei, grid, km, evs, kb, eb, adj_rtol=adj_rtol)
        epsbar = O.sliceinv_3x3_pullback(ei, eibar)
        out[f"adj_{ch}_wbar"] = wbar
        out[f"adj_{ch}_eibar_dirs"] = np.array([np.sum(eibar * d) for d in dirs])
        out[f"adj_{ch}_epsbar_dirs"] = np.array([np.sum(epsbar * d) for d in dirs])
        out[f"adj_{ch}_eibar_blocks"] = block_sums(eibar)
        out[f"adj_{ch}_epsbar_blocks"] = block_sums(epsbar)
        print(f"  adjoint channel {ch}: wbar {wbar:.12g}, {time.time() - t:.1f}s", flush=True)
    return out


def fields_fixture(grid, km, evs):
    """Eigenvector fingerprints: projections onto seeded random vectors and the lowest plane
    waves (after the oracle's phase canonicalisation)."""
    R = fingerprint_vectors(2 * grid.N)
    proj = np.stack([R.conj() @ v for v in evs])
    low = np.stack([low_planewaves(v, grid.Nx, grid.Ny) for v in evs])
    return dict(fp_proj=proj, fp_low=low, fp_low_energy=np.array([np.sum(np.abs(l) ** 2) for l in low]))


CACHE = None  # directory for the oracle's solutions (--cache DIR): a rerun skips the solve


def solve_and_save(name, grid, ei, nev, eig_tol, k_tol, extra=None):
    t = time.time()
    cpath = os.path.join(CACHE, name + "_solution.npz") if CACHE else None
    if cpath and os.path.exists(cpath):
        c = np.load(cpath)
        km, evs = c["kmags"], list(c["evecs"])
        info = [dict(n_eigsolves=n, lobpcg_iters=[it]) for n, it in zip(c["newton_steps"], c["lobpcg_iters"])]
        dt = float(c["oracle_seconds"])
    else:
        km, evs, info = O.solve_k(W0, ei, grid, nev=nev, k_tol=k_tol, eig_tol=eig_tol, maxiter=300,
                                  return_info=True, lock_band_only=True)
        dt = time.time() - t
        if cpath:
            os.makedirs(CACHE, exist_ok=True)
            np.savez(cpath, kmags=km, evecs=np.stack(evs), oracle_seconds=dt,
                     newton_steps=np.array([i["n_eigsolves"] for i in info]),
                     lobpcg_iters=np.array([sum(i["lobpcg_iters"]) for i in info]))
    slopes = np.array([2 * O.HMkH(v, k, ei, grid) for k, v in zip(km, evs)])
    ng = np.array([O.group_index(k, v, W0, ei, np.zeros_like(ei), grid) for k, v in zip(km, evs)])
    res = np.array([np.linalg.norm(O.apply_M(v, k, ei, grid) - W0 ** 2 * v) for k, v in zip(km, evs)])
    out = dict(omega=W0, kmags=km, dw2dk=slopes, ng=ng, resid=res, nx=grid.Nx, ny=grid.Ny, Dx=grid.Dx,
               Dy=grid.Dy, nev=nev, eig_tol=eig_tol, k_tol=k_tol, epsinv_sha256=sha(ei), oracle_seconds=dt,
               newton_steps=np.array([i["n_eigsolves"] for i in info]),
               lobpcg_iters=np.array([sum(i["lobpcg_iters"]) for i in info]))
    if extra:
        out.update(extra(km, evs))
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    print(name, "k =", km, "n_eff =", km / W0, "resid", res, f"{dt:.1f}s", flush=True)


def small():
    g = O.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], g)
    rng = np.random.default_rng(1234)
    X = rng.standard_normal((2 * g.N, 3)) + 1j * rng.standard_normal((2 * g.N, 3))
    op = O.HelmholtzOracle(1.25, ei, g)

    def extra(km, evs):
        return dict(evecs=np.stack(evs), X=X, MX=op.M(X), PX=op.P(X),
                    HMkH_X=np.array([op.HMkH(X[:, j]) for j in range(3)]), k_apply=1.25)
    solve_and_save("small_gauss_24x16", g, ei, 2, 1e-11, 1e-13, extra)


def _config_extra(g, ei, adj_rtol=1e-13):
    def extra(km, evs):
        out = fields_fixture(g, km, evs)
        out.update(adjoint_fixture(g, ei, km, evs, adj_rtol))
        return out
    return extra


def c1():
    g = O.Grid(6.0, 4.0, 128, 128)
    ei = G.rect_waveguide(g)
    solve_and_save("c1_rect_128", g, ei, 4, 1e-10, 1e-13, _config_extra(g, ei))


def c2():
    g = O.Grid(8.0, 4.0, 256, 256)
    ei = G.tfln_rib(g)
    # GMRES of the eigenvector channel at 256^2 reaches 1e-11, not 1e-13, in FP64
    solve_and_save("c2_tfln_256", g, ei, 8, 1e-10, 1e-13, _config_extra(g, ei, adj_rtol=1e-11))


if __name__ == "__main__":
    args = sys.argv[1:]
    if "--cache" in args:
        j = args.index("--cache")
        CACHE = args[j + 1]
        del args[j:j + 2]
    which = args[0] if args else "all"
    if which in ("small", "all"):
        small()
    if which in ("c1", "all"):
        c1()
    if which in ("c2", "all"):
        c2()
