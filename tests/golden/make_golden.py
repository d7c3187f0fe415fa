"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no golden values for this path (its tests are invariants), so the
fixtures are produced by the oracle, which tests/test_oracle.py pins against closed forms,
an independent dense assembly and finite differences.  Usage:
    python tests/golden/make_golden.py [small|c1|c2|all]
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


def solve_and_save(name, grid, ei, nev, eig_tol, k_tol, extra=None):
    t = time.time()
    km, evs, info = O.solve_k(W0, ei, grid, nev=nev, k_tol=k_tol, eig_tol=eig_tol, maxiter=300,
                              return_info=True, lock_band_only=True)
    dt = time.time() - t
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


def c1():
    g = O.Grid(6.0, 4.0, 128, 128)
    ei = G.rect_waveguide(g)
    solve_and_save("c1_rect_128", g, ei, 4, 1e-10, 1e-13)


def c2():
    g = O.Grid(8.0, 4.0, 256, 256)
    ei = G.tfln_rib(g)
    solve_and_save("c2_tfln_256", g, ei, 8, 1e-10, 1e-13)


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("small", "all"):
        small()
    if which in ("c1", "all"):
        c1()
    if which in ("c2", "all"):
        c2()
