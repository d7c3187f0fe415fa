"""GPU HMH, group index and GVD against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
import omx
from test_oracle_analysis import aniso_dispersive

pytestmark = pytest.mark.gpu

P_WG = [4.2, 2.1, 1.0, 0.6]
W0 = 1 / 1.55


def _sym_field(og, seed):
    """A smooth symmetric per-pixel 3x3 field (not diagonal, not eps^-1)."""
    rng = np.random.default_rng(seed)
    xs, ys = og.x(), og.y()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    B = rng.standard_normal((3, 3))
    A = np.einsum("ab,xy->abxy", B + B.T, 0.3 + np.exp(-(X ** 2 + Y ** 2)))
    A += np.einsum("ab,xy->abxy", np.eye(3), 1.5 + 0.2 * np.cos(X))
    return A


def test_hmh_matches_oracle_and_rayleigh_quotient():
    print("HMH vs oracle", flush=True)
    og = O.Grid(6.0, 4.0, 24, 16)
    g = omx.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi(P_WG, og)
    k = 1.2
    solver = omx.HIPSolver(precision="fp64")
    Mg = omx.HelmholtzMap(k, ei, g, solver)
    rng = np.random.default_rng(5)
    H = rng.standard_normal((2 * og.N, 3)) + 1j * rng.standard_normal((2 * og.N, 3))
    A = _sym_field(og, 1)
    ref = np.array([O.HMH(H[:, j], k, A, og) for j in range(3)])
    assert np.allclose(Mg.HMH(H, A), ref, rtol=1e-12, atol=0)
    # A = eps^-1: <H, M H>; on a unit eigenvector it is the eigenvalue (runtests.jl:61)
    ref_e = np.array([O.HMH(H[:, j], k, ei, og) for j in range(3)])
    assert np.allclose(Mg.HMH(H), ref_e, rtol=1e-12, atol=0)
    w2, evs = omx.solve_omega2(k, ei, g, solver, nev=2, tol=1e-11, maxiter=300)
    for lam, v in zip(w2, evs):
        assert Mg.HMH(v) == pytest.approx(lam, rel=1e-11)


def test_group_index_matches_oracle():
    print("group index vs oracle", flush=True)
    og = O.Grid(6.0, 4.0, 24, 16)
    g = omx.Grid(6.0, 4.0, 24, 16)
    ei = O.gaussian_wg_epsi(P_WG, og)
    solver = omx.HIPSolver(precision="fp64")
    km, evs = omx.solve_k(W0, ei, g, solver, nev=2, k_tol=1e-12, eig_tol=1e-11)
    # non-dispersive: n_g = ω / HMkH
    ng0 = omx.group_index(km, evs, W0, ei, None, g, solver)
    zero = np.zeros_like(ei)
    ref0 = np.array([O.group_index(k, v, W0, ei, zero, og) for k, v in zip(km, evs)])
    assert np.allclose(ng0, ref0, rtol=1e-11, atol=0)
    # a synthetic dispersive term: symmetric and dominated by a positive multiple of ε
    eps = O.sliceinv_3x3(ei)
    deps = 0.05 * eps + 0.005 * _sym_field(og, 2)
    ng = omx.group_index(km, evs, W0, ei, deps, g, solver)
    ref = np.array([O.group_index(k, v, W0, ei, deps, og) for k, v in zip(km, evs)])
    assert np.allclose(ng, ref, rtol=1e-11, atol=0)
    assert np.all(ng > ng0)  # positive material dispersion raises the group index
    # single-mode call form
    assert omx.group_index(km[0], evs[0], W0, ei, deps, g, solver) == pytest.approx(ng[0], rel=1e-14)


def _gpu_mode(profile, omega, g, og, nev):
    eps, de, dde = profile(omega, og)
    ei = O.sliceinv_3x3(eps)
    solver = omx.HIPSolver(precision="fp64")
    km, evs = omx.solve_k(omega, ei, g, solver, nev=nev, k_tol=1e-12, eig_tol=1e-11)
    return solver, km, evs, ei, de, dde


@pytest.mark.parametrize("profile,shape", [(O.gaussian_wg_eps_dispersive, (16, 16)),
                                           (aniso_dispersive, (24, 16)),
                                           (aniso_dispersive, (64, 64))],
                         ids=["isotropic-16x16", "aniso-24x16", "aniso-64x64"])
def test_ng_gvd_matches_oracle(profile, shape):
    """n_g, GVD and E of every mode against the oracle's restatement of ng_gvd_E evaluated on
    the same (k, ev); the GPU takes the derivative forwards, the oracle backwards."""
    print(f"ng_gvd vs oracle {profile.__name__} {shape}", flush=True)
    og = O.Grid(6.0, 4.0, *shape)
    g = omx.Grid(6.0, 4.0, *shape)
    solver, km, evs, ei, de, dde = _gpu_mode(profile, W0, g, og, 2)
    ng, gvd, Es = omx.ng_gvd(W0, km, evs, ei, de, dde, g, solver, return_E=True)
    for j in range(2):
        # GMRES on the oracle's undeflated system stalls near 1e-13 with an eigenvector from
        # another solver (residual ~1e-11); 1e-10 moves GVD by ~1e-11 relative
        rng_, rgvd, rE = O.ng_gvd(W0, km[j], evs[j], ei, de, dde, og, adj_rtol=1e-10, return_E=True)
        print(f"  mode {j}: ng {ng[j]:.15f} ({rng_:.15f}) gvd {gvd[j]:.12e} ({rgvd:.12e})", flush=True)
        assert ng[j] == pytest.approx(rng_, rel=1e-12)
        assert gvd[j] == pytest.approx(rgvd, rel=1e-9)
        assert np.allclose(Es[j], rE, rtol=0, atol=1e-12 * np.abs(rE).max())
        # the n_g of ng_gvd is the n_g of group_index
        assert ng[j] == pytest.approx(omx.group_index(km[j], evs[j], W0, ei, de, g, solver), rel=1e-12)
    # single-mode call form
    assert omx.ng_gvd(W0, km[1], evs[1], ei, de, dde, g, solver) == pytest.approx((ng[1], gvd[1]), rel=1e-12)


def test_ng_gvd_matches_fd_of_gpu_resolves_128():
    """At 128 x 128 (register-FFT operator): GVD = d n_g / d omega against a five-point
    finite difference of n_g from full GPU re-solves of the dispersive guide."""
    print("ng_gvd vs FD 128x128", flush=True)
    og = O.Grid(6.0, 4.0, 128, 128)
    g = omx.Grid(6.0, 4.0, 128, 128)

    def mode(w):
        return _gpu_mode(O.gaussian_wg_eps_dispersive, w, g, og, 1)

    solver, km, evs, ei, de, dde = mode(W0)
    ng, gvd = omx.ng_gvd(W0, km[0], evs[0], ei, de, dde, g, solver)
    h = 2e-3
    ngs, ks = [], []
    for s in (-2, -1, 1, 2):
        sv, k_, e_, ei_, de_, _ = mode(W0 + s * h)
        ks.append(k_[0])
        ngs.append(omx.group_index(k_[0], e_[0], W0 + s * h, ei_, de_, g, sv))
    fd = lambda v: (v[0] - 8 * v[1] + 8 * v[2] - v[3]) / (12 * h)  # noqa: E731
    print(f"  ng {ng} fd {fd(ks)}; gvd {gvd} fd {fd(ngs)}", flush=True)
    assert ng == pytest.approx(fd(ks), rel=1e-8)
    assert gvd == pytest.approx(fd(ngs), rel=1e-6)


def test_ng_gvd_rejects_bad_input():
    og = O.Grid(6.0, 4.0, 16, 16)
    g = omx.Grid(6.0, 4.0, 16, 16)
    solver, km, evs, ei, de, dde = _gpu_mode(O.gaussian_wg_eps_dispersive, W0, g, og, 1)
    bad = de.copy()
    bad[0, 1] += 1.0  # not symmetric
    with pytest.raises(omx.OmxError):
        omx.ng_gvd(W0, km[0], evs[0], ei, bad, dde, g, solver)
    with pytest.raises(omx.OmxError):
        omx.ng_gvd(-W0, km[0], evs[0], ei, de, dde, g, solver)
