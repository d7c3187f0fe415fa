"""GPU EME slices (omx_build_modes, omx_overlap_matrix, previous-solve warm start) against
the oracle's restatement of modes.jl:95-170, and the EME chain on coupler cross-sections
(the reference's checks, lib/EigenmodeExpansion/test/runtests.jl:66-103 and :213-227)."""
import numpy as np
import pytest

import oracle as O
from oracle.eme import build_mode as o_build_mode, overlap_matrix as o_overlap
import omx
from omx import eme as X
from omx import geometry as GEO

pytestmark = pytest.mark.gpu
W0 = 1 / 1.55


def _coupler(g, gap, w1=0.5, w2=0.5, h=0.3, backend="gpu"):
    return GEO.coupler_slice(g, gap, w1, w2, h=h, backend=backend)


@pytest.mark.parametrize("shape", [(32, 24), (64, 64)])
def test_modes_match_oracle(shape):
    print(f"eme modes vs oracle {shape}", flush=True)
    og = O.Grid(6.0, 4.0, *shape)
    g = omx.Grid(6.0, 4.0, *shape)
    ei = O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], og)
    solver = omx.HIPSolver(precision="fp64")
    km, evs = omx.solve_k(W0, ei, g, solver, nev=3, k_tol=1e-11, eig_tol=1e-10)
    for conj in (False, True):
        norms = X.build_modes(km, None, ei, g, solver, slot=1, conjugate=conj)   # device eigenvectors
        E, H = X.mode_fields(solver, 1, 3, g)
        ref = [o_build_mode(k, v, ei, og, conjugate=conj) for k, v in zip(km, evs)]
        for j in range(3):
            scale = np.abs(ref[j][0]).max()
            assert np.abs(E[j] - ref[j][0]).max() <= 1e-12 * scale
            assert np.abs(H[j] - ref[j][1]).max() <= 1e-12 * np.abs(ref[j][1]).max()
            assert norms[j] == pytest.approx(ref[j][2], rel=1e-12)
        Om = X.overlap_matrix(solver, 1, 1, 3, 3, conj)
        Oref = o_overlap([(r[0], r[1]) for r in ref], [(r[0], r[1]) for r in ref], og.dV(), conj)
        assert np.abs(Om - Oref).max() <= 1e-12
        if not conj:
            assert np.abs(Om - np.eye(3)).max() < 1e-7   # normalised and mutually orthogonal
    # host eigenvectors give the same mode set
    X.build_modes(km, evs, ei, g, solver, slot=2)
    X.build_modes(km, None, ei, g, solver, slot=3)
    assert np.allclose(X.mode_fields(solver, 2, 3, g)[0][2], X.mode_fields(solver, 3, 3, g)[0][2], rtol=0, atol=1e-14)


def test_cross_slice_overlaps_match_oracle():
    print("eme cross overlaps", flush=True)
    og = O.Grid(6.0, 3.0, 64, 32)
    g = omx.Grid(6.0, 3.0, 64, 32)
    solver = omx.HIPSolver(precision="fp64")
    built = []
    for s, gap in enumerate((0.9, 0.8)):
        ei = _coupler(g, gap)
        km, evs = omx.solve_k(W0, ei, g, solver, nev=2, k_tol=1e-11, eig_tol=1e-10)
        X.build_modes(km, None, ei, g, solver, slot=s)
        built.append([o_build_mode(k, v, ei, og) for k, v in zip(km, evs)])
    for conj in (False, True):
        O_LR = X.overlap_matrix(solver, 0, 1, 2, 2, conj)
        O_RL = X.overlap_matrix(solver, 1, 0, 2, 2, conj)
        r_LR = o_overlap([(b[0], b[1]) for b in built[0]], [(b[0], b[1]) for b in built[1]], og.dV(), conj)
        r_RL = o_overlap([(b[0], b[1]) for b in built[1]], [(b[0], b[1]) for b in built[0]], og.dV(), conj)
        assert np.abs(O_LR - r_LR).max() <= 1e-12
        assert np.abs(O_RL - r_RL).max() <= 1e-12


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_warm_start_reproduces_cold_solution(precision):
    print(f"warm start {precision}", flush=True)
    g = omx.Grid(6.0, 3.0, 128, 64)
    e0, e1 = _coupler(g, 0.60), _coupler(g, 0.62)
    cold = omx.HIPSolver(precision=precision)
    kc, evc, ic = omx.solve_k(W0, e1, g, cold, nev=4, k_tol=1e-10, eig_tol=1e-9, return_info=True)
    warm = omx.HIPSolver(precision=precision, warm_start=True)
    omx.solve_k(W0, e0, g, warm, nev=4, k_tol=1e-10, eig_tol=1e-9)
    kw, evw, iw = omx.solve_k(W0, e1, g, warm, nev=4, k_tol=1e-10, eig_tol=1e-9, return_info=True)
    print(f"  cold {ic}\n  warm {iw}", flush=True)
    assert np.allclose(kw, kc, rtol=1e-9, atol=0)
    for a, b in zip(evw, evc):
        assert abs(np.vdot(a, b)) > 1 - 1e-7
    # the previous solution skips the shared first eigensolve and starts Newton close by
    assert iw["n_newton"] < ic["n_newton"]


def test_eme_chain_on_coupler():
    """Four cells of a tapered coupler: warm and cold chains give the same device S-matrix,
    transmission is passive and the supermodes carry power (runtests.jl:66-81, :213-227)."""
    print("eme chain", flush=True)
    g = omx.Grid(6.0, 3.0, 64, 32)
    gaps = np.linspace(1.2, 0.3, 4)
    cells = [(lambda gp=gp: _coupler(g, gp)) for gp in gaps]
    lengths = [2.5] * 4
    r_warm = X.eme(cells, lengths, W0, g, omx.HIPSolver(precision="fp64"), nev=2, warmstart=True, k_tol=1e-10,
                   eig_tol=1e-9)
    r_cold = X.eme(cells, lengths, W0, g, omx.HIPSolver(precision="fp64"), nev=2, warmstart=False, k_tol=1e-10,
                   eig_tol=1e-9)
    T = r_warm.S.transmission()
    print("  |T|", np.abs(T), "\n  iters warm", [s["lobpcg_iters_fp64"] for s in r_warm.stats],
          "cold", [s["lobpcg_iters_fp64"] for s in r_cold.stats], flush=True)
    assert T.shape == (2, 2)
    assert np.all(np.abs(T) <= 1 + 1e-6)
    assert X.power_coupling(r_warm, in_mode=1, out_mode=1) > 1e-3
    assert np.allclose(T, r_cold.S.transmission(), atol=1e-5)
    assert X.power_coupling(r_warm) == pytest.approx(X.power_coupling(r_cold), rel=1e-5)
    for km_w, km_c in zip(r_warm.kmags, r_cold.kmags):
        assert np.allclose(km_w, km_c, rtol=1e-9)


def test_eme_slices_at_scale_512x256_16_modes():
    """Three cross-sections of the adiabatic coupler at 512 x 256 with 16 modes per slice,
    mixed precision, previous-slice warm start and GPU smoothing: every slice's mode set is
    power-normalised and orthogonal, and neighbouring slices overlap near the identity."""
    print("eme 512x256 nev16", flush=True)
    g = omx.Grid(8.0, 4.0, 512, 256)
    svals = [0.0, 0.005, 0.01]
    cells = [(lambda s=s: GEO.coupler_slice(g, 0.3 + 1.2 * s, 0.8 + 0.4 * s, 1.2 - 0.4 * s, h=0.8, backend="gpu"))
             for s in svals]
    solver = omx.HIPSolver()
    res = X.eme(cells, [1.0] * 3, W0, g, solver, nev=16, warmstart=True, k_tol=1e-9, eig_tol=1e-8)
    for st in res.stats:
        print(f"  slice: {st['seconds']:.3f} s, newton {st['n_newton']}, iters fp32 {st['lobpcg_iters_fp32']} "
              f"fp64 {st['lobpcg_iters_fp64']}", flush=True)
    assert all(st["converged"] for st in res.stats)
    # the last slice's modes (slot 0 after three cells) are power-normalised and orthogonal
    O_self = X.overlap_matrix(solver, 0, 0, 16, 16)
    print("  max |O_self - I|", np.abs(O_self - np.eye(16)).max(), flush=True)
    assert np.abs(O_self - np.eye(16)).max() < 1e-5
    for O_LR, O_RL in res.overlaps:
        best = np.abs(O_LR).max(axis=1)
        print("  min_i max_j |O_LR|", best.min(), flush=True)
        assert best.min() > 0.9
    assert np.all(np.abs(res.S.transmission()) <= 1 + 1e-6)
