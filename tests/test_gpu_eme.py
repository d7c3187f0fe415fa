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


# ---------------------------------------------------------------------------------------
# pull-backs of the EME chain (the reverse mode the reference gets from Zygote, eme.jl:80-88)
# ---------------------------------------------------------------------------------------

from oracle.eme import (mode_fields_pullback as o_fields_pb, normalization_pullback as o_norm_pb,  # noqa: E402
                        overlap_pullback as o_overlap_pb)


@pytest.mark.parametrize("conj", [False, True])
def test_mode_set_pullbacks_match_oracle(conj):
    """omx_overlap_pullback + omx_mode_pullback on the oracle's own eigenvectors against the
    oracle's overlap, normalisation and mode-field pull-backs (k̄, ēv, direct ε̄⁻¹), for seeded
    overlap cotangents on both sides of an interface."""
    print(f"eme pullbacks vs oracle conj={conj}", flush=True)
    og, g = O.Grid(6.0, 4.0, 32, 24), omx.Grid(6.0, 4.0, 32, 24)
    eis = [O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], og), O.gaussian_wg_epsi([4.0, 2.1, 1.2, 0.55], og)]
    sol = [O.solve_k(W0, e, og, nev=3, k_tol=1e-12, eig_tol=1e-10, maxiter=300, seed=3) for e in eis]
    solver = omx.HIPSolver(precision="fp64")
    rng = np.random.default_rng(12)
    obLR = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    obRL = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    X.build_modes(sol[1][0], sol[1][1], eis[1], g, solver, slot=1, conjugate=conj)
    X.build_modes(sol[0][0], sol[0][1], eis[0], g, solver, slot=0, conjugate=conj)
    # cell 0 on the left: O_LR = overlap(0, 1) reaches its E, O_RL = overlap(1, 0) its H
    X.overlap_pullback(solver, 0, 1, obLR, conj, which=1)
    X.overlap_pullback(solver, 1, 0, obRL, conj, which=2)
    kb, evb, direct = X.mode_pullback(solver, 0, 3, g)
    dA = og.dV()
    b0 = [o_build_mode(k, v, eis[0], og, conjugate=conj) for k, v in zip(*sol[0])]
    b1 = [o_build_mode(k, v, eis[1], og, conjugate=conj) for k, v in zip(*sol[1])]
    m0, m1 = [(b[0], b[1]) for b in b0], [(b[0], b[1]) for b in b1]
    ca, _ = o_overlap_pb(m0, m1, dA, obLR, conj)
    _, cb = o_overlap_pb(m1, m0, dA, obRL, conj)
    ref_direct = np.zeros_like(eis[0])
    for j, (E, H, nrm) in enumerate(b0):
        Erb, Hrb = o_norm_pb(E, H, nrm, dA, ca[j][0] + cb[j][0], ca[j][1] + cb[j][1], conj)
        kr, er, dr = o_fields_pb(sol[0][0][j], sol[0][1][j], eis[0], og, Erb, Hrb)
        ref_direct += dr
        print(f"  mode {j}: kbar {kb[j]:.12g} oracle {kr:.12g}", flush=True)
        assert kb[j] == pytest.approx(kr, rel=1e-9, abs=1e-12 * abs(kr) + 1e-14)
        assert np.linalg.norm(evb[j] - er) <= 1e-9 * np.linalg.norm(er)
    assert np.abs(direct - ref_direct).max() <= 1e-9 * np.abs(ref_direct).max()


def test_eme_gradient_matches_finite_differences():
    """eme_gradient (the cascade's power coupling |T11|² through three asymmetric coupler
    cells, two modes each) against central differences of the GPU EME forward along symmetric
    ε⁻¹ directions of the middle cell (a diagonal and an off-diagonal entry) and in ω.  The
    cells are well conditioned (|T11|² ≈ 1.06, overlap matrices far from singular); with the
    symmetric coupler's three modes the overlaps are near singular, |T11|² reaches 130 without
    the passivity projection and central differences of the forward lose their digits (the
    oracle's own differences agree with its adjoint to 1.6% there)."""
    print("eme gradient vs fd", flush=True)
    g = omx.Grid(6.0, 4.0, 64, 48)
    eis = [_coupler(g, gap, w2=0.7) for gap in (0.35, 0.5, 0.7)]
    lengths = [1.0, 2.5, 1.0]
    kw = dict(nev=2, passivity="none", k_tol=1e-12, eig_tol=1e-10, warmstart=False)
    solver = omx.HIPSolver(precision="fp64")
    obj = X.power_coupling_objective(1, 1)
    val, grads, wbar, parts = X.eme_gradient(eis, lengths, W0, g, solver, obj, return_parts=True, **kw)

    def value(e, w=W0):
        S = X.eme(e, lengths, w, g, omx.HIPSolver(precision="fp64"), nev=2, passivity="none", k_tol=1e-12,
                  eig_tol=1e-10, warmstart=False).S
        return obj(S)[0]
    assert val == pytest.approx(value(eis), rel=1e-9)
    rng = np.random.default_rng(4)
    for (a, b) in ((1, 1), (0, 1)):
        r = rng.standard_normal((g.Nx, g.Ny)) * 0.1
        D = np.zeros_like(eis[1])
        D[a, b] = r
        D[b, a] = r
        h = 1e-3
        ep = [e.copy() for e in eis]
        em = [e.copy() for e in eis]
        ep[1] = ep[1] + h * D
        em[1] = em[1] - h * D
        fd = (value(ep) - value(em)) / (2 * h)
        # the solve_k rule mirrors off-diagonals, the field rule does not: the directional
        # derivative along the symmetric D is the field part's contraction with D plus one
        # entry of each mirrored pair of the solve_k part
        sk, direct = parts[1]
        pred = np.sum(direct * D) + np.sum(sk[a, b] * r)
        assert np.array_equal(grads[1], sk + direct)
        print(f"  ({a},{b}) pred {pred:.10g} fd {fd:.10g}", flush=True)
        assert pred == pytest.approx(fd, rel=1e-3, abs=1e-8)
    hw = 1e-4
    fdw = (value(eis, W0 + hw) - value(eis, W0 - hw)) / (2 * hw)
    print(f"  omega bar {wbar:.10g} fd {fdw:.10g}", flush=True)
    assert wbar == pytest.approx(fdw, rel=1e-3, abs=1e-8)


def test_eme_chain_learned_cold_start_matches_plain():
    """The shipped coupler network (512x256, 16 modes) starts the cold cells of an EME chain:
    the same modes and device S-matrix as the plain start (the start never changes the
    converged result), every cell converged, and the learned guess used (fewer FP32 LOBPCG
    iterations summed over the cold cells is printed, not asserted: it is a speed property)."""
    from omx import warmstart as WS
    print("eme chain learned cold start", flush=True)
    learned = WS.load_coupler()
    assert learned is not None and learned.nev >= 16
    g = omx.Grid(*WS.COUPLER_GRID)
    cells = [(lambda gp=gp: GEO.coupler_slice(g, gp, 1.0, 1.0, h=0.8, backend="gpu")) for gp in (0.6, 0.9)]
    kw = dict(nev=16, warmstart=False, k_tol=1e-10, eig_tol=1e-9)
    rp = X.eme(cells, [1.0, 1.0], W0, g, omx.HIPSolver(), **kw)
    rl = X.eme(cells, [1.0, 1.0], W0, g, omx.HIPSolver(), learned=learned, **kw)
    print("  fp32 iterations plain", [s["lobpcg_iters_fp32"] for s in rp.stats], "learned",
          [s["lobpcg_iters_fp32"] for s in rl.stats], flush=True)
    assert all(s["converged"] for s in rp.stats + rl.stats)
    for a, b in zip(rl.kmags, rp.kmags):
        assert np.allclose(a, b, rtol=1e-9, atol=0)
    assert np.allclose(np.abs(rl.S.transmission()), np.abs(rp.S.transmission()), atol=1e-6)
    with pytest.raises(ValueError):
        X.eme(cells, [1.0, 1.0], W0, g, omx.HIPSolver(), nev=17, learned=learned)
