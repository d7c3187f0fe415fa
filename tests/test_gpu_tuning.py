"""The autotuner (omx_autotune, the counterpart of FFTW PATIENT planning in the reference's
ModeSolver, src/maxwell.jl:940-943): every plan it can pick computes the same operator as
the oracle, and a tuned handle solves to the same k as an untuned one."""
import numpy as np
import pytest

import oracle as O
import omx
from omx import tuning

pytestmark = pytest.mark.gpu
P_WG = [4.2, 2.1, 1.0, 0.6]


def _relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("size", [(6.0, 4.0, 192, 192), (8.0, 3.0, 256, 64), (6.0, 4.0, 128, 96)],
                         ids=["192x192", "256x64", "128x96"])
def test_tuned_plans_match_oracle(size):
    """Stockham plans with the radix-4 variant and other tile shapes (what the tuner may
    choose) against the oracle's operator, FP32 tolerance."""
    Dx, Dy, nx, ny = size
    print(f"tuned plans {nx}x{ny}", flush=True)
    og, g = O.Grid(Dx, Dy, nx, ny), omx.Grid(Dx, Dy, nx, ny)
    ei = O.gaussian_wg_epsi(P_WG, og)
    rng = np.random.default_rng(nx + ny)
    H = rng.standard_normal((2 * og.N, 3)) + 1j * rng.standard_normal((2 * og.N, 3))
    ref = O.HelmholtzOracle(1.3, ei, og)
    M_ref, P_ref = ref.M(H), ref.P(H)
    solver = omx.HIPSolver(reg_ops=False)
    h = solver.handle(g, ei)
    L = omx._lib.lib()
    tried = 0
    for variant in (0, 1):
        for rows in (8, 4, 2, 1):
            for cols in (16, 8, 4, 1):
                if ny % rows or nx % cols or 3 * rows * nx > 3072 or 3 * cols * ny > 3072:
                    continue
                for key, v in (("fft_variant32", variant), ("tile_rows32", rows), ("tile_cols32", cols)):
                    assert L.omx_set_option(h.h, key.encode(), float(v)) == 0, L.omx_last_error(h.h)
                Mg = omx.HelmholtzMap(1.3, ei, g, solver, precision=omx.OMX_PREC_FP32)
                assert _relerr(Mg * H, M_ref) < 2e-6, (variant, rows, cols)
                assert _relerr(Mg.precondition(H), P_ref) < 2e-6, (variant, rows, cols)
                tried += 1
    assert tried >= 4


@pytest.mark.parametrize("nx,nev", [(192, 4), (256, 8)])
def test_autotune_then_solve(nx, nev, tmp_path):
    print(f"autotune {nx} nev {nev}", flush=True)
    og, g = O.Grid(8.0, 4.0, nx, nx), omx.Grid(8.0, 4.0, nx, nx)
    ei = O.gaussian_wg_epsi([4.2, 2.1, 1.6, 0.6], og)
    w = 1 / 1.55
    plain = omx.HIPSolver()
    k0, _ = omx.solve_k(w, ei, g, plain, nev=nev, k_tol=1e-10, eig_tol=1e-8, fetch_evecs=False)
    tuned = omx.HIPSolver()
    path = str(tmp_path / "wisdom.json")
    c = tuning.autotune(tuned, g, ei, nev, wisdom_path=path)
    print("choice", c, flush=True)
    assert not c["cached"] and c["op_us"] > 0 and c["gram_us"] > 0 and c["combine_us"] > 0
    assert c["gram_blocks"] in (128, 256, 512, 1024, 2048, 4096)
    assert c["combine_blocks"] in (256, 512, 1024, 2048, 4096, 8192)
    if nx == 192:
        assert c["reg_ops"] == 0 or c["reg_ops"] == 1  # no register path at 192: the value is unused
    k1, _ = omx.solve_k(w, ei, g, tuned, nev=nev, k_tol=1e-10, eig_tol=1e-8, fetch_evecs=False)
    assert np.allclose(k1, k0, rtol=1e-9, atol=0)
    # a second tuner call finds the stored choice and applies it to a fresh handle
    again = omx.HIPSolver()
    c2 = tuning.autotune(again, g, ei, nev, wisdom_path=path)
    assert c2["cached"] and all(c2[k] == c[k] for k in tuning.KEYS)
    k2, _ = omx.solve_k(w, ei, g, again, nev=nev, k_tol=1e-10, eig_tol=1e-8, fetch_evecs=False)
    assert np.array_equal(k2, k1)
