"""omx_solve_k_batch: problems of uneven grid, ω and nev solved concurrently give exactly the
results of individual omx_solve_k calls, and a failing problem does not stop the others."""
import numpy as np
import pytest

import oracle as O
import omx
from omx import geometry as GEO

pytestmark = pytest.mark.gpu
W0 = 1 / 1.55


def _problems():
    out = []
    for (nx, ny, w, nev, kind) in [(64, 64, W0, 3, "gauss"), (96, 64, 0.7, 2, "gauss"), (128, 128, W0, 4, "rect"),
                                   (64, 64, 0.8, 2, "gauss"), (128, 64, W0, 4, "coupler"), (192, 128, 0.6, 5, "rect")]:
        g = omx.Grid(6.0, 4.0, nx, ny)
        if kind == "gauss":
            ei = O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], O.Grid(6.0, 4.0, nx, ny))
        elif kind == "rect":
            ei = GEO.rect_waveguide(g, backend="gpu")
        else:
            ei = GEO.coupler_slice(g, 0.5, 0.9, 1.1, h=0.8, backend="gpu")
        out.append(dict(grid=g, epsinv=ei, omega=w, nev=nev, k_tol=1e-10, eig_tol=1e-9))
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_matches_individual_solves(precision):
    print(f"batch vs individual ({precision})", flush=True)
    probs = _problems()
    res = omx.solve_k_batch(probs, nstreams=3, precision=precision, fetch_evecs=True)
    for p, r in zip(probs, res):
        assert r["status"] == 0 and r["info"]["converged"] == 1
        km, evs = omx.solve_k(p["omega"], p["epsinv"], p["grid"], omx.HIPSolver(precision=precision), nev=p["nev"],
                              k_tol=p["k_tol"], eig_tol=p["eig_tol"])
        print(f"  {p['grid'].Nx}x{p['grid'].Ny} nev {p['nev']}: k {r['kmags']}", flush=True)
        assert np.allclose(r["kmags"], km, rtol=1e-13, atol=0)
        for a, b in zip(r["evecs"], evs):
            assert abs(np.vdot(a, b)) > 1 - 1e-12


def test_batch_reports_per_problem_errors():
    print("batch errors", flush=True)
    probs = _problems()[:3]
    g7 = omx.Grid(6.0, 4.0, 7, 64)  # a prime factor 7: unsupported grid
    probs.insert(1, dict(grid=g7, epsinv=np.tile(np.eye(3)[:, :, None, None] * 0.5, (1, 1, 7, 64)), omega=W0, nev=2))
    res = omx.solve_k_batch(probs, nstreams=2)
    assert res[1]["status"] == omx._lib.OMX_ERR_BAD_SHAPE
    assert all(res[j]["status"] == 0 for j in (0, 2, 3))
