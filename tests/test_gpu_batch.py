"""omx_solve_k_batch: problems of uneven grid, ε⁻¹, ω and nev, run as launch groups on one
stream (every band of every problem of a group an instance of the same launches), converge
to the results of individual omx_solve_k calls, and a failing problem does not stop the
others."""
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
                                   (64, 64, 0.8, 2, "gauss"), (128, 64, W0, 4, "coupler"), (192, 128, 0.6, 5, "rect"),
                                   (64, 64, 0.75, 4, "gauss2"), (128, 128, 0.7, 3, "rect")]:
        g = omx.Grid(6.0, 4.0, nx, ny)
        if kind == "gauss":
            ei = O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], O.Grid(6.0, 4.0, nx, ny))
        elif kind == "gauss2":
            ei = O.gaussian_wg_epsi([3.9, 2.0, 1.3, 0.7], O.Grid(6.0, 4.0, nx, ny))
        elif kind == "rect":
            ei = GEO.rect_waveguide(g, backend="gpu")
        else:
            ei = GEO.coupler_slice(g, 0.5, 0.9, 1.1, h=0.8, backend="gpu")
        out.append(dict(grid=g, epsinv=ei, omega=w, nev=nev, k_tol=1e-11, eig_tol=1e-9))
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_matches_individual_solves(precision):
    """Grouped solves (three 64x64 problems with different ε⁻¹, ω and nev share one group)
    against individual solves: k to 1e-10 relative, eigenvectors |<a, b>| >= 1 - 1e-9."""
    print(f"batch vs individual ({precision})", flush=True)
    probs = _problems()
    res = omx.solve_k_batch(probs, precision=precision, fetch_evecs=True)
    for p, r in zip(probs, res):
        assert r["status"] == 0 and r["info"]["converged"] == 1
        km, evs = omx.solve_k(p["omega"], p["epsinv"], p["grid"], omx.HIPSolver(precision=precision), nev=p["nev"],
                              k_tol=p["k_tol"], eig_tol=p["eig_tol"])
        print(f"  {p['grid'].Nx}x{p['grid'].Ny} nev {p['nev']}: k {r['kmags']} rel {np.abs(r['kmags'] / km - 1).max():.2e}",
              flush=True)
        assert np.allclose(r["kmags"], km, rtol=1e-10, atol=0)
        for a, b in zip(r["evecs"], evs):
            assert abs(abs(np.vdot(a, b)) - 1) < 1e-9
            assert abs(np.linalg.norm(a) - 1) < 1e-12


def test_batch_groups_of_one_and_of_many_agree():
    print("batch group sizes", flush=True)
    probs = [p for p in _problems() if (p["grid"].Nx, p["grid"].Ny) in ((64, 64), (128, 128))]
    many = omx.solve_k_batch(probs)
    one = omx.solve_k_batch(probs, max_group=1)
    for a, b in zip(many, one):
        assert a["status"] == 0 and b["status"] == 0
        assert np.allclose(a["kmags"], b["kmags"], rtol=1e-10, atol=0)


def test_batch_reports_per_problem_errors():
    print("batch errors", flush=True)
    probs = _problems()[:3]
    g7 = omx.Grid(6.0, 4.0, 7, 64)  # a prime factor 7: unsupported grid
    probs.insert(1, dict(grid=g7, epsinv=np.tile(np.eye(3)[:, :, None, None] * 0.5, (1, 1, 7, 64)), omega=W0, nev=2))
    res = omx.solve_k_batch(probs)
    assert res[1]["status"] == omx._lib.OMX_ERR_BAD_SHAPE
    assert all(res[j]["status"] == 0 for j in (0, 2, 3))
