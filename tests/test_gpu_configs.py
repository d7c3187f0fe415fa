"""GPU solve_k on the configuration geometries against the FP64 oracle's golden k values:
every mode's ω² relative error (through the oracle's slope dω²/dk) must be <= 1e-8."""
import os

import numpy as np
import pytest

import omx
from omx import geometry as G

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = [("c1_rect_128", lambda g: G.rect_waveguide(g)), ("c2_tfln_256", lambda g: G.tfln_rib(g))]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("name,builder", CASES, ids=[c[0] for c in CASES])
def test_solve_k_matches_golden(name, builder, precision):
    print(f"golden {name} {precision}", flush=True)
    f = np.load(os.path.join(HERE, name + ".npz"))
    g = omx.Grid(float(f["Dx"]), float(f["Dy"]), int(f["nx"]), int(f["ny"]))
    ei = builder(g)
    w = float(f["omega"])
    nev = int(f["nev"])
    solver = omx.HIPSolver(precision=precision)
    km, evs, info = omx.solve_k(w, ei, g, solver, nev=nev, k_tol=1e-10, eig_tol=1e-8, return_info=True)
    rel = np.abs(f["dw2dk"] * (km - f["kmags"])) / w ** 2
    print(name, precision, "max rel err", rel.max(), info, flush=True)
    assert np.all(rel <= 1e-8), rel
    for v in evs:
        assert abs(np.linalg.norm(v) - 1) < 1e-10
