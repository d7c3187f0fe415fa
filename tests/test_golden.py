"""The committed golden fixtures against the oracle that wrote them (CPU), and the
structural invariants they must satisfy."""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from omx import geometry as G

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(HERE, name + ".npz"))


def test_small_fixture_operator_vectors():
    f = _load("small_gauss_24x16")
    g = O.Grid(float(f["Dx"]), float(f["Dy"]), int(f["nx"]), int(f["ny"]))
    ei = O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], g)
    op = O.HelmholtzOracle(float(f["k_apply"]), ei, g)
    assert np.allclose(op.M(f["X"]), f["MX"], rtol=0, atol=1e-12 * np.abs(f["MX"]).max())
    assert np.allclose(op.P(f["X"]), f["PX"], rtol=0, atol=1e-12 * np.abs(f["PX"]).max())
    assert np.allclose([op.HMkH(f["X"][:, j]) for j in range(3)], f["HMkH_X"], rtol=1e-12)


def test_small_fixture_modes_are_eigenpairs():
    f = _load("small_gauss_24x16")
    g = O.Grid(float(f["Dx"]), float(f["Dy"]), int(f["nx"]), int(f["ny"]))
    ei = O.gaussian_wg_epsi([4.2, 2.1, 1.0, 0.6], g)
    w = float(f["omega"])
    for k, v in zip(f["kmags"], f["evecs"]):
        assert np.linalg.norm(O.apply_M(v, k, ei, g) - w ** 2 * v) < 1e-9
        # phase canonicalisation is idempotent
        assert np.allclose(O.canonicalize_phase(v, k, ei, g), v, atol=1e-12)


@pytest.mark.parametrize("name,builder", [("c1_rect_128", lambda g: G.rect_waveguide(g)),
                                          ("c2_tfln_256", lambda g: G.tfln_rib(g))])
def test_config_fixtures_consistent(name, builder):
    f = _load(name)
    g = O.Grid(float(f["Dx"]), float(f["Dy"]), int(f["nx"]), int(f["ny"]))
    ei = builder(g)
    sha = hashlib.sha256(np.ascontiguousarray(ei, dtype=np.float64).tobytes()).hexdigest()
    assert sha == str(f["epsinv_sha256"])
    w = float(f["omega"])
    km = f["kmags"]
    assert np.all(np.diff(km) < 0)                        # bands in order of decreasing k
    eps_max = max(np.linalg.eigvalsh(np.linalg.inv(np.moveaxis(ei, (0, 1), (-2, -1)))).max(), 1.0)
    assert np.all(km / w < np.sqrt(eps_max)) and np.all(km / w > 1.0)
    assert np.all(f["resid"] < 1e-9)
    assert np.all(f["dw2dk"] > 0)
    # n_g = d k / d omega from the same fixture, consistent with k: n_g > n_eff for guided modes
    assert np.all(f["ng"] > km / w)
