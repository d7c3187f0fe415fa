"""GPU Kottke smoothing (omx_smooth_epsinv) against the host implementation it restates,
on the configuration geometries and on planar interfaces with closed forms."""
import os
import time

import numpy as np
import pytest

from omx import geometry as G

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class _Grid:
    def __init__(self, Dx, Dy, Nx, Ny):
        self.Dx, self.Dy, self.Nx, self.Ny = Dx, Dy, Nx, Ny


def _fixture_grid(name):
    f = np.load(os.path.join(HERE, name + ".npz"))
    return _Grid(float(f["Dx"]), float(f["Dy"]), int(f["nx"]), int(f["ny"]))


def _close(a, b):
    assert a.shape == b.shape
    err = np.abs(a - b).max() / np.abs(b).max()
    assert err <= 1e-13, err


CASES = [
    ("rect_c1", lambda g, be, ns: G.rect_waveguide(g, nsub=ns, backend=be), "c1_rect_128"),
    ("tfln_c2", lambda g, be, ns: G.tfln_rib(g, nsub=ns, backend=be), "c2_tfln_256"),
    ("coupler", lambda g, be, ns: G.coupler_slice(g, 0.35, 1.1, 0.9, nsub=ns, backend=be), "c1_rect_128"),
    ("ridge_aniso", lambda g, be, ns: G.ridge_waveguide(g, 1.3, 0.6, 70.0, G.diag_eps(*G.EPS_LN), nsub=ns,
                                                        backend=be), "c2_tfln_256"),
]


@pytest.mark.parametrize("nsub", [8, 3])
@pytest.mark.parametrize("name,builder,grid", CASES, ids=[c[0] for c in CASES])
def test_gpu_smoothing_matches_host(name, builder, grid, nsub):
    print(f"smooth {name} nsub={nsub}", flush=True)
    g = _fixture_grid(grid)
    t0 = time.perf_counter()
    host = builder(g, "host", nsub)
    t1 = time.perf_counter()
    gpu = builder(g, "gpu", nsub)
    t2 = time.perf_counter()
    print(f"  {g.Nx}x{g.Ny}: host {1e3 * (t1 - t0):.1f} ms, gpu call {1e3 * (t2 - t1):.1f} ms", flush=True)
    _close(gpu, host)


def test_gpu_smoothing_closed_forms():
    print("smooth closed forms", flush=True)
    g = _Grid(4.0, 4.0, 16, 16)
    d = 0.25
    yc, xc = -2.0 + 9 * d, -2.0 + 11 * d
    core = G.Polygon([(-10.0, -10.0), (xc, -10.0), (xc, yc), (-10.0, yc)], np.diag([5.0, 4.0, 4.5]))
    eps = np.linalg.inv(np.moveaxis(G.smooth_epsinv([core], 3.0, g, nsub=8, backend="gpu"), (0, 1), (-2, -1)))
    assert eps[4, 9][1, 1] == pytest.approx(1 / (0.5 / 4.0 + 0.5 / 3.0), rel=1e-12)
    assert eps[4, 9][0, 0] == pytest.approx(4.0, rel=1e-12)
    assert eps[11, 3][0, 0] == pytest.approx(1 / (0.5 / 5.0 + 0.5 / 3.0), rel=1e-12)
    c = -2.0 + 7 * d
    tri = G.Polygon([(c - 20, c - 20), (c + 20, c - 20), (c + 20, c + 20)], 4.0)
    eps = np.linalg.inv(np.moveaxis(G.smooth_epsinv([tri], 2.0, g, nsub=8, backend="gpu"), (0, 1), (-2, -1)))
    f = 36 / 64
    n = np.array([1.0, -1.0, 0.0]) / np.sqrt(2)
    nn = np.outer(n, n)
    ref = (f * 4.0 + (1 - f) * 2.0) * (np.eye(3) - nn) + 1 / (f / 4.0 + (1 - f) / 2.0) * nn
    assert np.allclose(eps[7, 7], ref, rtol=1e-12, atol=1e-13)


def test_gpu_smoothing_many_shapes_and_precedence():
    """15 overlapping anisotropic rectangles (the maximum): earlier shapes win where they
    overlap, including pixels where three or more materials meet."""
    print("smooth 15 shapes", flush=True)
    g = _Grid(6.0, 4.0, 96, 64)
    rng = np.random.default_rng(11)
    shapes = []
    for s in range(15):
        B = rng.standard_normal((3, 3))
        e = 0.2 * (B @ B.T) + (1.5 + s * 0.2) * np.eye(3)
        e[0, 2] = e[2, 0] = e[1, 2] = e[2, 1] = 0.0  # in-plane / z block structure
        shapes.append(G.Rect(rng.uniform(-2, 2), rng.uniform(-1.5, 1.5), rng.uniform(0.3, 2.0),
                             rng.uniform(0.3, 1.5), e))
    host = G.smooth_epsinv(shapes, 1.0, g, nsub=6)
    gpu = G.smooth_epsinv(shapes, 1.0, g, nsub=6, backend="gpu")
    _close(gpu, host)
