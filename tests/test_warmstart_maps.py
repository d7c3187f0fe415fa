"""Host-side maps of the learned warm start (no GPU): features and the start-vector
representation, pinned by closed forms."""
import numpy as np

import omx
from omx import warmstart as WS


def test_start_vectors_invert_field_targets_on_low_plane_waves():
    grid = omx.Grid(8.0, 4.0, 64, 32)
    hw = 16
    rng = np.random.default_rng(0)
    evs = []
    for _ in range(3):
        A = np.zeros((2, 64, 32), dtype=complex)
        sx, sy = WS._sel(64, hw), WS._sel(32, hw)
        A[:, sx[:, None], sy[None, :]] = rng.standard_normal((2, hw, hw)) + 1j * rng.standard_normal((2, hw, hw))
        v = A.ravel(order="F")
        evs.append(v / np.linalg.norm(v))
    back = WS.start_vectors(WS.field_targets(evs, grid, hw).astype(np.float64), grid, hw)
    for v, b in zip(evs, back):
        assert np.allclose(v, b, atol=1e-12)
    # a vector with energy above the kept plane waves comes back as its normalised projection
    A = np.zeros((2, 64, 32), dtype=complex)
    A[0, 0, 0], A[1, 20, 3] = 0.6, 0.8  # (20, 3) lies outside the 16 x 16 lowest waves
    t = WS.field_targets([A.ravel(order="F")], grid, hw)
    b = WS.start_vectors(t.astype(np.float64), grid, hw)[0].reshape((2, 64, 32), order="F")
    assert abs(b[0, 0, 0] - 1.0) < 1e-12 and np.count_nonzero(np.abs(b) > 1e-12) == 1


def test_features_block_means_and_omega():
    grid = omx.Grid(4.0, 4.0, 32, 32)
    e = np.zeros((3, 3, 32, 32))
    for a in range(3):
        e[a, a] = 0.55 + 0.25 * (a + 1)
    e[0, 1] = e[1, 0] = 0.05 * np.arange(32)[:, None] / 31
    f = WS.features(e, grid, 0.7, hw=8)
    assert f.shape == (7, 8, 8) and f.dtype == np.float32
    for a in range(3):
        assert np.allclose(f[a], a + 1)
    assert np.allclose(f[3][:, 0], (np.arange(32).reshape(8, 4).mean(1) / 31))
    assert np.allclose(f[6], (0.7 - WS.OMEGA0) / WS.OMEGA_SCALE)


def test_neff_targets_round_trip():
    k = np.array([1.3, 1.2, 1.05])
    t = WS.neff_targets(k, 0.65)
    assert np.allclose(0.65 * (WS.N_SHIFT + WS.N_SCALE * t.astype(float)), k, rtol=1e-6)


def test_lowres_vectors_are_the_kept_plane_waves_of_start_vectors():
    grid = omx.Grid(8.0, 4.0, 64, 32)
    hw = 16
    f = np.random.default_rng(1).standard_normal((12, hw, hw))
    full = WS.start_vectors(f, grid, hw)
    low = WS.lowres_vectors(f, hw)
    assert low.shape == (3, 2 * hw * hw)
    # entry c + 2(gx + hw gy) of a low vector sits at c + 2(ix + Nx iy) of the full one
    sx, sy = WS._sel(64, hw), WS._sel(32, hw)
    for c in range(2):
        for gx in (0, 3, hw - 1):
            for gy in (0, 5, hw - 2):
                assert np.allclose(low[:, c + 2 * (gx + hw * gy)], full[:, c + 2 * (sx[gx] + 64 * sy[gy])])
    assert np.allclose(np.linalg.norm(low, axis=1), 1.0)


def test_structure_families_draw_in_range_and_build_on_the_host():
    """Both training families: seeded draws are repeatable and inside their ranges, and each
    builds a symmetric ε⁻¹ of the grid's shape (host smoothing, no GPU)."""
    assert set(WS.FAMILIES) == {"rib", "coupler"}
    g = omx.Grid(8.0, 4.0, 64, 32)
    for fam, (draw, build) in WS.FAMILIES.items():
        a, b = draw(np.random.default_rng(7)), draw(np.random.default_rng(7))
        assert a == b
        assert 0.55 <= a["omega"] <= 0.75
        ei = build(g, a, backend="host")
        assert ei.shape == (3, 3, 64, 32)
        assert np.allclose(ei, ei.transpose(1, 0, 2, 3))
        assert np.all(np.linalg.eigvalsh(ei.transpose(2, 3, 0, 1)) > 0)
    c = WS.random_coupler(np.random.default_rng(3))
    assert 0.2 <= c["gap"] <= 1.7 and 0.7 <= c["w1"] <= 1.3 and 0.7 <= c["w2"] <= 1.3 and 0.7 <= c["h"] <= 0.9


def test_shipped_coupler_network_file():
    """The EME chain's network file: 16 modes on the 64 x 64 map, parameters finite (loaded
    with the non-executing npz reader; the GPU tests run it)."""
    import os
    path = os.path.join(WS.DATA_DIR, WS.COUPLER_FILE)
    with np.load(path, allow_pickle=False) as z:
        assert int(z["nev"]) == WS.COUPLER_NEV == 16 and int(z["hw"]) == 64
        assert np.all(np.isfinite(z["params"])) and z["params"].size > 10000


def test_align_targets_solves_the_procrustes_problem():
    """Outputs that are the targets rotated by a unitary W0 within each sample's band span (plus
    a little noise): the aligned targets recover the outputs to the noise level, the aligned
    loss never exceeds the unaligned one, and the rotation is unitary (norms kept)."""
    rng = np.random.default_rng(5)
    n, nev, hw = 3, 4, 8
    tf = rng.standard_normal((n, 4 * nev, hw, hw)).astype(np.float32)
    T = WS._complex_maps(tf, nev)
    Y = np.empty_like(T)
    for s in range(n):
        q, _ = np.linalg.qr(rng.standard_normal((nev, nev)) + 1j * rng.standard_normal((nev, nev)))
        Y[s] = np.einsum("jcab,jk->kcab", T[s], q)
    y = WS._channel_maps(Y) + 1e-4 * rng.standard_normal(tf.shape).astype(np.float32)
    al = WS.align_targets(y, tf, nev)
    assert np.abs(al - y).max() < 1e-2
    assert np.sum((al - y) ** 2) <= np.sum((tf - y) ** 2)
    assert np.allclose(np.sum(al.astype(np.float64) ** 2, axis=(1, 2, 3)), np.sum(tf.astype(np.float64) ** 2, axis=(1, 2, 3)),
                       rtol=1e-5)
    # round trip of the channel layout
    assert np.array_equal(WS._channel_maps(WS._complex_maps(tf, nev)), tf)
