"""The warm-start network on the GPU: forward and loss gradient against the CPU restatement
(itself pinned by torch autograd and finite differences), deterministic training, and the
learned start against the plain start on held-out rib geometries (same converged k, fewer
LOBPCG iterations)."""
import numpy as np
import pytest

import omx
from omx import warmstart as WS
from oracle import warmnet as W

pytestmark = pytest.mark.gpu


def _rand_case(hw, nev, B, seed):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(W.layout(hw, nev)[3]) * 0.2).astype(np.float32)
    x = rng.standard_normal((B, 7, hw, hw)).astype(np.float32)
    tf = rng.standard_normal((B, 4 * nev, hw, hw)).astype(np.float32)
    tk = rng.standard_normal((B, nev)).astype(np.float32)
    return p, x, tf, tk


@pytest.mark.parametrize("hw,nev,B", [(16, 3, 3), (32, 2, 5), (64, 8, 2)])
def test_forward_and_gradient_match_oracle(hw, nev, B):
    print(f"warmnet parity hw {hw} nev {nev} B {B}", flush=True)
    p, x, tf, tk = _rand_case(hw, nev, B, 1)
    net = WS.WarmNet(hw, nev)
    assert net.nparams == p.size
    net.params = p
    assert np.array_equal(net.params, p)
    y, k = net.forward(x)
    yr, kr = W.forward(p.astype(np.float64), x, hw, nev)
    assert np.max(np.abs(y - yr)) < 1e-5 * np.max(np.abs(yr))
    assert np.max(np.abs(k - kr)) < 1e-5 * np.max(np.abs(kr))
    wk = 0.5
    L, g = net.loss_grad(x, tf, tk, wk)
    Lr, gr = W.loss_grad(p.astype(np.float64), x, tf, tk, wk, hw, nev)
    assert abs(L - Lr) < 1e-5 * abs(Lr)
    assert np.max(np.abs(g - gr)) < 1e-4 * np.max(np.abs(gr))


@pytest.mark.parametrize("hw,nev,B", [(16, 3, 3), (32, 2, 5), (64, 8, 2), (64, 16, 1)])
def test_bf16_mfma_convolutions_match_oracle_to_bf16(hw, nev, B):
    """The network with its convolutions (forward and data gradient) in bf16 on the matrix
    cores, FP32 accumulation: outputs and loss gradient against the FP64 oracle to bf16
    accuracy (8 significand bits, six layers deep); a wrong operand layout or tap index
    would give O(1) errors."""
    print(f"warmnet bf16 parity hw {hw} nev {nev} B {B}", flush=True)
    p, x, tf, tk = _rand_case(hw, nev, B, 2)
    net = WS.WarmNet(hw, nev, bf16=True)
    net.params = p
    y, k = net.forward(x)
    yr, kr = W.forward(p.astype(np.float64), x, hw, nev)
    ey = np.max(np.abs(y - yr)) / np.max(np.abs(yr))
    ek = np.max(np.abs(k - kr)) / np.max(np.abs(kr))
    wk = 0.5
    L, g = net.loss_grad(x, tf, tk, wk)
    Lr, gr = W.loss_grad(p.astype(np.float64), x, tf, tk, wk, hw, nev)
    eg = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    print(f"  outputs {ey:.2e}, k {ek:.2e}, loss {abs(L - Lr) / abs(Lr):.2e}, gradient {eg:.2e} (cos {cos:.6f})",
          flush=True)
    assert ey < 3e-2 and ek < 3e-2
    assert abs(L - Lr) < 1e-2 * abs(Lr)
    assert eg < 5e-2 and cos > 0.998
    # the FP32 path of the same network is unchanged
    net.set_bf16(False)
    y32, _ = net.forward(x)
    assert np.max(np.abs(y32 - yr)) < 1e-5 * np.max(np.abs(yr))


def test_training_reduces_loss_and_is_deterministic():
    hw, nev = 16, 2
    rng = np.random.default_rng(4)
    teacher = (rng.standard_normal(W.layout(hw, nev)[3]) * 0.2)
    x = rng.standard_normal((48, 7, hw, hw)).astype(np.float32)
    tf, tk = W.forward(teacher, x, hw, nev)
    runs = []
    for _ in range(2):
        net = WS.WarmNet(hw, nev, seed=3)
        hist = net.train(x, tf.astype(np.float32), tk.astype(np.float32), wk=1.0, epochs=60, batch=8, lr=3e-3, seed=5)
        runs.append((hist, net.params))
    print(f"  loss {runs[0][0][0]:.4g} -> {runs[0][0][-1]:.4g}", flush=True)
    assert runs[0][0][-1] < 0.2 * runs[0][0][0]
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][0], runs[1][0])


def test_learned_start_matches_plain_and_saves_iterations():
    # rib draws at 128x128, 4 bands: train on 96 solver-converged draws, compare on 6 held out
    grid = omx.Grid(8.0, 4.0, 128, 128)
    nev, hw = 4, 32
    solver = omx.HIPSolver()
    X, TF, TK, _ = WS.generate(range(1000, 1096), grid, nev, solver, hw=hw)
    net = WS.WarmNet(hw, nev, seed=0)
    hist = net.train(X, TF, TK, wk=1.0, epochs=200, batch=16, lr=2e-3, seed=1)
    print(f"  training loss {hist[0]:.4g} -> {hist[-1]:.4g}", flush=True)
    assert hist[-1] < 0.5 * hist[0]
    rows = WS.compare(range(30000, 30006), grid, nev, WS.LearnedStart(net), solver, mode="seed")
    it_p = sum(r["it_plain"] for r in rows)
    it_l = sum(r["it_learned"] for r in rows)
    for r in rows:
        print(f"  seed {r['seed']}: it {r['it_plain']} -> {r['it_learned']}, newton {r['newton_plain']} -> "
              f"{r['newton_learned']}, ms {r['ms_plain']:.1f} -> {r['ms_learned']:.1f}", flush=True)
        assert r["conv_plain"] and r["conv_learned"]
        assert np.max(np.abs(r["k_learned"] - r["k_plain"]) / r["k_plain"]) < 1e-8
    print(f"  LOBPCG iterations plain {it_p}, learned {it_l}", flush=True)
    assert it_l < it_p


def test_lowres_guess_equals_full_vector_guess():
    # the same start block given as full vectors or by its lowest plane waves (scattered on
    # the GPU) gives the same solve, bit for bit
    grid = omx.Grid(8.0, 4.0, 64, 64)
    prm = WS.random_rib(np.random.default_rng(7))
    ei = WS.rib_epsinv(grid, prm)
    f = np.random.default_rng(8).standard_normal((12, 16, 16))
    full, low = WS.start_vectors(f, grid, 16), WS.lowres_vectors(f, 16)
    solver = omx.HIPSolver()
    kw = dict(nev=3, k_tol=1e-10, eig_tol=1e-8, return_info=True, fetch_evecs=False)
    for band in (False, True):
        kg = [1.2 * prm["omega"] * 2.0] if not band else list(prm["omega"] * np.array([2.0, 1.9, 1.8]))
        ka, _, ia = omx.solve_k(prm["omega"], ei, grid, solver, kguess=kg, Hguess=full, band_guess=band, **kw)
        kb, _, ib = omx.solve_k(prm["omega"], ei, grid, solver, kguess=kg, Hguess=low, band_guess=band, guess_hw=16,
                                **kw)
        assert np.array_equal(ka, kb) and ia["lobpcg_iters_fp32"] == ib["lobpcg_iters_fp32"]


def test_device_guess_matches_host_guess():
    # features, forward and inverse 64^2 DFT on the GPU (omx_warmnet_guess) against the host
    # maps of omx.warmstart (features, forward, numpy ifft2)
    grid = omx.Grid(8.0, 4.0, 256, 256)
    st = WS.LearnedStart.load(WS.DATA_DIR + "/warmnet_tfln256.npz")
    solver = omx.HIPSolver()
    for seed in (30001, 30004):
        prm = WS.random_rib(np.random.default_rng(seed))
        ei = WS.rib_epsinv(grid, prm)
        kd, lowd = WS.device_guess(st, ei, grid, prm["omega"], solver)
        y, kh = st.net.forward(WS.features(ei, grid, prm["omega"], st.hw)[None])
        lowh = WS.lowres_vectors(y[0], st.hw)
        assert np.max(np.abs(kd - kh[0])) < 1e-5
        assert np.max(np.abs(lowd - lowh)) < 1e-5 * np.max(np.abs(lowh))
        assert np.allclose(np.linalg.norm(lowd, axis=1), 1.0, atol=1e-12)


def test_solve_k_learned_on_device_matches_host_path():
    grid = omx.Grid(8.0, 4.0, 256, 256)
    st = WS.LearnedStart.load(WS.DATA_DIR + "/warmnet_tfln256.npz")
    solver = omx.HIPSolver()
    prm = WS.random_rib(np.random.default_rng(30006))
    ei = WS.rib_epsinv(grid, prm)
    kw = dict(nev=8, k_tol=1e-10, eig_tol=1e-8, return_info=True)
    kd, evd, idv = WS.solve_k_learned(prm["omega"], ei, grid, solver, st, on_device=True, **kw)
    kh, evh, ihv = WS.solve_k_learned(prm["omega"], ei, grid, solver, st, on_device=False, **kw)
    assert idv["converged"] and ihv["converged"]
    assert np.max(np.abs(kd - kh) / kh) < 1e-9
    for a, b in zip(evd, evh):
        assert abs(np.vdot(a, b)) > 1 - 1e-6
    with pytest.raises(omx.OmxError):  # more bands than the network predicts
        WS.solve_k_learned(prm["omega"], ei, grid, solver, st, nev=9, on_device=True)


def test_solve_k_learned_rejects_what_it_cannot_serve():
    st = WS.LearnedStart.load(WS.DATA_DIR + "/warmnet_tfln256.npz")
    solver = omx.HIPSolver()
    # a grid that is not a multiple of the network's 64 x 64 map
    g = omx.Grid(8.0, 4.0, 200, 192)
    ei = WS.rib_epsinv(g, WS.random_rib(np.random.default_rng(1)))
    with pytest.raises(omx.OmxError) as e:
        WS.solve_k_learned(0.65, ei, g, solver, st, nev=4)
    assert e.value.status == omx._lib.OMX_ERR_BAD_SHAPE
    # a 3D cell
    from oracle import periodic as P
    og3 = P.Grid3(2.4, 2.0, 0.45, 64, 64, 4)
    g3 = omx.Grid3(2.4, 2.0, 0.45, 64, 64, 4)
    ei3 = P.bragg_wg_epsi(og3, eps_core=4.0, eps_bg=2.1, w0=0.5, h0=0.4, mod=0.5)
    with pytest.raises(omx.OmxError) as e:
        WS.solve_k_learned(0.5, ei3, g3, solver, st, nev=2)
    assert e.value.status == omx._lib.OMX_ERR_UNSUPPORTED
