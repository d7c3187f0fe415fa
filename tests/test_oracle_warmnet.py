"""The warm-start network's CPU restatement pinned by torch's CPU autograd (an independent
implementation of the same layers) and by central finite differences of the loss."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import warmnet as W

HW, NEV, B = 8, 2, 3


def _case(seed=0, hw=HW, nev=NEV, b=B):
    rng = np.random.default_rng(seed)
    n = W.layout(hw, nev)[3]
    p = rng.standard_normal(n) * 0.2
    x = rng.standard_normal((b, W.NIN, hw, hw))
    tf = rng.standard_normal((b, 4 * nev, hw, hw))
    tk = rng.standard_normal((b, nev))
    return p, x, tf, tk


def _torch_loss(p, x, tf, tk, wk, hw, nev):
    convs, dw, db, n = W.layout(hw, nev)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    X = torch.tensor(x, dtype=torch.float64)

    def cv(i, t):
        ci, co, res, w, b = convs[i]
        return F.conv2d(t, P[w:b].reshape(co, ci, 3, 3), P[b:b + co], padding=1)

    a1 = F.relu(cv(0, X))
    a2 = F.relu(cv(1, F.avg_pool2d(a1, 2)))
    a3 = F.relu(cv(2, F.avg_pool2d(a2, 2)))
    k = a3.mean(dim=(2, 3)) @ P[dw:db].reshape(nev, 32).T + P[db:n]
    a4 = F.relu(cv(3, torch.cat([F.interpolate(a3, scale_factor=2, mode="nearest"), a2], 1)))
    a5 = F.relu(cv(4, torch.cat([F.interpolate(a4, scale_factor=2, mode="nearest"), a1], 1)))
    y = cv(5, a5)
    loss = ((y - torch.tensor(tf)) ** 2).mean() + wk * ((k - torch.tensor(tk)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), P.grad.numpy(), y.detach().numpy(), k.detach().numpy()


@pytest.mark.parametrize("hw,nev", [(8, 2), (16, 3)])
def test_forward_and_gradient_match_torch_autograd(hw, nev):
    p, x, tf, tk = _case(1, hw, nev)
    wk = 0.7
    L, g = W.loss_grad(p, x, tf, tk, wk, hw, nev)
    Lt, gt, yt, kt = _torch_loss(p, x, tf, tk, wk, hw, nev)
    y, k = W.forward(p, x, hw, nev)
    assert np.allclose(y, yt, rtol=1e-12, atol=1e-12) and np.allclose(k, kt, rtol=1e-12, atol=1e-12)
    assert abs(L - Lt) < 1e-12 * abs(Lt)
    assert np.max(np.abs(g - gt)) < 1e-11 * np.max(np.abs(gt))


def test_gradient_matches_finite_differences():
    p, x, tf, tk = _case(2)
    wk = 1.3
    _, g = W.loss_grad(p, x, tf, tk, wk, HW, NEV)
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(4):
        d = rng.standard_normal(p.size)
        fd = (W.loss_grad(p + h * d, x, tf, tk, wk, HW, NEV)[0] - W.loss_grad(p - h * d, x, tf, tk, wk, HW, NEV)[0]) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * abs(fd)


def test_layout_counts():
    convs, dw, db, n = W.layout(64, 8)
    assert [c[:2] for c in convs] == [(7, 16), (16, 32), (32, 32), (64, 32), (48, 32), (32, 32)]
    assert n == sum(ci * co * 9 + co for ci, co, *_ in convs) + 8 * 32 + 8
