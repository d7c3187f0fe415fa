"""CPU restatement (NumPy, FP64) of the learned warm-start network: test infrastructure, the
reference point of the GPU forward/backward kernels in the tests and nothing else.

The network has no counterpart in the reference (OptiMode starts every eigensolve from random
vectors and ``k_guess``, ``src/MaxwellEigenmodes.jl:63``), so nothing in the reference pins
it.  This restatement is pinned instead by torch's CPU autograd on the same network and by
central finite differences of the loss (``tests/test_oracle_warmnet.py``).

Tensors are ``[batch][channel][y][x]``.  Layers (3x3 convolutions, zero padding):
``c1`` 7->16 @hw, ReLU; 2x2 mean pool; ``c2`` 16->32 @hw/2, ReLU; pool; ``c3`` 32->32 @hw/4,
ReLU; head: pixel mean of c3 -> dense -> nev; ``c4`` [up2(c3) | c2] 64->32 @hw/2, ReLU;
``c5`` [up2(c4) | c1] 48->32 @hw, ReLU; ``c6`` 32->4nev @hw, linear.  Flat parameters: per
convolution its weights ``(co, ci, 3, 3)`` then its biases, then the dense ``(nev, 32)`` and
its biases.  Loss: ``mean((y - t_f)^2) + wk mean((k - t_k)^2)``.
"""
import numpy as np

NIN = 7


def layout(hw, nev):
    """``[(ci, co, res, w_off, b_off)] * 6, dense_w_off, dense_b_off, nparams``."""
    dims = [(NIN, 16, hw), (16, 32, hw // 2), (32, 32, hw // 4), (64, 32, hw // 2), (48, 32, hw), (32, 4 * nev, hw)]
    convs, off = [], 0
    for ci, co, res in dims:
        w = off
        b = w + ci * co * 9
        off = b + co
        convs.append((ci, co, res, w, b))
    dw = off
    db = dw + nev * 32
    return convs, dw, db, db + nev


def _unpack(p, hw, nev):
    convs, dw, db, n = layout(hw, nev)
    assert p.size == n
    Ws = [(p[w:b].reshape(co, ci, 3, 3), p[b:b + co]) for ci, co, res, w, b in convs]
    return Ws, p[dw:db].reshape(nev, 32), p[db:n]


def conv(x, W, b):
    B, ci, H, Wd = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    out = np.zeros((B, W.shape[0], H, Wd))
    for dy in range(3):
        for dx in range(3):
            out += np.einsum("oi,bihw->bohw", W[:, :, dy, dx], xp[:, :, dy:dy + H, dx:dx + Wd])
    return out + b[None, :, None, None]


def conv_bwd(x, W, dout, need_dx=True):
    B, ci, H, Wd = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp)
    for dy in range(3):
        for dx in range(3):
            dW[:, :, dy, dx] = np.einsum("bohw,bihw->oi", dout, xp[:, :, dy:dy + H, dx:dx + Wd])
            if need_dx:
                dxp[:, :, dy:dy + H, dx:dx + Wd] += np.einsum("oi,bohw->bihw", W[:, :, dy, dx], dout)
    return dW, dout.sum(axis=(0, 2, 3)), (dxp[:, :, 1:-1, 1:-1] if need_dx else None)


def pool(x):
    B, C, H, W = x.shape
    return x.reshape(B, C, H // 2, 2, W // 2, 2).mean(axis=(3, 5))


def pool_bwd(d):
    return np.repeat(np.repeat(d, 2, axis=2), 2, axis=3) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def up_bwd(d):
    B, C, H, W = d.shape
    return d.reshape(B, C, H // 2, 2, W // 2, 2).sum(axis=(3, 5))


def forward(p, x, hw, nev, cache=False):
    """``(fields (B, 4nev, hw, hw), k (B, nev))``, plus the activations when ``cache``."""
    (c1, c2, c3, c4, c5, c6), Wd, bd = _unpack(np.asarray(p, dtype=np.float64), hw, nev)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.maximum(conv(x, *c1), 0)
    p1 = pool(a1)
    a2 = np.maximum(conv(p1, *c2), 0)
    p2 = pool(a2)
    a3 = np.maximum(conv(p2, *c3), 0)
    g = a3.mean(axis=(2, 3))
    k = g @ Wd.T + bd
    cat4 = np.concatenate([up(a3), a2], axis=1)
    a4 = np.maximum(conv(cat4, *c4), 0)
    cat5 = np.concatenate([up(a4), a1], axis=1)
    a5 = np.maximum(conv(cat5, *c5), 0)
    y = conv(a5, *c6)
    if cache:
        return y, k, dict(x=x, a1=a1, p1=p1, a2=a2, p2=p2, a3=a3, g=g, cat4=cat4, a4=a4, cat5=cat5, a5=a5)
    return y, k


def loss_grad(p, x, tf, tk, wk, hw, nev):
    """``(loss, dloss/dp)`` for one batch."""
    p = np.asarray(p, dtype=np.float64)
    (c1, c2, c3, c4, c5, c6), Wd, bd = _unpack(p, hw, nev)
    y, k, a = forward(p, x, hw, nev, cache=True)
    tf = np.asarray(tf, dtype=np.float64)
    tk = np.asarray(tk, dtype=np.float64)
    loss = float(np.mean((y - tf) ** 2) + wk * np.mean((k - tk) ** 2))
    dy = 2.0 * (y - tf) / y.size
    dk = 2.0 * wk * (k - tk) / k.size
    grads = [None] * 6
    dW6, db6, da5 = conv_bwd(a["a5"], c6[0], dy)
    grads[5] = (dW6, db6)
    da5 = da5 * (a["a5"] > 0)
    dW5, db5, dcat5 = conv_bwd(a["cat5"], c5[0], da5)
    grads[4] = (dW5, db5)
    da4 = up_bwd(dcat5[:, :32]) * (a["a4"] > 0)
    da1 = dcat5[:, 32:]
    dW4, db4, dcat4 = conv_bwd(a["cat4"], c4[0], da4)
    grads[3] = (dW4, db4)
    da3 = up_bwd(dcat4[:, :32])
    da2 = dcat4[:, 32:]
    dWd = dk.T @ a["g"]
    dbd = dk.sum(axis=0)
    dg = dk @ Wd
    r2 = a["a3"].shape[2] * a["a3"].shape[3]
    da3 = (da3 + dg[:, :, None, None] / r2) * (a["a3"] > 0)
    dW3, db3, dp2 = conv_bwd(a["p2"], c3[0], da3)
    grads[2] = (dW3, db3)
    da2 = (da2 + pool_bwd(dp2)) * (a["a2"] > 0)
    dW2, db2, dp1 = conv_bwd(a["p1"], c2[0], da2)
    grads[1] = (dW2, db2)
    da1 = (da1 + pool_bwd(dp1)) * (a["a1"] > 0)
    dW1, db1, _ = conv_bwd(a["x"], c1[0], da1, need_dx=False)
    grads[0] = (dW1, db1)
    out = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads] + [dWd.ravel(), dbd])
    return loss, out
