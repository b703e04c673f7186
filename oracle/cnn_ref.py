"""FP64 numpy restatement of the residual CNN of the learned correction
(forward, backward, loss) -- TEST INFRASTRUCTURE ONLY: the reference of the HIP
kernels' gradients.  The network has no counterpart in the reference.

Parameter vector layout (as the C ABI): head W[C][9][2], b[C]; per block
W1[C][9][C], b1[C], W2[C][9][C], b2[C]; tail W[2][9][C], b[2].
Tensors NHWC; 3x3 convolutions with zero padding, tap t = 3*(dy+1) + (dx+1).
Pinned by finite differences in tests/test_oracle_cnn.py.
"""
import numpy as np


def layout(C, NB):
    off, shapes = 0, []

    def add(name, shape):
        nonlocal off
        n = int(np.prod(shape))
        shapes.append((name, off, shape))
        off += n
    add("head_w", (C, 9, 2))
    add("head_b", (C,))
    for b in range(NB):
        add(f"w1_{b}", (C, 9, C))
        add(f"b1_{b}", (C,))
        add(f"w2_{b}", (C, 9, C))
        add(f"b2_{b}", (C,))
    add("tail_w", (2, 9, C))
    add("tail_b", (2,))
    return shapes, off


def unpack(p, C, NB):
    shapes, n = layout(C, NB)
    assert p.size == n
    return {name: p[o:o + int(np.prod(s))].reshape(s) for name, o, s in shapes}


def _patches(x):
    """x [B,H,W,C] -> [B,H,W,9,C] of zero-padded 3x3 neighbourhoods"""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    out = np.empty((B, H, W, 9, C), x.dtype)
    for t in range(9):
        dy, dx = t // 3, t % 3
        out[:, :, :, t] = xp[:, dy:dy + H, dx:dx + W]
    return out


def conv(x, w, b):
    """y[p, o] = sum_{t,c} w[o, t, c] x[p + off(t), c] + b[o]"""
    return np.einsum("bhwtc,otc->bhwo", _patches(x), w) + b


def conv_grads(x, w, gy):
    """gradients of conv wrt x, w, b given gy = dL/dy"""
    P = _patches(x)
    gw = np.einsum("bhwtc,bhwo->otc", P, gy)
    gb = gy.sum(axis=(0, 1, 2))
    # dx[q, c] = sum_{o,t} w[o,t,c] gy[q - off(t), o]
    B, H, W, _ = gy.shape
    gyp = np.zeros((B, H + 2, W + 2, gy.shape[3]), gy.dtype)
    gyp[:, 1:-1, 1:-1] = gy
    gx = np.zeros(x.shape, gy.dtype)
    for t in range(9):
        dy, dx = t // 3, t % 3
        # y at q - off(t) lives at padded index (q - off + 1)
        g = gyp[:, 2 - dy:2 - dy + H, 2 - dx:2 - dx + W]
        gx += np.einsum("bhwo,oc->bhwc", g, w[:, t, :])
    return gx, gw, gb


def forward(p, x, C=32, NB=8, keep=False):
    q = unpack(p, C, NB)
    h = [conv(x, q["head_w"], q["head_b"])]
    a = [None]
    for b in range(NB):
        ab = np.maximum(conv(h[-1], q[f"w1_{b}"], q[f"b1_{b}"]), 0.0)
        a.append(ab)
        h.append(h[-1] + conv(ab, q[f"w2_{b}"], q[f"b2_{b}"]))
    y = x + conv(h[-1], q["tail_w"], q["tail_b"])
    return (y, h, a) if keep else y


def loss_and_grad(p, x, t, C=32, NB=8):
    """loss = mean (y - t)^2 and its gradient wrt the parameter vector"""
    y, h, a = forward(p, x, C, NB, keep=True)
    q = unpack(p, C, NB)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    gy = 2.0 * (y - t) / n
    g = {}
    gh, g["tail_w"], g["tail_b"] = conv_grads(h[NB], q["tail_w"], gy)
    for b in reversed(range(NB)):
        ga, g[f"w2_{b}"], g[f"b2_{b}"] = conv_grads(a[b + 1], q[f"w2_{b}"], gh)
        ga = ga * (a[b + 1] > 0)
        gx, g[f"w1_{b}"], g[f"b1_{b}"] = conv_grads(h[b], q[f"w1_{b}"], ga)
        gh = gh + gx
    _, g["head_w"], g["head_b"] = conv_grads(x, q["head_w"], gh)
    shapes, ntot = layout(C, NB)
    out = np.zeros(ntot)
    for name, o, s in shapes:
        out[o:o + int(np.prod(s))] = g[name].ravel()
    return loss, out
