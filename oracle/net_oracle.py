"""NumPy restatement of the learned-proposal network (test infrastructure).

The reference has no network; this is the CPU check for the device kernels of
bafqmc_amd/csrc/net.hpp: periodic 3x3x3 convolutions over (t, y, x), ReLU on
hidden layers, linear output, mean-squared-error loss.  Parameters are the
concatenation over layers of W[cout][cin][3][3][3] and b[cout].  The
gradient is pinned by central differences in float64 (tests/test_learned.py).
"""
from __future__ import annotations

import numpy as np


def layer_shapes(channels, hidden, layers):
    out = []
    for l in range(layers):
        cin = channels if l == 0 else hidden
        cout = channels if l == layers - 1 else hidden
        out.append((cout, cin))
    return out


def unpack(params, channels, hidden, layers):
    ws, bs, off = [], [], 0
    for cout, cin in layer_shapes(channels, hidden, layers):
        n = cout * cin * 27
        ws.append(params[off:off + n].reshape(cout, cin, 3, 3, 3))
        off += n
        bs.append(params[off:off + cout])
        off += cout
    assert off == params.size
    return ws, bs


def conv(x, w, b):
    """x: [B, cin, T, Y, X] -> [B, cout, T, Y, X]; out[p] = b + sum_d w[d] x[p + d] (periodic)."""
    out = np.zeros((x.shape[0], w.shape[0]) + x.shape[2:], dtype=np.result_type(x, w))
    for dt in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                sh = np.roll(x, (-dt, -dy, -dx), axis=(2, 3, 4))
                out += np.einsum("oc,bctyx->botyx", w[:, :, dt + 1, dy + 1, dx + 1], sh)
    return out + b[None, :, None, None, None]


def forward(params, x, channels, hidden, layers, keep=False):
    ws, bs = unpack(params, channels, hidden, layers)
    acts = [x]
    a = x
    for l in range(layers):
        a = conv(a, ws[l], bs[l])
        if l < layers - 1:
            a = np.maximum(a, 0.0)
        acts.append(a)
    return (a, acts) if keep else a


def loss_and_grad(params, x, target, channels, hidden, layers):
    """loss = |f(x) - target|^2 / size and d loss / d params."""
    ws, bs = unpack(params, channels, hidden, layers)
    y, acts = forward(params, x, channels, hidden, layers, keep=True)
    n = y.size
    loss = float(np.sum((y - target) ** 2) / n)
    g = 2.0 * (y - target) / n
    grads = []
    for l in range(layers - 1, -1, -1):
        a_in = acts[l]
        gw = np.zeros_like(ws[l])
        for dt in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    sh = np.roll(a_in, (-dt, -dy, -dx), axis=(2, 3, 4))
                    gw[:, :, dt + 1, dy + 1, dx + 1] = np.einsum("botyx,bctyx->oc", g, sh)
        gb = g.sum(axis=(0, 2, 3, 4))
        grads.append((gw, gb))
        if l > 0:
            gin = np.zeros_like(a_in)
            for dt in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        gin += np.einsum("oc,botyx->bctyx", ws[l][:, :, dt + 1, dy + 1, dx + 1],
                                         np.roll(g, (dt, dy, dx), axis=(2, 3, 4)))
            g = gin * (a_in > 0)
    flat = []
    for gw, gb in reversed(grads):
        flat += [gw.ravel(), gb.ravel()]
    return loss, np.concatenate(flat)
