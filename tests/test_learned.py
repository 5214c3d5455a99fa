"""The learned-proposal network: device forward/backward/Adam against the NumPy
oracle (oracle/net_oracle.py), the oracle pinned by an independent periodic
correlation (scipy.ndimage.correlate, mode='wrap') and by central differences
of its loss in float64.  The reference has no network, so none of this has a
reference counterpart; the proposal's exactness is checked in
tests/test_learned_gpu.py against exact diagonalisation.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import net_oracle as NO
from bafqmc_amd.proposal import ConvNet

DESC = dict(lt=4, ly=3, lx=5, channels=2, hidden=3, layers=3)


def _case(seed, batch, desc=DESC):
    rng = np.random.default_rng(seed)
    shape = (batch, desc["channels"], desc["lt"], desc["ly"], desc["lx"])
    return rng.normal(size=shape).astype(np.float32), rng.normal(size=shape).astype(np.float32)


def _params(seed, desc=DESC):
    n = sum(co * ci * 27 + co for co, ci in NO.layer_shapes(desc["channels"], desc["hidden"], desc["layers"]))
    return (np.random.default_rng(seed).normal(size=n) * 0.3).astype(np.float32)


def test_oracle_conv_is_periodic_correlation():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 4, 3, 5))
    w = rng.normal(size=(2, 3, 3, 3, 3))
    b = rng.normal(size=2)
    got = NO.conv(x, w, b)
    for bi in range(2):
        for co in range(2):
            ref = b[co] + sum(ndimage.correlate(x[bi, ci], w[co, ci], mode="wrap") for ci in range(3))
            np.testing.assert_allclose(got[bi, co], ref, rtol=1e-12, atol=1e-12)


def test_oracle_gradient_central_differences():
    x, t = _case(2, 2)
    p = _params(3).astype(np.float64)
    x, t = x.astype(np.float64), t.astype(np.float64)
    c, h, L = DESC["channels"], DESC["hidden"], DESC["layers"]
    _, g = NO.loss_and_grad(p, x, t, c, h, L)
    rng = np.random.default_rng(4)
    idx = rng.choice(p.size, 40, replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (NO.loss_and_grad(pp, x, t, c, h, L)[0] - NO.loss_and_grad(pm, x, t, c, h, L)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def _check_net(lib, desc=DESC, batch=3):
    x, t = _case(5, batch, desc)
    p = _params(6, desc)
    c, h, L = desc["channels"], desc["hidden"], desc["layers"]
    net = ConvNet(desc["lt"], desc["ly"], desc["lx"], c, h, L, max_batch=4, lib=lib)
    net.set_params(p)
    np.testing.assert_array_equal(net.get_params(), p)
    y = net.forward(x)
    p64, x64, t64 = p.astype(np.float64), x.astype(np.float64), t.astype(np.float64)
    yref = NO.forward(p64, x64, c, h, L)
    np.testing.assert_allclose(y, yref, rtol=1e-4, atol=1e-4 * np.abs(yref).max())
    loss, g = net.gradient(x, t)
    lref, gref = NO.loss_and_grad(p64, x64, t64, c, h, L)
    assert abs(loss - lref) <= 1e-4 * lref
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-4 * np.abs(gref).max())
    # one Adam step (Kingma & Ba, bias-corrected): |dp| = lr for every nonzero gradient
    lr = 1e-2
    net.set_params(p)
    net.train(x, t, 1, lr)
    p1 = net.get_params().astype(np.float64)
    gd = g.astype(np.float64)
    m, v = 0.1 * gd, 0.001 * gd * gd
    ref1 = p64 - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    np.testing.assert_allclose(p1, ref1, rtol=1e-5, atol=1e-6)
    # training decreases the loss
    net.set_params(p)
    l0 = net.gradient(x, t)[0]
    net.train(x, t, 30, 1e-2)
    assert net.gradient(x, t)[0] < 0.8 * l0
    net.close()


def test_net_emulator(emu_lib):
    _check_net(emu_lib)


def test_net_rejects_bad_batch(emu_lib):
    net = ConvNet(2, 2, 2, 2, 2, 2, max_batch=2, lib=emu_lib)
    with pytest.raises(AssertionError):
        net.forward(np.zeros((3, 2, 2, 2, 2), np.float32))
    net.close()


@pytest.mark.gpu
def test_net_gpu(hip_lib):
    _check_net(hip_lib)
    _check_net(hip_lib, dict(lt=40, ly=3, lx=3, channels=2, hidden=8, layers=3), batch=4)
