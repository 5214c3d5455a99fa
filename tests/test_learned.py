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


def _params(seed, desc=DESC, scale=None):
    """Random parameters; scale None: 0.3 everywhere, else `scale` times the
    He-normal width sqrt(2 / fan_in) of each layer (activations of order one
    at any width)."""
    rng = np.random.default_rng(seed)
    if scale is None:
        n = sum(co * ci * 27 + co for co, ci in NO.layer_shapes(desc["channels"], desc["hidden"], desc["layers"]))
        return (rng.normal(size=n) * 0.3).astype(np.float32)
    out = []
    for co, ci in NO.layer_shapes(desc["channels"], desc["hidden"], desc["layers"]):
        out += [rng.normal(size=co * ci * 27) * scale * np.sqrt(2.0 / (ci * 27)), rng.normal(size=co) * 0.1]
    return np.concatenate(out).astype(np.float32)


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


def _check_net(lib, desc=DESC, batch=3, scale=None):
    x, t = _case(5, batch, desc)
    p = _params(6, desc, scale)
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


def test_net_emulator_matrix_core_blocking(emu_lib):
    """The matrix-core convolution at the 9x9 network's width: two time blocks
    (the second a tail of 3 slices), six position tiles per wavefront, two
    16-channel output blocks, zero-padded input channels in the first layer."""
    _check_net(emu_lib, dict(lt=12, ly=9, lx=9, channels=2, hidden=32, layers=4), batch=2, scale=1.0)


def test_net_rejects_bad_batch(emu_lib):
    net = ConvNet(2, 2, 2, 2, 2, 2, max_batch=2, lib=emu_lib)
    with pytest.raises(AssertionError):
        net.forward(np.zeros((3, 2, 2, 2, 2), np.float32))
    net.close()


@pytest.mark.gpu
def test_net_gpu(hip_lib):
    _check_net(hip_lib)
    _check_net(hip_lib, dict(lt=40, ly=3, lx=3, channels=2, hidden=8, layers=3), batch=4)
    # 4 layers x 32 channels: the forward kernel's widest register blocking
    _check_net(hip_lib, dict(lt=6, ly=3, lx=4, channels=2, hidden=32, layers=4), batch=3)
    # the 9x9 network: several time blocks and a tail block per chain
    _check_net(hip_lib, dict(lt=80, ly=9, lx=9, channels=2, hidden=32, layers=4), batch=4, scale=1.0)
    # 12x12: one time block holds fewer slices
    _check_net(hip_lib, dict(lt=20, ly=12, lx=12, channels=2, hidden=32, layers=4), batch=2, scale=1.0)


def _device_move_case(lib, pairing):
    import ctypes as C
    from bafqmc_amd import global_moves as GM
    from bafqmc_amd.model import ModelSpec
    from bafqmc_amd.proposal import from_net, to_net
    from bafqmc_amd.solver import BatchedSolver
    from tests.helpers import solver_params
    d = dict(pairing=pairing, u1=0.8, u2=-0.3, mu=-4.0, rdelta=0.4 if pairing else 0.0, nlx=3, nly=3, ltrot=8,
             beta=1.0, nwrap=4, shift_loc=1.0)
    p = solver_params(d)
    rng = np.random.default_rng(21)
    phis = [rng.normal(size=(8, 9, 2)) * 0.5 for _ in range(4)]
    s = BatchedSolver([ModelSpec(p)], [0] * 4, lib=lib)
    s.set_fields(phis, [5, 6, 7, 8])
    s.prepare()
    net = ConvNet(8, 3, 3, 2, 3, 3, max_batch=3, seed=4, lib=lib)    # 3 < 4 chains: two chunks
    lam, sigma = 0.7, np.array([0.05, 0.08])
    out = np.zeros((4, 5))
    sig = np.ascontiguousarray(sigma)
    s._check(lib.bafq_global_propose(s.h, net.h, lam, sig.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(99),
                                     out.ctypes.data_as(C.POINTER(C.c_double))), "propose")
    P, _ = s.get_fields()
    F = np.stack(phis)
    P = np.stack(P)
    def fwd(X):
        return np.concatenate([np.stack(from_net(net.forward(to_net(list(X[i:i + 3]), p)), p))
                               for i in range(0, len(X), 3)])
    f0 = lam * fwd(F)
    f1 = lam * fwd(P)
    w = 1.0 / (2.0 * sigma ** 2)
    lq = np.sum(w * ((P - F - f0) ** 2 - (F - P - f1) ** 2), axis=(1, 2, 3))
    np.testing.assert_allclose(out[:, 0], lq, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(out[:, 1], np.sum(F ** 2, axis=(1, 2, 3)), rtol=1e-12)
    np.testing.assert_allclose(out[:, 2], np.sum(P ** 2, axis=(1, 2, 3)), rtol=1e-12)
    np.testing.assert_allclose(out[:, 3], np.sum(F[..., 1], axis=(1, 2)), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out[:, 4], np.sum(P[..., 1], axis=(1, 2)), rtol=1e-12, atol=1e-12)
    xi = (P - F - f0) / sigma                                  # the device's normal numbers
    assert abs(xi.mean()) < 0.15 and abs(xi.var() - 1.0) < 0.2
    # G of the proposals is current, then rejected chains get phi and its G back
    g_prop = [s.green(c) for c in range(4)]
    accept = np.array([1, 0, 1, 0], dtype=np.int32)
    s._check(lib.bafq_global_finish(s.h, accept.ctypes.data_as(C.POINTER(C.c_int32))), "finish")
    fin, seeds = s.get_fields()
    ref = BatchedSolver([ModelSpec(p)], [0] * 4, lib=lib)
    ref.set_fields(phis, [5, 6, 7, 8])
    ref.prepare()
    for c in range(4):
        np.testing.assert_array_equal(fin[c], P[c] if accept[c] else F[c])
        g = s.green(c)
        if accept[c]:
            np.testing.assert_array_equal(g, g_prop[c])
        else:
            np.testing.assert_allclose(g, ref.green(c), rtol=0, atol=1e-12 * np.max(np.abs(g)))
    assert list(seeds) == [5, 6, 7, 8]
    # the acceptance built from these pieces is the host one's
    assert GM.metropolis_global_device(s, net, lam, sigma, np.random.default_rng(3)).dtype == bool
    s.close()
    ref.close()
    net.close()


@pytest.mark.parametrize("pairing", [False, True])
def test_device_move_emulator(emu_lib, pairing):
    _device_move_case(emu_lib, pairing)


@pytest.mark.gpu
@pytest.mark.parametrize("pairing", [False, True])
def test_device_move_gpu(hip_lib, pairing):
    _device_move_case(hip_lib, pairing)


def _rejected_move_is_invisible(lib):
    """A device move whose chains are all rejected leaves no trace: the next
    sweep is bit for bit that of a batch that never moved."""
    import ctypes as C
    from bafqmc_amd.model import ModelSpec
    from bafqmc_amd.solver import BatchedSolver
    from tests.helpers import solver_params
    d = dict(pairing=True, u1=0.8, u2=-0.3, mu=-4.0, rdelta=0.4, nlx=3, nly=3, ltrot=8, beta=1.0, nwrap=4,
             shift_loc=1.0)
    p = solver_params(d)
    rng = np.random.default_rng(8)
    phis = [rng.normal(size=(8, 9, 2)) * 0.5 for _ in range(3)]
    runs = []
    for move in (False, True):
        s = BatchedSolver([ModelSpec(p)], [0] * 3, lib=lib)
        s.set_fields(phis, [15, 16, 17])
        s.prepare()
        s.sweep(1)
        s.read_bin()
        if move:
            out = np.zeros((3, 5))
            sig = np.array([0.1, 0.1])
            s._check(lib.bafq_global_propose(s.h, None, 1.0, sig.ctypes.data_as(C.POINTER(C.c_double)),
                                             C.c_uint64(5), out.ctypes.data_as(C.POINTER(C.c_double))), "propose")
            flags = np.zeros(3, dtype=np.int32)
            s._check(lib.bafq_global_finish(s.h, flags.ctypes.data_as(C.POINTER(C.c_int32))), "finish")
        s.sweep(1)
        obs, st = s.read_bin_raw()
        runs.append((obs, st, [s.green(c) for c in range(3)], s.get_fields()))
        s.close()
    (o0, s0, g0, f0), (o1, s1, g1, f1) = runs
    np.testing.assert_array_equal(o0, o1)
    np.testing.assert_array_equal(s0[:, :8], s1[:, :8])
    for a, b in zip(g0, g1):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(f0[0], f1[0]):
        np.testing.assert_array_equal(a, b)


def test_rejected_device_move_is_invisible_emulator(emu_lib):
    _rejected_move_is_invisible(emu_lib)


@pytest.mark.gpu
def test_rejected_device_move_is_invisible_gpu(hip_lib):
    _rejected_move_is_invisible(hip_lib)
