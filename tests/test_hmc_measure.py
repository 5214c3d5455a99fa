"""Measurement passes and Hamiltonian global moves with a learned force.

* bafq_measure_pass: a left sweep with frozen fields that measures every
  slice.  With shiftLoc = 0 the reference's own sweep cannot change the
  fields, and its bin (both half-sweeps, 2 Ltrot measurements of the same
  Green's functions) must equal the measurement pass's (Ltrot
  measurements) -- checked against the oracle's Chain.bin.  The pass leaves
  fields and RNG streams untouched.
* bafq_hmc_propose: the device's momenta and kinetic energies are recovered
  on the host from phi, phi' and the network's forward pass, the Gaussian sums
  are recomputed, a fully rejected move leaves no trace, and a move without a
  network at steps * angle = pi / 2 is the exact rotation phi' = c + p.
* The f32 MFMA lane layout the convolutions rely on.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from oracle import bafq_oracle as O
from bafqmc_amd import global_moves as GM
from bafqmc_amd.model import ModelSpec
from bafqmc_amd.proposal import ConvNet, from_net, to_net
from bafqmc_amd.solver import BatchedSolver
from tests.helpers import fixed_fields, oracle_params, solver_params


def _mfma32_layout(lib):
    out = lib.debug_mfma32_probe()
    for p in range(64):
        ip, kp = p % 16, p // 16
        for lane in range(64):
            for r in range(4):
                row, col = 4 * (lane // 16) + r, lane % 16
                want = 1000.0 + col + 16 * kp if row == ip else 0.0
                assert out[p, lane, r] == want, (p, lane, r, out[p, lane, r], want)


def test_mfma32_layout_emulator(emu_lib):
    _mfma32_layout(emu_lib)


@pytest.mark.gpu
def test_mfma32_layout_gpu(hip_lib):
    _mfma32_layout(hip_lib)


MEASURE_CASES = {
    "nc": dict(pairing=False, u1=0.7, u2=-0.2, mu=-4.5, nlx=3, nly=3, ltrot=12, beta=1.2, nwrap=4, shift_loc=0.0),
    "pairing": dict(pairing=True, u1=0.7, u2=-0.2, mu=-4.5, rdelta=0.3, nlx=3, nly=3, ltrot=12, beta=1.2, nwrap=4,
                    shift_loc=0.0),
}


def _measure_pass_case(lib, name, schedule=None):
    d = MEASURE_CASES[name]
    p = solver_params(d)
    phi = fixed_fields(d["ltrot"], 9)
    ch = O.Chain(oracle_params(d), phi, 777)
    ch.pre()
    ref = ch.bin()
    s = BatchedSolver([ModelSpec(p)], [0], lib=lib, schedule=schedule)
    s.set_fields([phi], [777])
    s.prepare()
    s.measure_pass(1)
    got = s.read_bin()[0]
    assert got["_nobs"] == d["ltrot"]
    for k in ("density_total", "energy_density", "sf_K", "dw_K", "kinetic", "doubleOcc", "pair_equal"):
        if k not in ref:
            continue
        a, b = complex(got[k]), complex(ref[k])
        assert abs(a - b) <= 1e-9 * abs(b) + 1e-11, (k, a, b)
    f, seeds = s.get_fields()
    np.testing.assert_array_equal(f[0], phi)
    assert int(seeds[0]) == 777
    # a second pass over the state the first leaves measures the same
    s.measure_pass(1)
    again = s.read_bin()[0]
    assert abs(complex(again["sf_K"]) - complex(got["sf_K"])) <= 1e-10 * abs(complex(got["sf_K"]))
    s.close()


def _checkpoint_pass_case(lib, name, schedule=None):
    """Measurement at the stabilisation points only: the average of the
    oracle's estimators on the dense Green's function (1 - B(tau) ... B(tau+1))^-1
    at tau = Nwrap, 2 Nwrap, ..., Ltrot (no stabilisation, explicit products)."""
    d = MEASURE_CASES[name]
    p = solver_params(d)
    phi = fixed_fields(d["ltrot"], 9)
    ch = O.Chain(oracle_params(d), phi, 777)
    samples = []
    for tau in range(d["nwrap"], d["ltrot"] + 1, d["nwrap"]):
        ch.prop.g = O.dense_green(ch, tau % d["ltrot"])[0]
        samples.append(O.measure_pairing(ch) if d["pairing"] else O.measure_nc(ch))
    ref = {k: sum(smp[k] for smp in samples) / len(samples) for k in samples[0]}
    s = BatchedSolver([ModelSpec(p)], [0], lib=lib, schedule=schedule)
    s.set_fields([phi], [777])
    s.prepare()
    s.measure_pass(1, mode="checkpoints")
    got = s.read_bin()[0]
    assert got["_nobs"] == d["ltrot"] // d["nwrap"]
    for k in ("density_total", "energy_density", "sf_K", "dw_K", "kinetic", "doubleOcc", "pair_equal"):
        if k not in ref:
            continue
        a, b = complex(got[k]), complex(ref[k])
        assert abs(a - b) <= 1e-9 * abs(b) + 1e-11, (k, a, b)
    # the pass ends where sweep_L ends: a full measurement pass from here agrees
    # with one from a fresh rebuild
    s.measure_pass(1)
    after = s.read_bin()[0]
    s2 = BatchedSolver([ModelSpec(p)], [0], lib=lib, schedule=schedule)
    s2.set_fields([phi], [777])
    s2.prepare()
    s2.measure_pass(1)
    fresh = s2.read_bin()[0]
    for k in ("density_total", "sf_K", "dw_K"):
        a, b = complex(after[k]), complex(fresh[k])
        assert abs(a - b) <= 1e-9 * abs(b) + 1e-11, (k, a, b)
    s.close()
    s2.close()


def _current_measurement_case(lib, name, schedule=None):
    """mode "current": the oracle's estimators on G(beta) of the stabilised
    rebuild, once; fields, G and the RNG stream unchanged."""
    d = MEASURE_CASES[name]
    p = solver_params(d)
    phi = fixed_fields(d["ltrot"], 9)
    ch = O.Chain(oracle_params(d), phi, 777)
    ch.pre()
    ref = O.measure_pairing(ch) if d["pairing"] else O.measure_nc(ch)
    s = BatchedSolver([ModelSpec(p)], [0], lib=lib, schedule=schedule)
    s.set_fields([phi], [777])
    s.prepare()
    g0 = s.green(0)
    s.measure_pass(5, mode="current")
    got = s.read_bin()[0]
    assert got["_nobs"] == 1
    for k in ("density_total", "energy_density", "sf_K", "dw_K", "kinetic", "pair_equal"):
        if k not in ref:
            continue
        a, b = complex(got[k]), complex(ref[k])
        assert abs(a - b) <= 1e-9 * abs(b) + 1e-11, (k, a, b)
    np.testing.assert_array_equal(s.green(0), g0)
    s.close()


@pytest.mark.parametrize("name", sorted(MEASURE_CASES))
def test_current_measurement_emulator(emu_lib, name):
    _current_measurement_case(emu_lib, name)
    _current_measurement_case(emu_lib, name, (0, 0))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MEASURE_CASES))
def test_current_measurement_gpu(hip_lib, name):
    _current_measurement_case(hip_lib, name)
    _current_measurement_case(hip_lib, name, (0, 0))


@pytest.mark.parametrize("name", sorted(MEASURE_CASES))
def test_checkpoint_pass_emulator(emu_lib, name):
    _checkpoint_pass_case(emu_lib, name)
    _checkpoint_pass_case(emu_lib, name, (0, 0))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MEASURE_CASES))
@pytest.mark.parametrize("schedule", [None, (0, 0)])
def test_checkpoint_pass_gpu(hip_lib, name, schedule):
    _checkpoint_pass_case(hip_lib, name, schedule)


@pytest.mark.parametrize("name", sorted(MEASURE_CASES))
def test_measure_pass_emulator(emu_lib, name):
    _measure_pass_case(emu_lib, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MEASURE_CASES))
@pytest.mark.parametrize("schedule", [None, (0, 0)])
def test_measure_pass_gpu(hip_lib, name, schedule):
    _measure_pass_case(hip_lib, name, schedule)


def _hmc_case(lib, pairing, steps, use_net=True, nl=3, lt=8, hidden=3, layers=3, nchains=4, max_batch=3):
    d = dict(pairing=pairing, u1=0.8, u2=-0.3, mu=-4.0, rdelta=0.4 if pairing else 0.0, nlx=nl, nly=nl, ltrot=lt,
             beta=lt / 8.0, nwrap=4, shift_loc=1.0)
    p = solver_params(d)
    ns = nl * nl
    rng = np.random.default_rng(31)
    phis = [rng.normal(size=(lt, ns, 2)) * 0.5 for _ in range(nchains)]
    seeds = list(range(5, 5 + nchains))
    s = BatchedSolver([ModelSpec(p)], [0] * nchains, lib=lib)
    s.set_fields(phis, seeds)
    s.prepare()
    net = ConvNet(lt, nl, nl, 2, hidden, layers, max_batch=max_batch, seed=4, lib=lib) if use_net else None
    angle = 0.5 * math.pi / steps
    cen = GM.gaussian_centers(p)
    out = np.zeros((nchains, 5))
    fsc = np.array([0.7, 1.3])                    # per-channel force scale
    s._check(lib.bafq_hmc_propose(s.h, net.h if net else None, angle, steps,
                                  np.ascontiguousarray(cen).ctypes.data_as(C.POINTER(C.c_double)),
                                  fsc.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(99),
                                  out.ctypes.data_as(C.POINTER(C.c_double))), "hmc")
    P = np.stack(s.get_fields()[0])
    F = np.stack(phis)
    mask = fsc

    def force(X):
        if net is None:
            return np.zeros_like(X)
        return np.concatenate([np.stack(from_net(net.forward(to_net(list(X[i:i + max_batch]), p)), p))
                               for i in range(0, len(X), max_batch)])

    # replay the trajectory on the host from the device's first momentum:
    # with one step, p_half = (phi' - c - (phi - c) cos h) / sin h
    if steps == 1:
        h = angle
        ph = (P - cen - (F - cen) * math.cos(h)) / math.sin(h)
        p0 = ph - 0.5 * h * force(F) * mask
        p1 = ph * math.cos(h) - (F - cen) * math.sin(h) + 0.5 * h * force(P) * mask
        kin = 0.5 * np.sum(p0 ** 2 - p1 ** 2, axis=(1, 2, 3))
        np.testing.assert_allclose(out[:, 0], kin, rtol=1e-8, atol=1e-8 * np.abs(kin).max())
        assert abs(p0.mean()) < 0.1 and abs(p0.var() - 1.0) < 0.1      # the device's normal momenta
    np.testing.assert_allclose(out[:, 1], np.sum(F ** 2, axis=(1, 2, 3)), rtol=1e-12)
    np.testing.assert_allclose(out[:, 2], np.sum(P ** 2, axis=(1, 2, 3)), rtol=1e-12)
    np.testing.assert_allclose(out[:, 3], np.sum(F[..., 1], axis=(1, 2)), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out[:, 4], np.sum(P[..., 1], axis=(1, 2)), rtol=1e-12, atol=1e-12)
    if net is None and steps == 1:
        # exact quarter rotation: phi' - c is the momentum, the new momentum -(phi - c)
        np.testing.assert_allclose(out[:, 0], 0.5 * np.sum((P - cen) ** 2 - (F - cen) ** 2, axis=(1, 2, 3)),
                                   rtol=1e-9)
    # G of phi' is current: a fresh rebuild agrees
    ref = BatchedSolver([ModelSpec(p)], [0] * nchains, lib=lib)
    ref.set_fields(list(P), seeds)
    ref.prepare()
    for c in range(nchains):
        g = s.green(c)
        np.testing.assert_allclose(g, ref.green(c), rtol=0, atol=1e-10 * np.max(np.abs(g)))
    # reject every second chain
    accept = np.array([c % 2 for c in range(nchains)], dtype=np.int32)
    s._check(lib.bafq_global_finish(s.h, accept.ctypes.data_as(C.POINTER(C.c_int32))), "finish")
    fin, sd = s.get_fields()
    for c in range(nchains):
        np.testing.assert_array_equal(fin[c], P[c] if accept[c] else F[c])
    assert list(sd) == seeds
    if steps == 1:
        # a second move starts from the network output the first one left
        # current (the proposal's where accepted, the old fields' elsewhere)
        F2 = np.stack(fin)
        out2 = np.zeros((nchains, 5))
        s._check(lib.bafq_hmc_propose(s.h, net.h if net else None, angle, steps,
                                      np.ascontiguousarray(cen).ctypes.data_as(C.POINTER(C.c_double)),
                                      fsc.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(7),
                                      out2.ctypes.data_as(C.POINTER(C.c_double))), "hmc 2")
        P2 = np.stack(s.get_fields()[0])
        h = angle
        ph = (P2 - cen - (F2 - cen) * math.cos(h)) / math.sin(h)
        p0 = ph - 0.5 * h * force(F2) * mask
        p1 = ph * math.cos(h) - (F2 - cen) * math.sin(h) + 0.5 * h * force(P2) * mask
        kin = 0.5 * np.sum(p0 ** 2 - p1 ** 2, axis=(1, 2, 3))
        np.testing.assert_allclose(out2[:, 0], kin, rtol=1e-8, atol=1e-8 * np.abs(kin).max())
        flags = np.ones(nchains, dtype=np.int32)
        s._check(lib.bafq_global_finish(s.h, flags.ctypes.data_as(C.POINTER(C.c_int32))), "finish 2")
    # the exact acceptance helper runs end to end
    acc = GM.hmc_global_device(s, net, angle, steps, np.random.default_rng(3))
    assert acc.dtype == bool and acc.shape == (nchains,)
    s.close()
    ref.close()
    if net:
        net.close()


@pytest.mark.parametrize("pairing", [False, True])
def test_hmc_move_emulator(emu_lib, pairing):
    _hmc_case(emu_lib, pairing, 1)
    _hmc_case(emu_lib, pairing, 1, use_net=False)
    _hmc_case(emu_lib, pairing, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("pairing", [False, True])
def test_hmc_move_gpu(hip_lib, pairing):
    _hmc_case(hip_lib, pairing, 1)
    _hmc_case(hip_lib, pairing, 1, use_net=False)
    _hmc_case(hip_lib, pairing, 4)


@pytest.mark.gpu
def test_hmc_move_gpu_c4_size(hip_lib):
    """The 9x9, Ltrot = 80 lattice with the 4-layer, 32-channel network."""
    _hmc_case(hip_lib, False, 1, nl=9, lt=80, hidden=32, layers=4, nchains=6, max_batch=4)


@pytest.mark.gpu
def test_hmc_move_gpu_c5_size(hip_lib):
    """The 12x12 pairing lattice (N = 288), Ltrot = 100, 4-layer 32-channel network."""
    _hmc_case(hip_lib, True, 1, nl=12, lt=100, hidden=32, layers=4, nchains=2, max_batch=2)


def _rejected_hmc_is_invisible(lib):
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
            cen = np.ascontiguousarray(GM.gaussian_centers(p))
            one = np.ones(2)
            s._check(lib.bafq_hmc_propose(s.h, None, 0.3, 2, cen.ctypes.data_as(C.POINTER(C.c_double)),
                                          one.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(5),
                                          out.ctypes.data_as(C.POINTER(C.c_double))), "hmc")
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


def test_rejected_hmc_is_invisible_emulator(emu_lib):
    _rejected_hmc_is_invisible(emu_lib)


@pytest.mark.gpu
def test_rejected_hmc_is_invisible_gpu(hip_lib):
    _rejected_hmc_is_invisible(hip_lib)


def test_hmc_rejects_bad_arguments(emu_lib):
    p = solver_params(MEASURE_CASES["nc"])
    s = BatchedSolver([ModelSpec(p)], [0], lib=emu_lib)
    s.set_fields([fixed_fields(12, 9)], [3])
    s.prepare()
    out = np.zeros((1, 5))
    cen = np.zeros(2)
    for angle, steps in ((0.0, 1), (1.0, 0), (4.0, 1), (1.0, 65)):
        rc = emu_lib.bafq_hmc_propose(s.h, None, angle, steps, cen.ctypes.data_as(C.POINTER(C.c_double)),
                                      np.ones(2).ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(1),
                                      out.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc != 0
    s.close()
