"""Global moves with exact Metropolis-Hastings (bafqmc_amd.global_moves).

CPU tests run the device code on the emulator: the weight built from the
device's log|det G| must reproduce, for single-site moves, the local
Metropolis ratio that the reference's LocalU_metro uses, computed here by brute
force from dense determinants of the oracle's slice matrices (number conserving
localU.f90:55-61: r_gauss |det(1-B)/det(1-B')|^2; pairing localU.f90:59-66:
r_gauss r_const |det(1-B4)/det(1-B4')|^(1/2) on the 4Ns Nambu form).  The GPU
test interleaves global shift moves with local sweeps and checks the sampled
observables against exact diagonalisation.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import bafq_oracle as O
from bafqmc_amd import global_moves as GM
from bafqmc_amd.model import ModelSpec
from bafqmc_amd.solver import BatchedSolver
from tests.helpers import oracle_params, solver_params


def _brute_log_ratio(d, phi, phi2, nf):
    """log of the reference's acceptance ratio for phi -> phi2 (one site changed)."""
    ch = O.Chain(oracle_params(d), phi, 1)
    ch2 = O.Chain(oracle_params(d), phi2, 1)
    _, prod = O.dense_green(ch, 1)
    _, prod2 = O.dense_green(ch2, 1)
    n = prod.shape[0]
    ld = math.log(abs(np.linalg.det(np.eye(n) - prod)))
    ld2 = math.log(abs(np.linalg.det(np.eye(n) - prod2)))
    lg = -0.5 * float(np.sum(phi2 ** 2) - np.sum(phi ** 2))
    if not d["pairing"]:
        return lg + 2.0 * (ld - ld2)
    a2 = ch.alpha[1].real
    lc = -a2 * float(np.sum(phi2[:, :, 1]) - np.sum(phi[:, :, 1])) if nf == 2 else 0.0
    return lg + lc + 0.5 * (ld - ld2)


@pytest.mark.parametrize("pairing", [False, True])
def test_weight_reproduces_local_ratios(emu_lib, pairing):
    d = dict(pairing=pairing, u1=0.8, u2=-0.3, mu=-4.0, rdelta=0.4 if pairing else 0.0, nlx=3, nly=3, ltrot=8,
             beta=1.0, nwrap=4, shift_loc=1.0)
    rng = np.random.default_rng(11)
    phi = rng.normal(size=(8, 9, 2))
    p = solver_params(d)
    s = BatchedSolver([ModelSpec(p)], [0], lib=emu_lib)
    s.set_fields([phi], [123])
    s.prepare()
    lw0 = GM.log_weights(s)[0]
    for (t, i, nf) in ((0, 0, 1), (3, 4, 2), (7, 8, 2), (5, 2, 1)):
        phi2 = phi.copy()
        phi2[t, i, nf - 1] += 0.37
        s.set_fields([phi2], [123])
        s.prepare()
        lw1 = GM.log_weights(s)[0]
        brute = _brute_log_ratio(d, phi, phi2, nf)
        assert abs((lw1 - lw0) - brute) < 1e-9 * max(1.0, abs(brute)), ((t, i, nf), lw1 - lw0, brute)
    s.close()


def test_metropolis_global_keeps_consistent_states(emu_lib):
    """Accepted chains hold the proposal, rejected ones their old fields, and G
    is that of the final fields."""
    d = dict(pairing=False, u1=0.8, u2=-0.3, mu=-4.0, nlx=3, nly=3, ltrot=8, beta=1.0, nwrap=4, shift_loc=1.0)
    p = solver_params(d)
    rng = np.random.default_rng(5)
    phis = [rng.normal(size=(8, 9, 2)) for _ in range(6)]
    s = BatchedSolver([ModelSpec(p)], [0] * 6, lib=emu_lib)
    s.set_fields(phis, [11 + c for c in range(6)])
    s.prepare()
    props = GM.uniform_shift_proposals(s, rng, width=0.4)
    acc = GM.metropolis_global(s, props, np.random.default_rng(9))
    assert 0 < acc.sum() < 6                      # this seed accepts some and rejects some
    fields, seeds = s.get_fields()
    for c in range(6):
        np.testing.assert_array_equal(fields[c], props[c] if acc[c] else phis[c])
        ch = O.Chain(oracle_params(d), fields[c], 1)
        ch.pre()
        g = s.green(c)
        assert np.max(np.abs(g - ch.prop.g)) < 1e-10 * np.max(np.abs(ch.prop.g))
    assert list(seeds) == [11 + c for c in range(6)]   # chain RNG streams untouched
    s.close()


@pytest.mark.gpu
def test_global_moves_sample_the_ed_answer(hip_lib):
    """Local sweeps interleaved with global uniform-shift moves (exact MH) on the
    reference's number-conserving 3x3 U2 = -0.5 benchmark reproduce ED within
    2 standard errors, and the moves are accepted at a useful rate."""
    import json
    from pathlib import Path
    from bafqmc_amd.analysis import chain_seeds
    from bafqmc_amd.model import Params, conf_set_batch
    fx = json.loads((Path(__file__).resolve().parent / "golden" / "paper_3x3.json").read_text())
    case = fx["triangle_3x3_U2_-0.5_mu-7_b1"]
    d = case["params"]
    p = Params(pairing=False, u1=d["u1"], u2=d["u2"], mu=d["mu"], nlx=3, nly=3, ltrot=d["ltrot"], beta=d["beta"],
               nwrap=d["nwrap"], shift_loc=d["shift_loc"], ini_type=d["ini_type"], ini_ampl=d["ini_ampl"],
               is_warm=True, nwarm=d["nwarm"], nlx_therm=3, nly_therm=3, ltrot_therm=d["ltrot"]).validate()
    nch, meas = 256, 100
    phis, seeds = conf_set_batch(p, chain_seeds(777, nch))
    s = BatchedSolver([ModelSpec(p)], [0] * nch, lib=hip_lib)
    s.set_fields(list(phis), seeds)
    s.warmup(p.nwarm)
    s.prepare()
    s.sweep(50, measure=False)
    s.read_bin()
    rng = np.random.default_rng(2026)
    sums = {"density_total": np.zeros(nch), "energy_density": np.zeros(nch), "sf_K": np.zeros(nch),
            "dw_K": np.zeros(nch)}
    accepted = 0
    for _ in range(meas):
        acc = GM.metropolis_global(s, GM.uniform_shift_proposals(s, rng, width=0.1), rng)
        accepted += int(acc.sum())
        s.sweep(1)
        for c, b in enumerate(s.read_bin()):
            for k in sums:
                sums[k][c] += b[k].real if isinstance(b[k], complex) else b[k]
    s.close()
    rate = accepted / (nch * meas)
    print(f"global shift acceptance {rate:.3f}", flush=True)
    assert 0.05 < rate < 0.95
    names = {"density_total": "density_total", "energy_density": "energy_density", "sf_K": "S_SF_K",
             "dw_K": "S_DW_K"}
    for k, v in sums.items():
        x = v / meas
        m, se = float(np.mean(x)), float(np.std(x, ddof=1) / math.sqrt(nch))
        ed = case["observables"][names[k]]["ed"]
        print(f"{k}: {m:.6g} +- {se:.2g}  ED {ed:.6g}  z {(m - ed) / se:+.2f}", flush=True)
        assert abs(m - ed) <= 2.0 * se, (k, m, se, ed)


@pytest.mark.parametrize("pairing,path", [(False, -1), (True, -1), (True, 0)])
def test_grad_log_weight_central_differences(emu_lib, pairing, path):
    """bafq_grad_log_weight against central differences of log W (each from a
    fresh rebuild of G), on both channels and on both paths."""
    d = dict(pairing=pairing, u1=0.8, u2=-0.3, mu=-4.0, rdelta=0.4 if pairing else 0.0, nlx=3, nly=3, ltrot=8,
             beta=1.0, nwrap=4, shift_loc=1.0)
    p = solver_params(d)
    rng = np.random.default_rng(31)
    phis = [rng.normal(size=(8, 9, 2)) * 0.6 for _ in range(2)]
    s = BatchedSolver([ModelSpec(p)], [0, 0], lib=emu_lib, schedule=[(path, -1)])
    s.set_fields(phis, [3, 4])
    s.prepare()
    s.sweep(1, measure=False)                 # the pass must also work after a sweep
    fields, seeds = s.get_fields()
    grads = GM.grad_log_weights(s)
    after, seeds2 = s.get_fields()
    assert all(np.array_equal(a, b) for a, b in zip(fields, after)) and list(seeds) == list(seeds2)
    h = 1e-5
    for (t, i, nf) in ((0, 0, 0), (3, 4, 1), (7, 8, 1), (5, 2, 0)):
        lw = []
        for sgn in (1.0, -1.0):
            f2 = [f.copy() for f in fields]
            f2[1][t, i, nf] += sgn * h
            s.set_fields(f2, seeds)
            s.prepare()
            lw.append(GM.log_weights(s)[1])
        fd = (lw[0] - lw[1]) / (2 * h)
        assert abs(fd - grads[1][t, i, nf]) < 1e-6 * max(1.0, abs(fd)), ((t, i, nf), fd, grads[1][t, i, nf])
    s.close()
