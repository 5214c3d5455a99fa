"""The time-displaced pass (the reference's is_tau path).

The reference runs Dyn_reset + Dyn_sweep_R between sweep_L and sweep_R of
every bin after Nthermal (src/number_conserving/src/local_sweep.f90:119-123,
dynamics.f90:38-57); its estimators are empty (obser_tau.f90:39-54), so its
only output is info.txt's six Wrap_tau drift lines (dynamics.f90:60-82).

Pinning of the oracle (oracle/bafq_oracle.py, Chain.tau_pass, stab_green_big):
* with the reference's conventions (+D_L block, U2 propagated with the U1
  field) it reproduces the six drift lines of three Fortran runs built from the
  reference's sources (tests/golden/{nc_tau_3x3,nc_tau_6x6_u1,pairing_tau_3x3}.npz)
  -- numbers of O(1e-5 .. 1e4) that depend on every propagation step, the
  block layout, the Wrap_tau schedule and the normalisation;
* with the bosonic conventions every slice's G(0,0), G(tau,0), G(0,tau),
  G(tau,tau) equals the closed form from explicit products
  (oracle.dense_green_tau), which a fermionic sign, a wrong field, a swapped
  block or a wrong propagation order all fail.
The device pass (bafq_sweep_tau) is compared with the oracle on the CPU
emulator and on the GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

from bafqmc_amd import driver
from bafqmc_amd.model import ModelSpec, conf_set_batch
from bafqmc_amd.solver import BatchedSolver, TAU_BOSONIC, TAU_REFERENCE
from oracle import bafq_oracle as O
from oracle import run_ref
from tests.helpers import load_golden, golden_info_value, oracle_params, solver_params, reduced

GOLDEN_TAU = ["nc_tau_3x3", "nc_tau_6x6_u1", "pairing_tau_3x3"]
LABELS = ("GRtt", "GRt0", "GR0t")     # the reference's labels of the G0t, Gt0, Gtt drifts


def _oracle_golden_run(name, sign, u2_field):
    z, pd = load_golden(name)
    p = oracle_params(pd)
    chains = []
    for s in z["seeds"]:
        r = O.Ranf(int(s))
        phi = O.conf_set(r, pd.get("ini_type", 2), pd.get("ini_ampl", 1.0), (0.0, 0.0), p.nlx * p.nly, p.ltrot)
        ch = O.Chain(p, phi, r.seed)
        ch.pre()
        chains.append(ch)
    for b in range(pd["nbin"]):
        for ch in chains:
            ch.bin(tau=b + 1 > pd["nthermal"], tau_sign=sign, tau_u2_field=u2_field)
    n = len(chains) * pd["nbin"] * p.nst
    return z, pd, chains, np.max([c.tau_xmax for c in chains], axis=0), np.sum([c.tau_xmean for c in chains], 0) / n


@pytest.mark.parametrize("name", GOLDEN_TAU)
def test_oracle_reproduces_reference_tau_drift_lines(name):
    z, pd, chains, xmax, xmean = _oracle_golden_run(name, O.REFERENCE_TAU_SIGN, 1)
    for k, lab in enumerate(LABELS):
        for got, label in ((xmax[k], f"Max diff {lab}"), (xmean[k], f"Mean diff {lab}")):
            ref = golden_info_value(z, label)
            assert abs(got - ref) <= 1e-7 * abs(ref), (label, got, ref)
    # the pass leaves the chain alone: the reference's final fields and seeds
    for c, ch in enumerate(chains):
        assert ch.rng.seed == int(z["seeds_out"][c])
        np.testing.assert_allclose(ch.phi, z["fields_out"][c], rtol=0, atol=1e-13)


def test_oracle_bosonic_conventions_recompute_what_they_propagate():
    """With +D_L or the U1 field in the U2 operator (the reference's code) the
    recomputed functions differ from the propagated ones by O(1); with the
    bosonic conventions by rounding."""
    for name in ("nc_tau_3x3", "pairing_tau_3x3"):
        _, _, _, xmax_b, _ = _oracle_golden_run(name, O.BOSONIC_TAU_SIGN, 2)
        assert np.all(xmax_b < 1e-8), (name, xmax_b)
        _, _, _, xmax_s, _ = _oracle_golden_run(name, O.REFERENCE_TAU_SIGN, 2)
        assert np.max(xmax_s) > 1e-2, (name, xmax_s)
        _, _, _, xmax_f, _ = _oracle_golden_run(name, O.BOSONIC_TAU_SIGN, 1)
        assert np.max(xmax_f) > 1e-4, (name, xmax_f)


@pytest.mark.parametrize("pairing,u2,beta,tol", [(False, -0.5, 2.0, 1e-11), (False, 0.0, 2.0, 1e-11),
                                                 (True, -0.5, 1.0, 1e-8)])
def test_oracle_tau_pass_equals_dense_closed_form(pairing, u2, beta, tol):
    """Every slice's four functions against explicit products (the pairing
    closed form is limited by the conditioning of its dense 4Ns products)."""
    d = dict(pairing=pairing, u1=1.0, u2=u2, mu=-5.0, nlx=3, nly=3, ltrot=20, beta=beta, nwrap=5, shift_loc=1.0,
             rdelta=0.5 if pairing else 0.0)
    rng = np.random.default_rng(11)
    ch = O.Chain(oracle_params(d), rng.normal(size=(20, 9, 2)), 4242)
    ch.pre()
    ch.sweep_L()
    rec = []
    ch.tau_pass(O.BOSONIC_TAU_SIGN, rec)
    assert len(rec) == 20
    for tau, got in enumerate(rec):
        for k, (a, b) in enumerate(zip(got, O.dense_green_tau(ch, tau))):
            err = np.max(np.abs(a - b)) / np.max(np.abs(b))
            assert err < tol, (tau, k, err)


def _device_vs_oracle(lib, d, mode, phi, seed, schedule=None, tol=1e-10):
    """One sweep with the pass on the device and on the oracle: records at every
    slice, drift maxima, and the chain (observables, fields, seed) of a plain sweep."""
    p = oracle_params(d)
    ch = O.Chain(p, phi, seed)
    ch.pre()
    rec = []
    sign, fld = (O.BOSONIC_TAU_SIGN, 2) if mode == TAU_BOSONIC else (O.REFERENCE_TAU_SIGN, 1)
    ch.bin(tau=True, tau_sign=sign, record=rec, tau_u2_field=fld)
    sp = solver_params(d)
    s = BatchedSolver([ModelSpec(sp)], [0], lib=lib, schedule=schedule)
    plain = BatchedSolver([ModelSpec(sp)], [0], lib=lib, schedule=schedule)
    try:
        for x in (s, plain):
            x.set_fields([phi], [seed])
            x.prepare()
            x.read_bin()
        s.sweep_tau(1, measure=True, mode=mode, record=True)
        plain.sweep(1, measure=True)
        got, ref = s.read_bin()[0], plain.read_bin()[0]
        gt = s.green_tau(0)
        ns = p.nlx * p.nly
        for t in range(p.ltrot):
            # on the scale of the equal-time G (G(0,tau) starts at G - 1, whose
            # entries can be 1e-2 of G's while its rounding is G's)
            scale = np.max(np.abs(rec[t][0]))
            for k in range(4):
                want = reduced(rec[t][k], ns) if p.pairing else rec[t][k]
                err = np.max(np.abs(gt[t, k] - want)) / max(np.max(np.abs(want)), scale)
                assert err < tol, (t, k, err)
        st = got["_stats"]
        dev = np.array([st["tau_xmax_0t"], st["tau_xmax_t0"], st["tau_xmax_tt"]])
        if mode == TAU_REFERENCE:
            np.testing.assert_allclose(dev, ch.tau_xmax, rtol=1e-8)
        else:
            # rounding, summed over the N x N entries (1e-8 at the 12x12 pairing lattice)
            assert np.all(dev < 1e-6) and np.all(ch.tau_xmax < 1e-6), (dev, ch.tau_xmax)
        assert st["tau_nwrap"] == p.nst
        # the chain is untouched by the pass: the same fields, seeds and estimators
        f1, s1 = s.get_fields()
        f2, s2 = plain.get_fields()
        assert list(s1) == list(s2)
        np.testing.assert_array_equal(f1[0], f2[0])
        for key in ("density_total", "energy_density", "sf_K", "dw_K"):
            assert got[key] == ref[key], key
        np.testing.assert_allclose(s.green(0), plain.green(0), rtol=0, atol=0)
    finally:
        s.close()
        plain.close()


EMU_CASES = [
    (dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=3, nly=3, ltrot=20, beta=2.0, nwrap=5, shift_loc=1.5), None),
    (dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=3, nly=3, ltrot=20, beta=2.0, nwrap=5,
          shift_loc=1.0), None),
    (dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=3, nly=3, ltrot=20, beta=2.0, nwrap=10, shift_loc=1.5),
     (0, 4)),                                                  # device-memory path, narrow delayed updates
    # N = 36: the 72 x 72 block matrix takes the blocked reflector application and solve
    (dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=6, nly=6, ltrot=10, beta=1.0, nwrap=5, shift_loc=1.5), None),
]


@pytest.mark.parametrize("mode", [TAU_BOSONIC, TAU_REFERENCE])
@pytest.mark.parametrize("case", range(len(EMU_CASES)))
def test_emu_tau_pass_matches_oracle(emu_lib, case, mode):
    d, schedule = EMU_CASES[case]
    rng = np.random.default_rng(3 + case)
    ns = d["nlx"] * d["nly"]
    _device_vs_oracle(emu_lib, d, mode, rng.normal(size=(d["ltrot"], ns, 2)), 4242 + case, schedule)


def _write_case(tmp_path, name):
    z, pd = load_golden(name)
    d = tmp_path / name
    run_ref.write_inputs(d, dict(pd), [int(s) for s in z["seeds"]])
    return z, pd, d


def _check_tau_info(z, pd, d, mode):
    info = (d / "info.txt").read_text()

    def rows(text):
        return [line.split(":")[0].strip() for line in text.splitlines() if ":" in line]
    assert rows(info) == rows(str(z["info"]))
    for lab in LABELS:
        for label in (f"Max diff {lab}", f"Mean diff {lab}"):
            mine, ref = golden_info_value({"info": info}, label), golden_info_value(z, label)
            if mode == TAU_REFERENCE:
                assert abs(mine - ref) <= 1e-7 * abs(ref), (label, mine, ref)
            else:
                assert mine < 1e-8, (label, mine)
    # the equal-time outputs are the reference's (the pass changes no field)
    for key in z.files:
        if key.startswith("out_") and key[4:] in ("density_total", "energy_density"):
            got = run_ref.read_rows(d / key[4:])
            np.testing.assert_allclose(got, z[key], rtol=1e-9, atol=1e-13)
    toks = (d / "confout.txt").read_text().split()
    assert int(toks[0]) == int(z["seeds_out"][0])


@pytest.mark.parametrize("mode", [TAU_REFERENCE, TAU_BOSONIC])
def test_driver_writes_tau_drift_lines_emulator(tmp_path, emu_lib, mode):
    z, pd, d = _write_case(tmp_path, "nc_tau_3x3")
    driver.run(d, "number_conserving", len(z["seeds"]), lib=emu_lib, poles=False, tau_mode=mode)
    _check_tau_info(z, pd, d, mode)


@pytest.mark.gpu
@pytest.mark.parametrize("name,solver", [("nc_tau_3x3", "number_conserving"), ("nc_tau_6x6_u1", "number_conserving"),
                                         ("pairing_tau_3x3", "pairing")])
@pytest.mark.parametrize("mode", [TAU_REFERENCE, TAU_BOSONIC])
def test_driver_writes_tau_drift_lines_gpu(tmp_path, hip_lib, name, solver, mode):
    z, pd, d = _write_case(tmp_path, name)
    driver.run(d, solver, len(z["seeds"]), lib=hip_lib, poles=False, tau_mode=mode)
    _check_tau_info(z, pd, d, mode)


GPU_CASES = EMU_CASES + [
    # the pairing 6x6 lattice (N = 72, LDS path; its 2N x 2N block matrix 144 x 144)
    (dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=6, nly=6, ltrot=20, beta=1.0, nwrap=10,
          shift_loc=1.0), None),
    # NC 9x9 (N = 81) on the device-memory path
    (dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=9, nly=9, ltrot=20, beta=2.0, nwrap=10, shift_loc=1.5),
     (0, -1)),
    # the 12x12 pairing lattice (N = 288, device-memory path; its 576 x 576 block
    # matrix takes the QR's general loop and the blocked solve)
    (dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.3, nlx=12, nly=12, ltrot=4, beta=0.4, nwrap=2,
          shift_loc=1.0), None),
]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [TAU_BOSONIC, TAU_REFERENCE])
@pytest.mark.parametrize("case", range(len(GPU_CASES)))
def test_gpu_tau_pass_matches_oracle(hip_lib, case, mode):
    d, schedule = GPU_CASES[case]
    ns = d["nlx"] * d["nly"]
    rng = np.random.default_rng(30 + case)
    print(f"tau pass case {case} mode {mode}: {d}", flush=True)
    _device_vs_oracle(hip_lib, d, mode, rng.normal(size=(d["ltrot"], ns, 2)), 9090 + case, schedule,
                      tol=1e-9 if d["pairing"] else 1e-10)


@pytest.mark.gpu
def test_gpu_tau_pass_in_an_uneven_batch(hip_lib):
    """Chains of several sizes and both paths in one launch give each chain's
    single-run time-displaced functions and drifts bit for bit."""
    picks = (0, 1, 4, 5)               # NC 3x3, pairing 3x3, pairing 6x6 (LDS path), NC 9x9 (device memory)
    models = [solver_params(GPU_CASES[k][0]) for k in picks]
    sched = [(-1, -1), (-1, -1), (-1, -1), (0, -1)]
    chain_model = [0, 2, 1, 3, 0, 2]
    phis, seeds = [], []
    for c, k in enumerate(chain_model):
        ph, sd = conf_set_batch(models[k], [1000 + 17 * c])
        phis += list(ph)
        seeds += list(sd)
    s = BatchedSolver([ModelSpec(p) for p in models], chain_model, lib=hip_lib, schedule=sched)
    try:
        s.set_fields(phis, seeds)
        s.prepare()
        s.sweep_tau(2, measure=True, mode=TAU_BOSONIC, record=True)
        batch = [s.green_tau(c) for c in range(len(chain_model))]
        st_b = s.read_bin()
    finally:
        s.close()
    for c, k in enumerate(chain_model):
        one = BatchedSolver([ModelSpec(models[k])], [0], lib=hip_lib, schedule=[sched[k]])
        try:
            one.set_fields([phis[c]], [seeds[c]])
            one.prepare()
            one.sweep_tau(2, measure=True, mode=TAU_BOSONIC, record=True)
            np.testing.assert_array_equal(one.green_tau(0), batch[c])
            assert one.read_bin()[0]["_stats"]["tau_xmax_tt"] == st_b[c]["_stats"]["tau_xmax_tt"]
        finally:
            one.close()
