"""Parity of the HIP solver (through the C ABI) with the oracle and the reference.

G is compared on the scale of max|G| (1e-10, the bar stated for the sweep);
observables with tests.helpers.assert_obs_close.  Pairing runs compare the
solver's reduced {b, c+} block with the oracle's full 4Ns matrices.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bafq_oracle as O
from bafqmc_amd.model import ModelSpec, conf_set_batch, conf_transfer
from bafqmc_amd.solver import BatchedSolver, STATS
from tests.helpers import (fixed_fields, load_golden, golden_row, golden_info_value, oracle_params, solver_params,
                           reduced, rel_err, assert_obs_close, obs_atol)

pytestmark = pytest.mark.gpu
G_TOL = 1e-10


def check_lane_ops(lib):
    """DPP row/quad sums, 64-bit lane broadcast and wavefront argmax (values are
    not integers, so a broadcast that truncated to 32-bit integers would fail)."""
    rng = np.random.default_rng(5)
    v = rng.normal(size=64) * 10.0 ** rng.integers(-3, 4, size=64)
    top = float(np.max(np.abs(v))) + 123.456789
    v[[7, 50]] = top               # a tie for the maximum: the lowest index wins
    idx = np.arange(64, dtype=np.int32)
    out = lib.debug_lane_ops(v, idx)
    rows = v.reshape(4, 16).sum(axis=1)
    np.testing.assert_allclose(out[0:64], np.repeat(rows, 16), rtol=1e-14, atol=1e-14)
    quads = v.reshape(16, 4).sum(axis=1)
    np.testing.assert_allclose(out[64:128], np.repeat(quads, 4), rtol=1e-14, atol=1e-14)
    assert np.all(out[128:192] == v[37])
    assert np.all(out[192:256] == top) and np.all(out[256:320] == 7)


def test_lane_ops(hip_lib):
    check_lane_ops(hip_lib)


def check_mfma_layout(lib):
    """FP64 MFMA 16x16x4: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, result
    D[i][j] in lane j + 16 (i % 4), accumulator register i // 4."""
    out = lib.debug_mfma_probe()
    seen = {}
    for p in range(64):
        ip, kp = p % 16, p // 16
        for lane in range(64):
            for r in range(4):
                row, col = lane // 16 + 4 * r, lane % 16
                want = 1000.0 + col + 16 * kp if row == ip else 0.0
                v = out[p, lane, r]
                if v != 0.0:
                    q = int(round(v - 1000.0))
                    seen[(lane, r)] = (ip, q % 16)
                if v != want:
                    table = {k: seen.get(k) for k in sorted(seen)}
                    raise AssertionError(f"MFMA layout differs at p={p} lane={lane} r={r}: {v} != {want}; "
                                         f"(lane, reg) -> (row, col) seen so far: {table}")


def test_mfma_layout(hip_lib):
    check_mfma_layout(hip_lib)


@pytest.mark.parametrize("n", [9, 18, 36, 72, 81, 144, 150, 288])
def test_device_gemm(hip_lib, n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    b = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    for oa in (0, 1):
        for ob in (0, 1):
            A = a.conj().T if oa else a
            B = b.conj().T if ob else b
            c = hip_lib.debug_gemm(a, b, oa, ob)
            assert rel_err(c, A @ B) < 1e-13


def check_qrp(lib, n, variant):
    rng = np.random.default_rng(100 + n)
    # graded columns, like the stabilisation matrices
    a = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * np.logspace(0, 8, n)[None, :]
    r, tau, perm, d = lib.debug_qrp(a, variant)
    R = np.triu(r) * d[:, None]
    Q = np.eye(n, dtype=complex)
    for k in range(n - 1, -1, -1):
        v = np.zeros(n, complex)
        v[k] = 1.0
        v[k + 1:] = r[k + 1:, k]
        Q = (np.eye(n) - tau[k] * np.outer(v, v.conj())) @ Q
    assert sorted(perm.tolist()) == list(range(n))
    assert rel_err(Q @ R, a[:, perm]) < 1e-13
    assert np.max(np.abs(Q.conj().T @ Q - np.eye(n))) < 1e-13
    assert np.all(np.diff(d) <= 1e-12 * d[0])          # pivoting orders |R_kk|
    # same |R_kk| as LAPACK's pivoted QR
    import scipy.linalg as sla
    _, rl, pl = sla.qr(a, pivoting=True)
    np.testing.assert_allclose(d, np.abs(np.diag(rl)), rtol=1e-9)
    # pivots by exact column norms: the same order as LAPACK's on these (tie-free) inputs
    np.testing.assert_array_equal(perm, pl)


@pytest.mark.parametrize("n,variant", [(9, 0), (18, 0), (72, 0), (81, 0), (144, 0), (200, 0), (288, 0),
                                       (9, 1), (18, 1), (72, 1), (81, 1), (88, 1)])
def test_device_qrp(hip_lib, n, variant):
    check_qrp(hip_lib, n, variant)


def tied_matrices(n):
    """Inputs whose column norms tie exactly (the identity, a diagonal with
    repeated magnitudes in scrambled order) and a matrix of repeated columns
    perturbed at 1e-9, numerically of rank 4.  (Exactly repeated columns make
    |R_kk| = 0, and V = D^-1 R is not finite, in the reference as here.)"""
    rng = np.random.default_rng(7 + n)
    vals = np.array([1.0, 3.0, 3.0, 0.5, 2.0, 2.0, 3.0, 1.0])[np.arange(n) % 8]
    diag = np.diag(vals * np.exp(2j * np.pi * rng.random(n)))[:, rng.permutation(n)]
    base = rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4))
    noise = 1e-9 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    return {"identity": np.eye(n, dtype=complex), "diagonal": diag, "repeated": np.tile(base, (1, n // 4)) + noise}


def check_qrp_ties(lib, n, variant):
    """Pivoted QR on exactly tied column norms: a valid factorization A P = Q R
    with nonincreasing |R_kk| equal to LAPACK's (ZGEQP3 breaks a tie to the
    lowest current position, these routines to the lowest original column, so
    the pivot order itself may differ on ties)."""
    import scipy.linalg as sla
    for name, a in tied_matrices(n).items():
        r, tau, perm, d = lib.debug_qrp(a, variant)
        R = np.triu(r) * d[:, None]
        Q = np.eye(n, dtype=complex)
        for k in range(n - 1, -1, -1):
            v = np.zeros(n, complex)
            v[k] = 1.0
            v[k + 1:] = r[k + 1:, k]
            Q = (np.eye(n) - tau[k] * np.outer(v, v.conj())) @ Q
        assert sorted(perm.tolist()) == list(range(n)), name
        assert np.max(np.abs(Q @ R - a[:, perm])) < 1e-13 * max(1.0, np.abs(a).max()), name
        assert np.max(np.abs(Q.conj().T @ Q - np.eye(n))) < 1e-13, name
        if name != "repeated":           # full rank: |R_kk| are the sorted magnitudes
            assert np.all(np.diff(d) <= 1e-13 * d[0]), name
            _, rl, _ = sla.qr(a, pivoting=True)
            np.testing.assert_allclose(d, np.abs(np.diag(rl)), rtol=1e-13, err_msg=name)
        else:                            # numerically rank 4: the first four carry the matrix
            assert np.all(np.diff(d) <= 1e-12 * d[0]), name
            assert d[3] > 1e-3 * d[0] and np.max(d[4:]) < 1e-7 * d[0], name


@pytest.mark.parametrize("n,variant", [(16, 0), (72, 0), (16, 1), (72, 1), (88, 1)])
def test_device_qrp_ties(hip_lib, n, variant):
    check_qrp_ties(hip_lib, n, variant)


CASES = [
    dict(pairing=False, u1=0.7, u2=-0.1, mu=-5.0, nlx=3, nly=3, ltrot=6, beta=0.6, nwrap=2, shift_loc=0.0),
    dict(pairing=False, u1=0.7, u2=-0.1, mu=-5.0, nlx=3, nly=3, ltrot=6, beta=0.6, nwrap=2, shift_loc=1.5),
    dict(pairing=True, u1=0.7, u2=-0.1, mu=-5.0, rdelta=0.3, nlx=3, nly=3, ltrot=6, beta=0.6, nwrap=2,
         shift_loc=0.0),
    dict(pairing=True, u1=0.7, u2=-0.1, mu=-5.0, rdelta=0.3, nlx=3, nly=3, ltrot=6, beta=0.6, nwrap=3,
         shift_loc=1.5),
    dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=6, nly=6, ltrot=20, beta=2.0, nwrap=10, shift_loc=1.5),
    dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=6, nly=6, ltrot=20, beta=2.0, nwrap=10,
         shift_loc=1.0),
    # beyond the LDS-resident size (N > 96): global-memory path
    dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=12, nly=12, ltrot=10, beta=1.0, nwrap=5, shift_loc=1.5),
    dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.3, nlx=12, nly=12, ltrot=4, beta=0.4, nwrap=2,
         shift_loc=1.0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_sweep_matches_oracle(hip_lib, case):
    check_sweep_vs_oracle(hip_lib, case)


def check_sweep_vs_oracle(lib, case):
    hip_lib = lib
    d = CASES[case]
    print(f"case {case}: {d}", flush=True)
    ns = d["nlx"] * d["nly"]
    phi = fixed_fields(d["ltrot"], ns)
    ch = O.Chain(oracle_params(d), phi, 1357911)
    ch.pre()
    s = BatchedSolver([ModelSpec(solver_params(d))], [0], lib=hip_lib)
    s.set_fields([phi], [1357911])
    s.prepare()
    go = reduced(ch.prop.g, ns) if d["pairing"] else ch.prop.g
    assert rel_err(s.green(0), go) < G_TOL
    for b in range(2):
        ref = ch.bin()
        s.sweep(1)
        mine = s.read_bin()[0]
        for k, v in ref.items():
            assert_obs_close(mine[k], v, f"case {case} bin {b} {k}", atol=obs_atol(k, d["pairing"]))
        go = reduced(ch.prop.g, ns) if d["pairing"] else ch.prop.g
        assert rel_err(s.green(0), go) < G_TOL
        f, sd = s.get_fields()
        np.testing.assert_allclose(f[0], ch.phi, rtol=0, atol=1e-13)
        assert sd[0] == ch.rng.seed
        st = mine["_stats"]
        assert (st["acc_u1"], st["acc_u2"]) == (ch.acc[0, 1], ch.acc[1, 1])
    s.close()


GOLDEN = ["nc_fixed_3x3_w1", "nc_fixed_3x3_w6", "pairing_fixed_3x3_w2", "nc_traj_3x3_c1",
          "nc_traj_3x3_u2", "pairing_traj_3x3", "nc_warm_3x3", "pairing_warm_3x3", "pairing_traj_6x6_c2",
          "nc_traj_9x9_c4", "nc_transfer_6x6", "pairing_transfer_6x6", "nc_traj_12x12", "pairing_traj_12x12_c5",
          "pairing_traj_3x3_d0", "pairing_traj_3x3_d03"]


@pytest.mark.parametrize("name", GOLDEN)
def test_sweep_matches_reference_fortran(hip_lib, name):
    check_golden(hip_lib, name)


def check_golden(lib, name):
    """Chain-averaged per-bin outputs and final fields/seeds of the reference binaries."""
    hip_lib = lib
    z, pd = load_golden(name)
    print(f"golden {name}", flush=True)
    p = solver_params(pd)
    seeds = [int(x) for x in z["seeds"]]
    if "fields_in" in z.files:
        phis, seeds_run = list(z["fields_in"]), seeds
    else:
        # conf_set on the thermalisation lattice, then conf_transfer (fields.f90:35-45)
        phis_t, seeds_run = conf_set_batch(p, seeds)
        phis = list(conf_transfer(p, phis_t))
    s = BatchedSolver([ModelSpec(p)], [0] * len(seeds), lib=hip_lib)
    s.set_fields(phis, [int(x) for x in seeds_run])
    if pd.get("nwarm", 0):
        # Local_sweep_therm sweeps; acceptance as in the reference's info.txt
        s.warmup(pd["nwarm"])
        therm = s.read_bin()
        for k, chan in ((1, "U1"), (2, "U2")):
            if abs(pd[f"u{k}"]) < 1e-10:
                continue                   # the reference reports coupled channels only
            want = golden_info_value(z, f"Thermalize {chan} Accept Ratio")
            prop = sum(c["_stats"][f"therm_prop_u{k}"] for c in therm)
            acc = sum(c["_stats"][f"therm_acc_u{k}"] for c in therm)
            assert abs(acc / prop - want) < 1e-12, (name, chan, acc / prop, want)
    s.prepare()
    for b in range(pd["nbin"]):
        s.sweep(1)
        per = s.read_bin()
        for key in z.files:
            if not key.startswith("out_"):
                continue
            k = key[4:]
            if k not in per[0]:
                continue
            mine = sum(c[k] for c in per) / len(per)
            assert_obs_close(mine, golden_row(z, k, b), f"{name} bin {b} {k}", atol=obs_atol(k, pd["pairing"]))
    f, sd = s.get_fields()
    for c in range(len(seeds)):
        np.testing.assert_allclose(f[c], z["fields_out"][c], rtol=0, atol=1e-12)
        assert sd[c] == int(z["seeds_out"][c])
    # compare_mat drift (stabilization.f90:256-266) against the reference's
    # "Max diff Matrix" (main.f90:59-73): rounding of two different codes, so an
    # order of magnitude; below 1e-8 both are rounding noise of G
    drift = float(np.max(s.stats()[:, 4]))
    ref_drift = golden_info_value(z, "Max diff Matrix")
    print(f"golden {name}: drift {drift:.3e} reference {ref_drift:.3e}", flush=True)
    if ref_drift >= 1e-8:
        assert ref_drift / 10 <= drift <= 10 * ref_drift, (name, drift, ref_drift)
    else:
        assert drift < 1e-8, (name, drift, ref_drift)
    if not pd["pairing"]:
        # log_weight = -1/2 sum phi^2 + 2 log|det G| (pole_diagnostics.f90:130-131)
        for c in range(len(seeds)):
            sv = np.linalg.svd(s.green(c), compute_uv=False)
            lw = -0.5 * float(np.sum(f[c] ** 2)) + 2.0 * float(np.sum(np.log(sv)))
            if len(seeds) == 1:
                np.testing.assert_allclose(lw, z["out_log_weight"][-1][0], rtol=1e-10)
    s.close()


def check_blowup(lib):
    """Pairing with U2 < 0 has no finite trace (docs/algorithm.md:152-156): the
    golden chain (3x3, U1 = 1, U2 = -0.5, Delta = 0.5, mu = -5, beta = 1) reaches
    a pole of (1 - B)^-1 in its 15th sweep in the reference Fortran, where its
    density jumps from 0.35 to 1.5e3.  The same seed here is the same Markov
    chain: the same rows before the jump (1e-9), the jump in the same sweep, and
    after it the same accept/reject sequence (rows to 1e-4: the drift of the
    stabilisation is O(1) there), final fields and seed."""
    z, pd = load_golden("pairing_blowup_3x3")
    p = solver_params(pd)
    phis, seeds = conf_set_batch(p, [int(x) for x in z["seeds"]])
    s = BatchedSolver([ModelSpec(p)], [0], lib=lib)
    s.set_fields(list(phis), seeds)
    s.prepare()
    ref = z["out_density_total"][:, 0]
    jump = int(np.argmax(np.abs(ref) > 1.0))
    assert jump == 14 and np.all(np.abs(ref[:jump]) < 1.0)
    for b in range(pd["nbin"]):
        s.sweep(1)
        got = s.read_bin()[0]["density_total"]
        rtol = 1e-9 if b < jump else 1e-4
        assert abs(got - ref[b]) <= rtol * abs(ref[b]), (b, got, ref[b])
    f, sd = s.get_fields()
    assert sd[0] == int(z["seeds_out"][0])
    np.testing.assert_allclose(f[0], z["fields_out"][0], rtol=0, atol=1e-9)
    assert float(s.stats()[0, 4]) > 5.5e-5        # the drift flags it, as the reference's Max diff Matrix does
    assert golden_info_value(z, "Max diff Matrix") > 5.5e-5
    s.close()


def test_blowup_at_negative_u2_is_the_references(hip_lib):
    check_blowup(hip_lib)


@pytest.mark.parametrize("delayed", [-1, 0])
@pytest.mark.parametrize("name", ["nc_traj_3x3_u2", "pairing_traj_3x3", "pairing_traj_6x6_c2"])
def test_global_memory_path_matches_reference(hip_lib, name, delayed, monkeypatch):
    """The device-memory path (used for N > 96) on the golden cases, with
    delayed (-1: widest) and per-site (0) local updates."""
    monkeypatch.setattr(BatchedSolver, "default_schedule", (0, delayed))
    check_golden(hip_lib, name)


@pytest.mark.parametrize("width", [0, 4])
@pytest.mark.parametrize("name", ["nc_traj_3x3_u2", "pairing_traj_3x3", "pairing_traj_6x6_c2"])
def test_local_update_variants_match_reference(hip_lib, name, width, monkeypatch):
    """Per-site rank updates (0) and delayed updates flushed every 4 columns (4)."""
    monkeypatch.setattr(BatchedSolver, "default_schedule", (1, width))
    check_golden(hip_lib, name)


@pytest.mark.parametrize("width", [20, 32])
@pytest.mark.parametrize("name", ["pairing_traj_6x6_c2", "nc_traj_9x9_c4", "pairing_traj_12x12_c5"])
def test_wide_device_memory_updates_match_reference(hip_lib, name, width, monkeypatch):
    """Delayed-update buffers wider than 16 columns (device-memory path only,
    buffers in the chain's work matrices): the reference's trajectory."""
    monkeypatch.setattr(BatchedSolver, "default_schedule", (0, width))
    check_golden(hip_lib, name)


@pytest.mark.parametrize("case,width,nchains", [(4, 16, 3), (5, 32, 2), (6, 16, 2), (7, 32, 3)])
def test_paired_workgroups_equal_single(hip_lib, case, width, nchains, monkeypatch):
    """Two workgroups per chain (device-memory path): the same chain, bit for bit."""
    check_pairs_bitwise(hip_lib, case, width, nchains, monkeypatch)


@pytest.mark.parametrize("mixed", [False, True])
def test_uneven_batch_equals_single_chains(hip_lib, mixed):
    check_uneven_batch(hip_lib, mixed)


def test_uneven_batch_across_paths_equals_single_chains(hip_lib):
    """Lattices from 3x3 to 12x12 (N = 9 ... 288) in one handle with the
    library's own schedules: the 12x12 chains run on the device-memory path,
    the others in LDS, as two persistent kernels of one launch graph; every
    chain ends bit for bit as when run alone."""
    specs = [CASES[1], CASES[4], CASES[6], CASES[7], CASES[5]]
    models = [ModelSpec(solver_params(d)) for d in specs]
    chain_model = [2, 0, 3, 1, 4, 2, 0, 3]
    phis = [fixed_fields(specs[m]["ltrot"], specs[m]["nlx"] * specs[m]["nly"]) for m in chain_model]
    seeds = [2222 + 11 * c for c in range(len(chain_model))]
    s = BatchedSolver(models, chain_model, lib=hip_lib)
    assert {p for p, _ in s.schedule()} == {0, 1}
    s.set_fields(phis, seeds)
    s.prepare()
    s.sweep(2)
    batch = s.read_bin()
    gb = [s.green(c) for c in range(len(chain_model))]
    fb, sb = s.get_fields()
    s.close()
    for c, m in enumerate(chain_model):
        s1 = BatchedSolver([models[m]], [0], lib=hip_lib)
        s1.set_fields([phis[c]], [seeds[c]])
        s1.prepare()
        s1.sweep(2)
        one = s1.read_bin()[0]
        for k in ("density_total", "energy_density", "sf_K", "dw_K"):
            assert batch[c][k] == one[k], (c, k)
        assert np.array_equal(gb[c], s1.green(0))
        f1, s1d = s1.get_fields()
        assert np.array_equal(fb[c], f1[0]) and sb[c] == s1d[0]
        s1.close()


def test_uneven_workload_batch_equals_single_chains(hip_lib):
    """The uneven batch at its full size: 256 number-conserving chains, 16 each
    of lattices 3x3, 6x6, 9x9, 12x12 x beta 4, 8 (dtau 0.1, Ltrot 40, 80) x
    U2 0, -0.5 (U1 = 1, mu = -5), in one handle with schedules from the
    autotuner; after two sweeps one chain of every model (16 chains) equals a
    run of that chain alone under the same schedule, bit for bit (observables,
    G, fields, seed)."""
    from bafqmc_amd import autotune
    from bafqmc_amd.analysis import chain_seeds
    params = []
    for L in (3, 6, 9, 12):
        for beta in (4.0, 8.0):
            for u2 in (0.0, -0.5):
                params.append(solver_params(dict(pairing=False, u1=1.0, u2=u2, mu=-5.0, nlx=L, nly=L,
                                                 ltrot=int(round(beta / 0.1)), beta=beta, nwrap=10, shift_loc=1.5,
                                                 ini_type=2, ini_ampl=1.0)))
    chain_model = np.arange(256) % len(params)
    choices = autotune.tune(hip_lib, params, chain_model, sweeps=1, max_sample=16)
    schedules = [c.schedule for c in choices]
    print("uneven workload schedules:", schedules, flush=True)
    assert {sc[0] for sc in schedules} == {0, 1}          # both paths in one launch graph
    models = [ModelSpec(p) for p in params]
    phis, seeds = [], []
    for c in range(256):
        f, sd = conf_set_batch(params[chain_model[c]], chain_seeds(777 + c, 1))
        phis.append(f[0])
        seeds.append(sd[0])
    s = BatchedSolver(models, chain_model, lib=hip_lib, schedule=schedules)
    s.set_fields(phis, seeds)
    s.prepare()
    s.sweep(2)
    batch = s.read_bin()
    sample = [int(np.flatnonzero(chain_model == m)[m % 16]) for m in range(len(params))]
    gb = {c: s.green(c) for c in sample}
    fb, sb = s.get_fields()
    s.close()
    for c in sample:
        m = chain_model[c]
        s1 = BatchedSolver([models[m]], [0], lib=hip_lib, schedule=[schedules[m]])
        s1.set_fields([phis[c]], [seeds[c]])
        s1.prepare()
        s1.sweep(2)
        one = s1.read_bin()[0]
        for k in ("density_total", "energy_density", "sf_K", "dw_K", "kinetic"):
            assert np.array_equal(batch[c][k], one[k], equal_nan=True), (c, m, k)
        assert np.array_equal(gb[c], s1.green(0), equal_nan=True), (c, m)
        f1, s1d = s1.get_fields()
        assert np.array_equal(fb[c], f1[0]) and sb[c] == s1d[0], (c, m)
        s1.close()


def check_pairs_bitwise(lib, case, width, nchains, monkeypatch, sweeps=2):
    """Two workgroups per chain on the device-memory path (BAFQ_PAIRS=1) give
    the same chain as one workgroup per chain (BAFQ_PAIRS=0), bit for bit:
    fields, seeds, G, every estimator and every counter except the timers,
    over the rebuild, the sweeps and the three measurement passes."""
    d = CASES[case] if isinstance(case, int) else case
    ns = d["nlx"] * d["nly"]
    phis = [fixed_fields(d["ltrot"], ns) * (1.0 + 0.05 * k) for k in range(nchains)]
    seeds = [24680 + 11 * k for k in range(nchains)]
    timers = [v for k, v in STATS.items() if k.startswith("t_")]
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("BAFQ_PAIRS", mode)
        s = BatchedSolver([ModelSpec(solver_params(d))], [0] * nchains, lib=lib, schedule=[(0, width)])
        if mode == "0":
            assert s.pairs() == 0
        else:
            assert s.pairs() >= 1
        s.set_fields(phis, seeds)
        s.prepare()
        for _ in range(sweeps):
            s.sweep(1)
        for pm in ("checkpoints", "all", "current"):
            s.measure_pass(1, mode=pm)
        obs, st = s.read_bin_raw()
        st = np.delete(np.asarray(st), timers, axis=1)
        phi, sd = s.get_fields()
        out[mode] = ([s.green(c) for c in range(nchains)], np.asarray(obs), st, phi, np.asarray(sd))
        s.close()
    a, b = out["0"], out["1"]
    for c in range(nchains):
        assert np.array_equal(a[0][c], b[0][c]), f"G of chain {c}"
        assert np.array_equal(a[3][c], b[3][c]), f"fields of chain {c}"
    for k in (1, 2, 4):
        assert np.array_equal(a[k], b[k])


def check_uneven_batch(hip_lib, mixed=False, light=False):
    """Chains of different lattices/solvers in one launch evolve exactly as alone
    (mixed: half of the models on the device-memory path, so both persistent
    kernels run in one launch, against the single chains on their own paths;
    light: fewer chains, for the CPU emulator)."""
    specs = [CASES[1], CASES[3], CASES[4], CASES[5]]
    models = [ModelSpec(solver_params(d)) for d in specs]
    chain_model = [0, 1, 2, 3, 0] if light else [0, 1, 2, 3, 2, 0, 1, 3, 3]
    sched = [(0, -1), (1, -1), (0, -1), (1, 4)] if mixed else None
    phis = [fixed_fields(specs[m]["ltrot"], specs[m]["nlx"] * specs[m]["nly"]) for m in chain_model]
    seeds = [1111 + 7 * c for c in range(len(chain_model))]
    s = BatchedSolver(models, chain_model, lib=hip_lib, schedule=sched)
    s.set_fields(phis, seeds)
    s.prepare()
    s.sweep(1)
    s.sweep(1)          # replays the captured launch
    batch = s.read_bin()
    gb = [s.green(c) for c in range(len(chain_model))]
    s.close()
    for c, m in enumerate(chain_model):
        s1 = BatchedSolver([models[m]], [0], lib=hip_lib, schedule=[sched[m]] if sched else None)
        s1.set_fields([phis[c]], [seeds[c]])
        s1.prepare()
        s1.sweep(2)
        one = s1.read_bin()[0]
        for k in ("density_total", "energy_density", "sf_K", "dw_K"):
            assert batch[c][k] == one[k]
        assert np.array_equal(gb[c], s1.green(0))
        s1.close()


@pytest.mark.parametrize("pairing,delta", [(False, 0.0), (True, 0.2)])
def test_gaussian_closed_form_gpu(hip_lib, pairing, delta):
    check_gaussian(hip_lib, pairing, delta)


def check_gaussian(hip_lib, pairing, delta):
    d = dict(pairing=pairing, u1=0.0, u2=0.0, mu=-5.0, rdelta=delta, nlx=3, nly=3, ltrot=100, beta=1.0,
             nwrap=10, shift_loc=1.0)
    s = BatchedSolver([ModelSpec(solver_params(d))], [0], lib=hip_lib)
    s.set_fields([np.zeros((100, 9, 2))], [24681357])
    s.prepare()
    s.sweep(1)
    b = s.read_bin()[0]
    exact = O.gaussian_closed_form(delta, 1.0, -5.0)
    for k, v in exact.items():
        if k == "pair_equal" and not pairing:
            continue
        got = b[k].real if isinstance(b[k], complex) else b[k]
        assert abs(got - v) <= 2e-9 + 2e-8 * abs(v), (k, got, v)
    s.close()
