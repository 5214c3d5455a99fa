"""The HIP kernels on the CPU emulator back end (fibers per workgroup).

Same device code and workgroup sizes as the gfx950 build, run on the same
cases as tests/test_gpu_parity.py (except the largest).  The emulator runs the
threads of a workgroup one after another between barriers, so a read that is
not separated by a barrier from the write it needs shows up as a wrong result
here even where GPU timing hides it.  Lane-level builtins (DPP, MFMA layout)
have their own GPU tests.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import rel_err, fixed_fields, solver_params
from tests.test_gpu_parity import CASES
from bafqmc_amd.model import ModelSpec
from bafqmc_amd.solver import BatchedSolver
from tests.test_gpu_parity import (check_sweep_vs_oracle, check_golden, check_mfma_layout, check_lane_ops, check_qrp,
                                   check_qrp_ties, check_pairs_bitwise,
                                   check_uneven_batch, check_gaussian, check_blowup, GOLDEN)


def test_emu_mfma_layout(emu_lib):
    check_mfma_layout(emu_lib)


def test_emu_gemm(emu_lib):
    rng = np.random.default_rng(3)
    for n in (5, 18, 33, 150):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        b = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        for oa in (0, 1):
            for ob in (0, 1):
                A = a.conj().T if oa else a
                B = b.conj().T if ob else b
                assert rel_err(emu_lib.debug_gemm(a, b, oa, ob), A @ B) < 1e-13


@pytest.mark.parametrize("n,variant", [(9, 0), (18, 0), (33, 0), (150, 0), (9, 1), (18, 1), (32, 1)])
def test_emu_qrp(emu_lib, n, variant):
    check_qrp(emu_lib, n, variant)


@pytest.mark.parametrize("n,variant", [(16, 0), (16, 1), (72, 1)])
def test_emu_qrp_ties(emu_lib, n, variant):
    check_qrp_ties(emu_lib, n, variant)


@pytest.mark.parametrize("case", range(8))
def test_emu_sweep_matches_oracle(emu_lib, case):
    """Case 7 (12x12 pairing, N = 288) keeps its delayed-update buffers in device
    memory: they do not fit beside the rest in LDS at width 16."""
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("delayed", [-1, 0])
@pytest.mark.parametrize("case", [1, 3, 5])
def test_emu_global_path_matches_oracle(emu_lib, case, delayed, monkeypatch):
    monkeypatch.setattr(BatchedSolver, "default_schedule", (0, delayed))
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("width", [20, 32])
@pytest.mark.parametrize("case", [4, 5])
def test_emu_wide_device_memory_updates(emu_lib, case, width, monkeypatch):
    """Delayed-update buffers wider than 16 columns in device memory: the
    decisions sum the correction over two blocks of 16, the flush runs its
    MFMA k-steps over both."""
    monkeypatch.setattr(BatchedSolver, "default_schedule", (0, width))
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("case,width,nchains", [(3, 8, 3), (4, 16, 2), (5, 32, 1)])
def test_emu_paired_workgroups_equal_single(emu_lib, case, width, nchains, monkeypatch):
    """Two workgroups per chain, run as concurrent OS threads; with 3 chains
    the emulator's 3 resident workgroups hold one pair, which takes the chains
    one after another from the queue."""
    check_pairs_bitwise(emu_lib, case, width, nchains, monkeypatch)


def test_emu_paired_workgroups_split_stabilisation(emu_lib, monkeypatch):
    """N = 81: large enough for the pair to split the stabilisation (blocked
    reflectors, both members' work buffers), the QR factors passed from the
    leader to its partner through device memory."""
    d = dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=9, nly=9, ltrot=4, beta=0.4, nwrap=2, shift_loc=1.5)
    check_pairs_bitwise(emu_lib, d, 16, 2, monkeypatch)


def test_emu_pair_wait_timeout_is_an_error(emu_lib, monkeypatch):
    """A pair barrier's bounded wait: with a zero limit the first wait that is
    not met at once sets the pair's flag, both members stop waiting, the kernel
    ends and the sweep returns an error instead of hanging."""
    from bafqmc_amd.solver import SolverError
    monkeypatch.setenv("BAFQ_PAIRS", "1")
    monkeypatch.setenv("BAFQ_PAIR_WAIT_MS", "0")
    d = CASES[5]
    s = BatchedSolver([ModelSpec(solver_params(d))], [0], lib=emu_lib, schedule=[(0, 32)])
    try:
        assert s.pairs() == 1
        s.set_fields([fixed_fields(d["ltrot"], d["nlx"] * d["nly"])], [4242])
        with pytest.raises(SolverError, match="BAFQ_PAIR_WAIT_MS"):
            s.prepare()
            s.sweep(2)
    finally:
        s.close()


def test_wide_updates_rejected_on_the_lds_path(emu_lib):
    with pytest.raises(RuntimeError, match="above 16 need the device-memory path"):
        BatchedSolver([ModelSpec(solver_params(CASES[4]))], [0], lib=emu_lib, schedule=(1, 32))


# the 12x12 goldens (N = 144, 288) run on the GPU only: minutes on the emulator;
# of the two Ltrot = 400 single-channel pairing goldens the emulator runs one
EMU_GOLDEN = [g for g in GOLDEN if "12x12" not in g and g != "pairing_traj_3x3_d0"]


@pytest.mark.parametrize("name", EMU_GOLDEN)
def test_emu_matches_reference_fortran(emu_lib, name):
    check_golden(emu_lib, name)


def test_emu_blowup_at_negative_u2_is_the_references(emu_lib):
    check_blowup(emu_lib)


@pytest.mark.parametrize("width", [0, 4, 8])
@pytest.mark.parametrize("case", [3, 5])
def test_emu_local_update_variants(emu_lib, case, width, monkeypatch):
    monkeypatch.setattr(BatchedSolver, "default_schedule", (1, width))
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("mixed", [False, True])
def test_emu_uneven_batch(emu_lib, mixed):
    check_uneven_batch(emu_lib, mixed, light=True)


@pytest.mark.parametrize("pairing,delta", [(False, 0.0), (True, 0.2)])
def test_emu_gaussian_closed_form(emu_lib, pairing, delta):
    check_gaussian(emu_lib, pairing, delta)


def test_emu_lane_ops(emu_lib):
    check_lane_ops(emu_lib)


@pytest.mark.parametrize("case,path", [(3, -1), (5, -1), (4, 0)])
def test_emu_thread_order_independence(emu_lib, case, path, monkeypatch):
    """Race check: the same sweep with the threads of every workgroup run in
    ascending and in descending order between barriers gives identical bits."""
    monkeypatch.setattr(BatchedSolver, "default_schedule", (path, -1))
    d = CASES[case]
    ns = d["nlx"] * d["nly"]
    phi = fixed_fields(d["ltrot"], ns)
    out = []
    for rev in ("0", "1"):
        monkeypatch.setenv("BAFQ_EMU_REVERSE", rev)
        s = BatchedSolver([ModelSpec(solver_params(d))], [0], lib=emu_lib)
        s.set_fields([phi], [1357911])
        s.prepare()
        s.sweep(1)
        obs, st = s.read_bin_raw()
        out.append((s.green(0), obs, s.get_fields()[0][0]))
        s.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_emu_wavefronts_agree_on_a_divergent_chain(emu_lib):
    """Regression: with mu above the band bottom the chain's G grows without
    bound (drift ~1e6 after one sweep), and the decisions of trailing rejected
    sites are sensitive to every bit of G.  A wavefront's final delayed-update
    flush used to overwrite G while other wavefronts were still deciding those
    sites; their decisions then differed and so did their barrier counts (the
    emulator reports that as an error)."""
    from bafqmc_amd.analysis import chain_seeds
    from bafqmc_amd.model import conf_set_batch
    d = dict(pairing=True, u1=1.0, u2=-0.5, mu=-3.0, rdelta=0.5, nlx=6, nly=6, ltrot=40, beta=4.0, nwrap=10,
             shift_loc=1.0)
    p = solver_params(d)
    phis, seeds = conf_set_batch(p, chain_seeds(97, 1))
    s = BatchedSolver([ModelSpec(p)], [0], lib=emu_lib, schedule=[(1, 16)])
    s.set_fields(list(phis), seeds)
    s.prepare()
    s.sweep(1, measure=False)
    assert s.read_bin()[0]["_stats"]["xmax"] > 1.0       # the chain has blown up
    s.close()


@pytest.mark.parametrize("schedule", [(1, 16), (1, 0), (0, 0)])
def test_emu_nonfinite_fields(emu_lib, schedule):
    """Non-finite fields give non-finite results, never an index out of range:
    a NaN column norm wins the QR pivot search, so every step has a pivot."""
    d = dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=3, nly=3, ltrot=8, beta=1.0, nwrap=4,
             shift_loc=1.0)
    phis = [np.full((8, 9, 2), np.nan), np.random.default_rng(1).normal(size=(8, 9, 2))]
    phis[1][3, 5, 0] = np.inf
    s = BatchedSolver([ModelSpec(solver_params(d))], [0, 0], lib=emu_lib, schedule=[schedule])
    s.set_fields(phis, [11, 12])
    s.prepare()
    s.sweep(1)
    b = s.read_bin()
    assert not np.isfinite(b[0]["density_total"]) and not np.isfinite(b[1]["density_total"])
    s.close()
