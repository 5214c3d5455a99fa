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
                                   check_uneven_batch, check_gaussian, GOLDEN)


def test_emu_mfma_layout(emu_lib):
    check_mfma_layout(emu_lib)


def test_emu_gemm(emu_lib):
    rng = np.random.default_rng(3)
    for n in (5, 18, 33):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        b = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        for oa in (0, 1):
            for ob in (0, 1):
                A = a.conj().T if oa else a
                B = b.conj().T if ob else b
                assert rel_err(emu_lib.debug_gemm(a, b, oa, ob), A @ B) < 1e-13


@pytest.mark.parametrize("n,variant", [(9, 0), (18, 0), (33, 0), (9, 1), (18, 1), (32, 1)])
def test_emu_qrp(emu_lib, n, variant):
    check_qrp(emu_lib, n, variant)


@pytest.mark.parametrize("case", range(7))
def test_emu_sweep_matches_oracle(emu_lib, case):
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("case", [1, 3, 5])
def test_emu_global_path_matches_oracle(emu_lib, case, monkeypatch):
    monkeypatch.setenv("BAFQ_FORCE_GLOBAL", "1")
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("name", GOLDEN)
def test_emu_matches_reference_fortran(emu_lib, name):
    check_golden(emu_lib, name)


@pytest.mark.parametrize("width", ["0", "4", "8"])
@pytest.mark.parametrize("case", [3, 5])
def test_emu_local_update_variants(emu_lib, case, width, monkeypatch):
    monkeypatch.setenv("BAFQ_DELAYED", width)
    check_sweep_vs_oracle(emu_lib, case)


def test_emu_uneven_batch(emu_lib):
    check_uneven_batch(emu_lib)


@pytest.mark.parametrize("pairing,delta", [(False, 0.0), (True, 0.2)])
def test_emu_gaussian_closed_form(emu_lib, pairing, delta):
    check_gaussian(emu_lib, pairing, delta)


def test_emu_lane_ops(emu_lib):
    check_lane_ops(emu_lib)


@pytest.mark.parametrize("case,force_global", [(3, "0"), (5, "0"), (4, "1")])
def test_emu_thread_order_independence(emu_lib, case, force_global, monkeypatch):
    """Race check: the same sweep with the threads of every workgroup run in
    ascending and in descending order between barriers gives identical bits."""
    monkeypatch.setenv("BAFQ_FORCE_GLOBAL", force_global)
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
