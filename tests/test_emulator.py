"""The HIP kernels on the CPU emulator back end (one OS thread per GPU thread).

Same device code as the gfx950 build; small cases only.  Checks logic, indexing
and barriers without a GPU (the MFMA lane layout itself is checked on the GPU).
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import rel_err
from tests.test_gpu_parity import check_sweep_vs_oracle, check_golden, check_mfma_layout, check_lane_ops


def test_emu_mfma_layout(emu_lib):
    check_mfma_layout(emu_lib)


def test_emu_gemm_qrp(emu_lib):
    rng = np.random.default_rng(3)
    for n in (5, 18, 33):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        b = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        for oa in (0, 1):
            for ob in (0, 1):
                A = a.conj().T if oa else a
                B = b.conj().T if ob else b
                assert rel_err(emu_lib.debug_gemm(a, b, oa, ob), A @ B) < 1e-13
        r, tau, perm, d = emu_lib.debug_qrp(a)
        R = np.triu(r) * d[:, None]
        Q = np.eye(n, dtype=complex)
        for k in range(n - 1, -1, -1):
            v = np.zeros(n, complex)
            v[k] = 1.0
            v[k + 1:] = r[k + 1:, k]
            Q = (np.eye(n) - tau[k] * np.outer(v, v.conj())) @ Q
        assert rel_err(Q @ R, a[:, perm]) < 1e-13


@pytest.mark.parametrize("case", [1, 3])
def test_emu_lds_path_matches_oracle(emu_lib, case):
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("case", [1, 3])
def test_emu_global_path_matches_oracle(emu_lib, case, monkeypatch):
    monkeypatch.setenv("BAFQ_FORCE_GLOBAL", "1")
    check_sweep_vs_oracle(emu_lib, case)


@pytest.mark.parametrize("name", ["pairing_fixed_3x3_w2", "pairing_warm_3x3"])
def test_emu_matches_reference_fortran(emu_lib, name):
    check_golden(emu_lib, name)


@pytest.mark.parametrize("width", ["0", "4"])
def test_emu_local_update_variants(emu_lib, width, monkeypatch):
    monkeypatch.setenv("BAFQ_DELAYED", width)
    check_sweep_vs_oracle(emu_lib, 3)


def test_emu_lane_ops(emu_lib):
    check_lane_ops(emu_lib)
