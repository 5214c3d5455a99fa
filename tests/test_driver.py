"""The drop-in run-directory driver: reference input grammar in, reference output layout out.

CPU tests run the kernels on the emulator back end on the smallest golden cases;
the GPU test runs a two-chain case (seeds.txt field generation, chain averaging).
"""
from __future__ import annotations

import io

import json

import numpy as np
import pytest

from bafqmc_amd import driver
from bafqmc_amd.model import Params, read_params
from oracle import run_ref
from tests.helpers import load_golden, golden_row, golden_info_value, assert_obs_close, obs_atol


def _write_case(tmp_path, name):
    z, pd = load_golden(name)
    p = dict(pd)
    d = tmp_path / name
    fields = list(z["fields_in"]) if "fields_in" in z.files else None
    run_ref.write_inputs(d, p, [int(s) for s in z["seeds"]], fields)
    return z, pd, d


def _check_outputs(z, pd, d):
    for key in z.files:
        if not key.startswith("out_"):
            continue
        name = key[4:]
        rows = run_ref.read_rows(d / name)
        ref = z[key]
        if name == "pole_z":
            # eigenvalue order is not defined; compare the sorted moduli
            got = rows.reshape(rows.shape[0], -1, 2)
            want = ref.reshape(ref.shape[0], -1, 2)
            gm = np.sort(np.hypot(got[..., 0], got[..., 1]), axis=1)
            wm = np.sort(np.hypot(want[..., 0], want[..., 1]), axis=1)
            np.testing.assert_allclose(gm, wm, rtol=1e-9, err_msg=name)
            continue
        assert rows.shape == ref.shape, (name, rows.shape, ref.shape)
        for b in range(rows.shape[0]):
            want = ref[b]
            got = rows[b]
            if want.size == 2 and name not in ("density_site_total",):
                assert_obs_close(complex(got[0], got[1]), complex(want[0], want[1]), f"{name} row {b}",
                                 atol=obs_atol(name, pd["pairing"]))
            else:
                assert_obs_close(got, want, f"{name} row {b}", atol=max(1e-11, obs_atol(name, pd["pairing"])))
    toks = (d / "confout.txt").read_text().split()
    assert int(toks[0]) == int(z["seeds_out"][0])
    nvals = z["fields_out"][0].size
    np.testing.assert_allclose(np.array([float(t) for t in toks[1:1 + nvals]]), z["fields_out"][0].ravel(),
                               rtol=0, atol=1e-12)
    info = (d / "info.txt").read_text()
    assert "Tot CPU time" in info and "Max diff Matrix" in info
    labels = ["Accept_U1_shift"] + (["Thermalize U1 Accept Ratio"] if pd.get("nwarm", 0) else [])
    if abs(pd["u2"]) > 1e-10:
        labels.append("Accept_U2_shift")
        if pd.get("nwarm", 0):
            labels.append("Thermalize U2 Accept Ratio")
    for label in labels:
        mine = golden_info_value({"info": info}, label)
        assert abs(mine - golden_info_value(z, label)) < 1e-12, label
    # the same rows in the same order as the reference's info.txt
    # (write_info, Therm/Local_control_print, main.f90:59-76)
    def rows(text):
        return [line.split(":")[0].strip() for line in text.splitlines() if ":" in line]
    assert rows(info) == rows(str(z["info"]))
    # compare_mat drift (stabilization.f90:256-266, main.f90:59-73): rounding of
    # two codes, so an order of magnitude; below 1e-8 both are rounding noise
    for label in ("Max diff Matrix", "Mean diff Matrix"):
        mine, ref = golden_info_value({"info": info}, label), golden_info_value(z, label)
        if ref >= 1e-8:
            assert ref / 10 <= mine <= 10 * ref, (label, mine, ref)
        else:
            assert mine < 1e-8, (label, mine, ref)
    for label in [r for r in rows(str(z["info"])) if r.startswith("Phase_")]:
        # the reference's phases are its rounding (~1e-12); zero in the reduced form
        tol = 1e-9 if "mean_abs_arg" in label else 1e-12
        assert abs(golden_info_value({"info": info}, label) - golden_info_value(z, label)) < tol, label


def test_read_params_roundtrip(tmp_path):
    for pairing in (False, True):
        p = Params(pairing=pairing, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5 if pairing else 0.0, nlx=6, nly=6,
                   ltrot=80, beta=8.0, nlx_therm=6, nly_therm=6, ltrot_therm=80, nwrap=10, nbin=3, nsweep=2,
                   shift_loc=1.0, is_warm=True, nwarm=5)
        f = tmp_path / f"p{int(pairing)}.txt"
        f.write_text(p.to_text())
        q = read_params(f, pairing)
        assert q == p


def test_read_params_validation(tmp_path):
    p = Params(u1=-1.0)
    f = tmp_path / "bad.txt"
    f.write_text(p.to_text())
    with pytest.raises(ValueError, match="U1 must be non-negative"):
        read_params(f, False)
    p = Params(ltrot=40, nwrap=7, nlx_therm=3, nly_therm=3, ltrot_therm=40)
    f.write_text(p.to_text())
    with pytest.raises(ValueError, match="multiple of Nwrap"):
        read_params(f, False)


@pytest.mark.parametrize("name,solver", [("nc_fixed_3x3_w2", "number_conserving")])
def test_driver_matches_reference_outputs_emulator(tmp_path, emu_lib, name, solver):
    z, pd, d = _write_case(tmp_path, name)
    driver.run(d, solver, len(z["seeds"]), lib=emu_lib)
    _check_outputs(z, pd, d)


def test_driver_autotuned_schedule_gives_the_same_outputs_emulator(tmp_path, emu_lib):
    """--autotune picks the fastest schedule; every schedule is the same Markov chain."""
    z, pd, d = _write_case(tmp_path, "nc_traj_3x3_u2")
    log = io.StringIO()
    driver.run(d, "number_conserving", len(z["seeds"]), lib=emu_lib, autotune=True, log=log)
    assert "schedule" in log.getvalue()
    _check_outputs(z, pd, d)


@pytest.mark.gpu
@pytest.mark.parametrize("name,solver", [("nc_traj_3x3_c1", "number_conserving"),
                                         ("pairing_traj_3x3", "pairing"),
                                         ("nc_warm_3x3", "number_conserving"),
                                         ("pairing_warm_3x3", "pairing"),
                                         ("pairing_traj_3x3_d03", "pairing")])
def test_driver_matches_reference_outputs_gpu(tmp_path, hip_lib, name, solver):
    z, pd, d = _write_case(tmp_path, name)
    driver.run(d, solver, len(z["seeds"]), lib=hip_lib)
    _check_outputs(z, pd, d)
