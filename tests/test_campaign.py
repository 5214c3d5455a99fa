"""Campaign glue (bafqmc_amd.campaign): the cases of a reference-format
manifest run as batched launches write, case by case, the files a run of that
case alone writes -- every observable row, the pole diagnostics, confout.txt
and info.txt (all but its wall-time row) identical -- and cases with another
warm-up length, Nbin or Nsweep go into a batch of their own."""
from __future__ import annotations

import io
import json

import pytest

from bafqmc_amd import campaign, driver
from bafqmc_amd.model import Params
from oracle import run_ref

CASES = {
    "nc_a": dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=3, nly=3, ltrot=12, beta=1.2, nwrap=6, nbin=2,
                 shift_loc=1.5, ini_type=2, ini_ampl=1.0),
    "nc_b": dict(pairing=False, u1=0.5, u2=-0.2, mu=-6.0, nlx=3, nly=3, ltrot=8, beta=0.8, nwrap=4, nbin=2,
                 shift_loc=1.5, ini_type=2, ini_ampl=1.0),
    "nc_warm": dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=3, nly=3, ltrot=10, beta=1.0, nwrap=5, nbin=3,
                    shift_loc=1.5, ini_type=2, ini_ampl=1.0, nwarm=2),
}


def make_manifest(root):
    cases = []
    for k, (cid, p) in enumerate(CASES.items()):
        d = root / "inputs" / cid
        run_ref.write_inputs(d, p, [1234567 + 1000 * k, 7654321 + 1000 * k])
        cases.append({"id": cid, "input_dir": f"inputs/{cid}"})
    (root / "manifest.json").write_text(json.dumps({"cases": cases}))
    return root / "manifest.json"


def check_campaign_equals_single_runs(tmp_path, lib):
    manifest = make_manifest(tmp_path)
    log = io.StringIO()
    campaign.run_campaign(manifest, tmp_path / "batched", "number_conserving", 2, lib=lib, log=log)
    assert log.getvalue().count("batch of") == 2            # nc_a + nc_b together, nc_warm alone
    for cid in CASES:
        alone = tmp_path / "alone" / cid
        alone.mkdir(parents=True)
        for name in campaign.INPUTS:
            (alone / name).write_bytes((tmp_path / "inputs" / cid / name).read_bytes())
        driver.run(alone, "number_conserving", 2, lib=lib, log=io.StringIO())
        batched = tmp_path / "batched" / cid / "dqmc"
        names = sorted(f.name for f in alone.iterdir())
        assert names == sorted(f.name for f in batched.iterdir())
        for name in names:
            a, b = (alone / name).read_text(), (batched / name).read_text()
            if name == "info.txt":
                a = "\n".join(l for l in a.splitlines() if "Tot CPU time" not in l)
                b = "\n".join(l for l in b.splitlines() if "Tot CPU time" not in l)
            assert a == b, (cid, name)


def test_campaign_equals_single_runs_emulator(tmp_path, emu_lib):
    check_campaign_equals_single_runs(tmp_path, emu_lib)


@pytest.mark.gpu
def test_campaign_equals_single_runs_gpu(tmp_path, hip_lib):
    check_campaign_equals_single_runs(tmp_path, hip_lib)


def test_campaign_refuses_nonempty_run_directories(tmp_path, emu_lib):
    manifest = make_manifest(tmp_path)
    (tmp_path / "out" / "nc_a" / "dqmc").mkdir(parents=True)
    (tmp_path / "out" / "nc_a" / "dqmc" / "density_total").write_text("1\n")
    with pytest.raises(ValueError, match="nonempty run directory"):
        campaign.run_campaign(manifest, tmp_path / "out", "number_conserving", 2, lib=emu_lib, log=io.StringIO())
    with pytest.raises(ValueError, match="unknown case IDs"):
        campaign.read_manifest(manifest, only=["missing"])
