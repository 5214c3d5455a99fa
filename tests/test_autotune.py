"""Schedules and the autotuner (bafqmc_amd.autotune).

Every schedule must compute the same Markov chain: a chain run under each
schedule the library accepts ends with the same fields, Green's function and
observables as under the default one (to rounding: the delayed-update width
changes the order of floating-point operations, and the two paths factor
with different pivoted QRs).  G agrees to 1e-10 of its largest entry; an
estimator is a sum over N^2 products of G's entries whose terms can cancel
(dw_K of the 6x6 pairing case differed by 1.07e-10 of its value between the
two paths on gfx950), so the estimators are held to 3e-10 of their value.  The tuner's choice must be the fastest admitted entry
of its own timing table.
"""
from __future__ import annotations

import numpy as np
import pytest

from bafqmc_amd import autotune
from bafqmc_amd.model import ModelSpec
from bafqmc_amd.solver import BatchedSolver, SolverError
from tests.helpers import fixed_fields, solver_params
from tests.test_gpu_parity import CASES


def _run(lib, d, schedule):
    ns = d["nlx"] * d["nly"]
    s = BatchedSolver([ModelSpec(solver_params(d))], [0, 0], lib=lib, schedule=schedule)
    s.set_fields([fixed_fields(d["ltrot"], ns), fixed_fields(d["ltrot"], ns)[::-1].copy()], [2468, 1357])
    s.prepare()
    s.sweep(1)
    out = s.read_bin(), [s.green(c) for c in range(2)], s.get_fields()[0], s.schedule()[0]
    s.close()
    return out


def check_schedules_agree(lib, case):
    d = CASES[case]
    ref = _run(lib, d, None)
    ran = 0
    for sched in autotune.SCHEDULES:
        try:
            bins, gs, fields, eff = _run(lib, d, [sched])
        except SolverError:
            continue
        ran += 1
        assert tuple(eff) == sched
        for c in range(2):
            assert np.array_equal(fields[c], ref[2][c])          # same accept/reject decisions
            assert np.max(np.abs(gs[c] - ref[1][c])) < 1e-10 * np.max(np.abs(ref[1][c]))
            for k in ("density_total", "energy_density", "sf_K", "dw_K"):
                assert abs(bins[c][k] - ref[0][c][k]) <= 3e-10 * max(1e-300, abs(ref[0][c][k])), (sched, k)
    assert ran >= 6


@pytest.mark.parametrize("case", [3])
def test_emu_schedules_agree(emu_lib, case):
    check_schedules_agree(emu_lib, case)


def test_emu_invalid_schedules_rejected(emu_lib):
    d = CASES[3]
    for bad in ((2, 0), (1, 6), (0, 6), (1, 20), (-2, 0)):
        with pytest.raises(SolverError):
            _run(emu_lib, d, [bad])


def test_emu_tuner_runs(emu_lib, tmp_path):
    d = dict(CASES[3])
    p = solver_params(d)
    cache = tmp_path / "tune.json"
    ch = autotune.tune(emu_lib, [p], [0, 0], sweeps=1, tune_nwrap=True, cache=cache)[0]
    assert ch.schedule in autotune.SCHEDULES
    assert p.ltrot % ch.nwrap == 0
    again = autotune.tune(emu_lib, [p], [0, 0], sweeps=1, tune_nwrap=True, cache=cache)[0]
    assert again.schedule == ch.schedule and again.nwrap == ch.nwrap


@pytest.mark.gpu
@pytest.mark.parametrize("case", [3, 5])
def test_schedules_agree_gpu(hip_lib, case):
    check_schedules_agree(hip_lib, case)


@pytest.mark.gpu
def test_tuner_picks_the_fastest_gpu(hip_lib):
    d = dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=6, nly=6, ltrot=40, beta=4.0, nwrap=10,
             shift_loc=1.0)
    ch = autotune.tune_model(hip_lib, solver_params(d), chains=64, sweeps=2, tune_nwrap=True)
    print("autotune", ch, flush=True)
    times = [v for v in ch.table.values() if v is not None]
    assert ch.ms_per_sweep == min(times)


def test_cache_key_separates_hamiltonians():
    """Two models that differ only in U2 (or beta, mu, Delta) get their own
    tuning entries: a schedule or Nwrap validated for one is never reused for
    the other."""
    import dataclasses
    from bafqmc_amd.autotune import _key
    from tests.helpers import solver_params
    base = solver_params(dict(pairing=True, u1=1.0, u2=0.0, mu=-5.0, rdelta=0.5, nlx=3, nly=3, ltrot=40,
                              beta=4.0, nwrap=10, shift_loc=1.0))
    keys = {_key(base, 16, "gfx950")}
    for change in (dict(u2=-0.5), dict(beta=4.5), dict(mu=-4.0), dict(rdelta=0.3), dict(u1=0.5)):
        keys.add(_key(dataclasses.replace(base, **change), 16, "gfx950"))
    assert len(keys) == 6
