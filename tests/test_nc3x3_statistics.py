"""The plain number-conserving chain on the 3x3 cluster against exact
diagonalisation and against the reference Fortran's own statistics.

Workload: triangular 3x3 lattice, t = 1, U1 = 1, U2 = 0, mu = -5, beta = 4,
dtau = 0.1 (Ltrot = 40), Nwrap = 10, shiftLoc = 1.5, fields i.i.d. N(0, 1) from
per-chain seeds, 2000 measured sweeps per chain, 256 chains (every chain
counted).  The standard error is the spread of the chains' time averages.

* ED: oracle/ed.py in the 4-boson space, exact for this dilute gas to 6e-11
  (tests/test_ed_oracle.py).  The Trotter error at dtau = 0.1 is below the
  statistical error here: chains at dtau = 0.1, 0.05, 0.025 and 0.0125 all
  land within 1.1 standard errors of ED.  Required: |mean - ED| <= 2 sigma.
* The reference: tests/golden/nc3x3_b4_stat_fortran.json, the same chain run
  by the reference Fortran built from source (8 MPI ranks x 12000 one-sweep
  bins, blocked standard error; tests/golden/make_stat_fixture.py).
  Required: |mean - Fortran| <= 2 sqrt(sigma^2 + sigma_F^2).
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from bafqmc_amd.model import Params
from oracle.ed import FockED
from tests.helpers import mean_and_error, run_statistics

pytestmark = pytest.mark.gpu

FORTRAN = json.loads((Path(__file__).resolve().parent / "golden" / "nc3x3_b4_stat_fortran.json").read_text())
CHAINS, THERM, MEAS = 256, 200, 2000


def params():
    p = Params(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=3, nly=3, ltrot=40, beta=4.0, nwrap=10,
               shift_loc=1.5, ini_type=2, ini_ampl=1.0)
    p.nlx_therm, p.nly_therm, p.ltrot_therm = 3, 3, 40
    return p.validate()


def test_nc3x3_chain_matches_ed_and_reference_statistics(hip_lib):
    print(f"nc 3x3 beta=4 dtau=0.1: {CHAINS} chains, {THERM} + {MEAS} sweeps", flush=True)
    means, diag = run_statistics(hip_lib, params(), CHAINS, THERM, MEAS, 4711)
    assert diag["finite"].all() and float(diag["drift"].max()) < 5.5e-5
    ed = FockED(3, 3, nmax=3, ncut=4).thermal(1.0, 0.0, -5.0, 0.0, 4.0)
    failures = []
    for k in ("density_total", "energy_density", "sf_K", "dw_K"):
        m, se = mean_and_error(means[k])
        f = FORTRAN["observables"][k]
        z_ed = (m - ed[k]) / se
        z_f = (m - f["mean"]) / math.hypot(se, f["stderr"])
        line = (f"{k}: {m:.6g} +- {se:.2g}  ED {ed[k]:.6g} (z {z_ed:+.2f})  "
                f"Fortran {f['mean']:.6g} +- {f['stderr']:.2g} (z {z_f:+.2f})")
        print(line, flush=True)
        if abs(z_ed) > 2.0 or abs(z_f) > 2.0:
            failures.append(line)
    assert not failures, failures
