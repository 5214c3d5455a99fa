"""Statistical agreement with exact diagonalisation and with the reference's own
published DQMC numbers on 3x3 clusters (benchmarks/paper/data/observables.csv,
collected in tests/golden/paper_3x3.json by tests/golden/make_paper_fixture.py).

Each case runs the reference's inputs (same Hamiltonian, Ltrot, Nwrap,
shiftLoc, initial field amplitude and Gaussian warm-up of Nwarm sweeps) as 256
independent chains on the GPU (one per CU), THERM local sweeps unmeasured, then
MEAS measured sweeps.  Every chain is counted; none may cross the reference's
"ortho unstable" drift threshold (stabilization.f90:320,358) or turn
non-finite.  The standard error comes from the spread of the chains' time
averages.  Required: |mean - reference DQMC| <= 2 sqrt(sigma^2 + sigma_ref^2)
for every observable, and |mean - ED| <= 2 sigma wherever the reference's own
published DQMC agrees with ED within 2 of its standard errors.

Where it does not (the pairing case at Delta = 0.3: the published DQMC is 7 of
its standard errors from its ED), the ED distance is reported, not asserted:
that ED is a finite space of at most 4 bosons, and at Delta = 0.3 the pair
field populates its edge -- on the free gas (U1 = U2 = 0) the same space misses
the closed form by 0.2 % (tests/test_ed_oracle.py), while the DQMC-ED gap here
is 0.09 %.

Chain means over 100 sweeps at Delta = 0 (density 3e-6 per site) spread from
0.3x to 2x the ensemble mean; with so few sweeps the spread underestimates the
error of the mean and runs landed 1.9 and 2.6 standard errors above ED.  Over
1000 sweeps per chain the means are close to normal.  The run is seeded, so
the outcome is reproducible.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from bafqmc_amd.model import Params
from tests.helpers import mean_and_error, run_statistics

pytestmark = pytest.mark.gpu

FIXTURE = json.loads((Path(__file__).resolve().parent / "golden" / "paper_3x3.json").read_text())
NAMES = {"density_total": "density_total", "energy_density": "energy_density", "S_SF_K": "sf_K",
         "S_DW_K": "dw_K"}
CHAINS, MEAS = 256, 1000
THERM = {False: 50, True: 1000}      # local sweeps before measuring, by pairing
UNSTABLE_DRIFT = 5.5e-5


def case_params(d):
    p = Params(pairing=d["pairing"], u1=d["u1"], u2=d["u2"], mu=d["mu"], rdelta=d.get("rdelta", 0.0),
               nlx=d["nlx"], nly=d["nly"], ltrot=d["ltrot"], beta=d["beta"], nwrap=d["nwrap"],
               shift_loc=d["shift_loc"], ini_type=d["ini_type"], ini_ampl=d["ini_ampl"],
               is_warm=d["nwarm"] > 0, nwarm=d["nwarm"])
    p.nlx_therm, p.nly_therm, p.ltrot_therm = p.nlx, p.nly, p.ltrot
    return p.validate()


@pytest.mark.parametrize("case", list(FIXTURE))
def test_agrees_with_ed_and_reference_dqmc(hip_lib, case):
    d = FIXTURE[case]["params"]
    print(f"paper case {case}: {CHAINS} chains, {THERM[d['pairing']]} + {MEAS} sweeps", flush=True)
    means, diag = run_statistics(hip_lib, case_params(d), CHAINS, THERM[d["pairing"]], MEAS,
                                 20260101 + len(case))
    unstable = int(np.sum(diag["drift"] >= UNSTABLE_DRIFT))
    print(f"{case}: largest drift {diag['drift'].max():.2e}, unstable chains {unstable}", flush=True)
    assert unstable == 0 and diag["finite"].all()
    lines, failures = [], []
    for k, ref in FIXTURE[case]["observables"].items():
        m, se = mean_and_error(means[NAMES[k]])
        z_ed = (m - ref["ed"]) / se
        z_ref = (m - ref["dqmc"]) / math.hypot(se, ref["stderr"])
        line = (f"{k}: {m:.6g} +- {se:.2g}  ED {ref['ed']:.6g} (z {z_ed:+.2f})  "
                f"ref {ref['dqmc']:.6g} +- {ref['stderr']:.2g} (z {z_ref:+.2f})")
        print(case, line, flush=True)
        if abs(z_ref) > 2.0:
            failures.append(line)
        if abs(ref["dqmc"] - ref["ed"]) <= 2.0 * ref["stderr"] and abs(z_ed) > 2.0:
            failures.append(line)
    assert not failures, failures
