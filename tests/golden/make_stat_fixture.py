"""Statistical fixture from the reference Fortran built from source.

Runs the reference number-conserving solver (oracle/_ref, built by
oracle/Makefile.ref) on the 3x3 benchmark chain -- triangular lattice, t = 1,
U1 = 1, U2 = 0, mu = -5, beta = 4, dtau = 0.1 (Ltrot = 40), Nwrap = 10,
shiftLoc = 1.5, fields i.i.d. N(0, 1) from seeds -- with R MPI ranks (one chain
each, one rank per host core: oversubscribed ranks spin in the per-bin
MPI_REDUCE) for Nbin one-sweep bins.  The seeds are states of the generator
(ranf, seed <- 48828125 seed mod 2^31, period 2^29 for odd seeds) 2^26 draws
apart, so the ranks' random streams (1440 draws per sweep) do not overlap.
The reference writes one rank-averaged row per bin (fourier_trans.f90:350-457);
the first 10 % of the rows are dropped as thermalisation and the rest blocked
into 10 blocks, the estimator of benchmarks/campaign_analysis.py:215-232
(restated as bafqmc_amd.analysis.block_statistics).  Writes
tests/golden/nc3x3_b4_stat_fortran.json (data only: parameters, seeds, mean and
standard error per observable).

    make -f oracle/Makefile.ref && python tests/golden/make_stat_fixture.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import run_ref  # noqa: E402
from bafqmc_amd.analysis import block_statistics  # noqa: E402

OUT = Path(__file__).resolve().parent / "nc3x3_b4_stat_fortran.json"
CHAIN = dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=3, nly=3, ltrot=40, beta=4.0, nwrap=10, shift_loc=1.5,
          ini_type=2, ini_ampl=1.0)
RANKS, NBIN = 8, 12000
SPACING = 1 << 26
FILES = {"density_total": "density_total", "energy_density": "energy_density", "sf_K": "sf_K", "dw_K": "dw_K"}


def main():
    p = dict(CHAIN, nbin=NBIN, nsweep=1)
    jump = pow(48828125, SPACING, 1 << 31)
    seeds = [86421 * pow(jump, k, 1 << 31) % (1 << 31) for k in range(RANKS)]
    d, wall = run_ref.run(p, seeds, nranks=RANKS, timeout=3600)
    out = {"params": p, "ranks": RANKS, "seeds": seeds, "wall_s": wall, "skip_rows": NBIN // 10,
           "blocks": 10, "observables": {}}
    for key, fname in FILES.items():
        rows = run_ref.read_rows(d / fname)
        vals = [float(r[0]) for r in rows]            # real part (sf_K, dw_K: real imag columns)
        st = block_statistics(vals, (NBIN - NBIN // 10) // 10, skip=NBIN // 10)
        out["observables"][key] = {"mean": st["actual"], "stderr": st["stderr"], "rows": len(vals)}
    run_ref.cleanup(d)
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out["observables"], indent=1), f"{wall:.1f} s")


if __name__ == "__main__":
    main()
