"""Command line: the drop-in for ``mpirun -np R bosonDQMC.out`` in a run directory.

    python -m bafqmc_amd run [--solver number_conserving|pairing] [--chains R] [DIR]
    python -m bafqmc_amd campaign --manifest M --output O [--solver S] [--np R] [--case ID ...]

DIR (default: the current directory) holds paramC_sets.txt, confin.txt and
seeds.txt; the per-bin observable files, info.txt and confout.txt are written
there, as the reference executable does (src/number_conserving/run_paper.py:113,162
launches it with cwd = the run directory).  R replaces the MPI rank count.
"""
from __future__ import annotations

import argparse
import os
import sys


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m bafqmc_amd")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run", help="run the sweep in a reference run directory")
    r.add_argument("directory", nargs="?", default=".")
    r.add_argument("--solver", default=os.environ.get("BAFQMC_SOLVER", "number_conserving"),
                   choices=["number_conserving", "pairing"])
    r.add_argument("--chains", type=int, default=int(os.environ.get("BAFQMC_CHAINS", "1")),
                   help="independent Markov chains (the reference's MPI ranks)")
    r.add_argument("--no-poles", action="store_true", help="skip the per-bin pole diagnostics")
    r.add_argument("--autotune", action="store_true",
                   help="time the execution schedules on a sample of the chains first and run the fastest "
                        "(the same Markov chain)")
    r.add_argument("--tune-nwrap", action="store_true",
                   help="with --autotune, also choose N_Ortho among the divisors of Ltrot near the input's, "
                        "admitting only intervals whose measured drift stays below 1e-7 (results change at "
                        "that level)")
    r.add_argument("--tau-conventions", choices=["bosonic", "reference"], default="bosonic",
                   help="time-displaced pass (is_tau): the bosonic block form, or the reference's code as written "
                        "(fermionic block sign, U2 propagated with the U1 field), whose info.txt drift lines it "
                        "reproduces")
    cp = sub.add_parser("campaign", help="run the cases of a reference manifest as batched launches")
    cp.add_argument("--manifest", required=True, help="the reference's manifest.json (cases: id, input_dir)")
    cp.add_argument("--output", required=True, help="run directories <output>/<case id>/dqmc are created here")
    cp.add_argument("--solver", default="number_conserving", choices=["number_conserving", "pairing"])
    cp.add_argument("--case", action="append", default=[], help="run only this case ID; repeatable")
    cp.add_argument("--np", type=int, default=1, help="chains per case (the reference's MPI ranks)")
    cp.add_argument("--autotune", action="store_true")
    cp.add_argument("--tune-nwrap", action="store_true")
    a = ap.parse_args(argv)
    if a.cmd == "campaign":
        from .campaign import run_campaign
        run_campaign(a.manifest, a.output, a.solver, a.np, only=a.case, autotune=a.autotune,
                     tune_nwrap=a.tune_nwrap)
        return 0
    from .driver import run
    from .solver import TAU_BOSONIC, TAU_REFERENCE
    run(a.directory, a.solver, a.chains, poles=False if a.no_poles else None, autotune=a.autotune,
        tune_nwrap=a.tune_nwrap, tau_mode=TAU_REFERENCE if a.tau_conventions == "reference" else TAU_BOSONIC)
    return 0


if __name__ == "__main__":
    sys.exit(main())
