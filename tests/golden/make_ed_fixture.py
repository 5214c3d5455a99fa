"""Collect the reference's stored 3x3 exact-diagonalisation results as a fixture.

Reads (as data) every JSON under benchmarks/paper/data/{pairing,number_conserving}/ed/
of the reference and writes tests/golden/ed_3x3.json: per case the Hamiltonian
parameters, the cutoff of the finite Fock space (pairing: nmax, ncut;
number conserving: the particle-number shells summed and the last shell's
relative contribution) and the observables the sweep estimates.

    python tests/golden/make_ed_fixture.py [/path/to/reference]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
DATA = REF / "benchmarks" / "paper" / "data"
OUT = Path(__file__).resolve().parent / "ed_3x3.json"
OBS = ("density_total", "energy_density", "S_SF_K", "S_DW_K", "interaction_energy_density", "pair_equal")


def value(v):
    return float(v["value"]) if isinstance(v, dict) else float(v)


def main():
    out = {}
    for model in ("pairing", "number_conserving"):
        for f in sorted((DATA / model / "ed").glob("*.json")):
            d = json.loads(f.read_text())
            par = d["parameters"]
            rec = {"model": model,
                   "params": {"u1": float(par["U1"]), "u2": float(par["U2"]), "mu": float(par["mu"]),
                              "beta": float(par["beta"]), "delta": float(par.get("Delta", 0.0)),
                              "lx": int(par["Lx"]), "ly": int(par["Ly"]), "t": float(par["t"])},
                   "observables": {k: value(d["observables"][k]) for k in OBS if k in d["observables"]}}
            if model == "pairing":
                rec["cutoff"] = {"nmax": int(d["cutoffs"]["nmax"]), "ncut": int(d["cutoffs"]["ncut"]),
                                 "basis_size": int(d["basis_size"])}
            else:
                rec["cutoff"] = {"max_total_particles": int(par["max_total_particles"]),
                                 "completed_shells": int(d["completed_shells"]), "status": d["status"],
                                 "relative_last_shell": {k: float(v) for k, v in d["relative_last_shell"].items()
                                                         if k in OBS}}
            out[f.stem] = rec
    OUT.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print(f"{len(out)} cases -> {OUT}")


if __name__ == "__main__":
    main()
