"""Collect the reference's published 3x3 benchmark numbers as a small fixture.

Reads (as data) the paper table benchmarks/paper/data/observables.csv (the
reference's own DQMC mean and standard error and the exact-diagonalisation
value per observable) and each case's DQMC input paramC_sets.txt, and writes
tests/golden/paper_3x3.json.

    python tests/golden/make_paper_fixture.py [/path/to/reference]
"""
from __future__ import annotations

import csv
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
DATA = REF / "benchmarks" / "paper" / "data"
OUT = Path(__file__).resolve().parent / "paper_3x3.json"

CASES = [
    ("number_conserving", "triangle_3x3_U1_1_mu-3.5_b4"),
    ("number_conserving", "triangle_3x3_U2_-0.5_mu-7_b1"),
    ("pairing", "triangle_pairing_delta_3x3_D0p3_mum5_b4"),
    ("pairing", "triangle_pairing_delta_3x3_D0_mum5_b4"),
]
OBS = ("density_total", "energy_density", "S_SF_K", "S_DW_K")


def params(model: str, case: str) -> dict:
    rows = (DATA / model / "inputs" / case / "paramC_sets.txt").read_text().split("\n")
    r = [line.split() for line in rows[:7]]
    p = {"pairing": model == "pairing", "u1": float(r[0][0]), "u2": float(r[0][1]), "mu": float(r[0][2]),
         "nlx": int(r[1][0]), "nly": int(r[1][1]), "ltrot": int(r[1][2]), "beta": float(r[1][3]),
         "nwrap": int(r[3][0]), "shift_loc": float(r[3][3]), "nwarm": int(r[5][1]) if r[5][0] == ".true." else 0,
         "ini_type": int(r[6][0]), "ini_ampl": float(r[6][1])}
    if model == "pairing":
        p["rdelta"] = float(r[0][3])
    return p


def main():
    table = list(csv.DictReader(open(DATA / "observables.csv")))
    out = {}
    for model, case in CASES:
        vals = {}
        for row in table:
            if row["model"] == model and row["case"] == case and row["observable"] in OBS:
                vals[row["observable"]] = {"dqmc": float(row["dqmc"]), "stderr": float(row["stderr"]),
                                           "ed": float(row["ed"])}
        out[case] = {"params": params(model, case), "observables": vals}
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
