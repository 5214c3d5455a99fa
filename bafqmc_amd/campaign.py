"""Campaign glue: many reference run directories in one batched launch.

The reference runs a campaign case by case: for every case of a manifest,
``run_paper.py --mode dqmc`` copies the case's inputs into a fresh run
directory and starts ``mpirun -np R bosonDQMC.out`` there
(src/number_conserving/run_paper.py:75-175; the same for src/pairing; the
paper's production loop over both manifests is
benchmarks/paper/production.py:35-125).  Here the cases of a manifest are
prepared the same way -- ``<output>/<case id>/dqmc`` holding copies of
``paramC_sets.txt``, ``confin.txt`` and ``seeds.txt`` -- and then run as a few
batched solvers instead of one process each: cases that share the warm-up
length, Nbin and Nsweep go into one handle (one model per case, R chains per
case), whatever their lattices, Ltrot or Hamiltonians.  Every case's output
files are those of ``python -m bafqmc_amd run`` on that directory alone
(bafqmc_amd.driver.run_cases), and the row counts are checked as run_paper.py
does (:49-65).
"""
from __future__ import annotations

import json
import re
import shutil
import sys
from pathlib import Path

from .driver import load_case, run_cases

INPUTS = ("paramC_sets.txt", "confin.txt", "seeds.txt")


def read_manifest(manifest: Path, only=()):
    """Cases of a reference manifest ({"cases": [{"id", "input_dir", ...}]},
    input_dir relative to the manifest file), optionally restricted to `only`."""
    manifest = Path(manifest).resolve()
    data = json.loads(manifest.read_text())
    cases = data.get("cases")
    if not isinstance(cases, list) or not cases:
        raise ValueError("manifest must define a nonempty cases list")
    ids = [c["id"] for c in cases]
    if len(ids) != len(set(ids)) or any(not re.fullmatch(r"[A-Za-z0-9_.-]+", i) or i in {".", ".."} for i in ids):
        raise ValueError("case IDs must be unique simple directory names")
    unknown = set(only) - set(ids)
    if unknown:
        raise ValueError(f"unknown case IDs: {sorted(unknown)}")
    out = []
    for c in cases:
        if only and c["id"] not in only:
            continue
        inputs = Path(c["input_dir"])
        out.append((c["id"], inputs if inputs.is_absolute() else manifest.parent / inputs))
    return out


def prepare(manifest: Path, output: Path, only=()):
    """Fresh run directory per case with copies of its inputs; a nonempty
    directory is refused, as the reference's drivers refuse it."""
    dirs = []
    for cid, inputs in read_manifest(manifest, only):
        for name in INPUTS:
            if not (inputs / name).is_file():
                raise FileNotFoundError(inputs / name)
        target = Path(output).resolve() / cid / "dqmc"
        if target.exists() and any(target.iterdir()):
            raise ValueError(f"nonempty run directory: {target}; choose a fresh --output")
        target.mkdir(parents=True, exist_ok=True)
        for name in INPUTS:
            shutil.copy2(inputs / name, target / name)
        dirs.append((cid, target))
    return dirs


def run_campaign(manifest: Path, output: Path, solver: str = "number_conserving", nchains: int = 1, only=(),
                 lib=None, log=sys.stdout, autotune: bool = False, tune_nwrap: bool = False):
    """Prepare and run every case of the manifest; returns {case id: run statistics}."""
    dirs = prepare(manifest, output, only)
    cases = {cid: load_case(d, solver, nchains) for cid, d in dirs}
    groups = {}
    for cid, c in cases.items():
        p = c.params
        groups.setdefault((p.nwarm if (p.is_warm and p.nwarm > 0) else 0, p.nbin, p.nsweep), []).append(cid)
    stats = {}
    for key, ids in sorted(groups.items()):
        print(f"bafqmc_amd: batch of {len(ids)} case(s) (Nwarm {key[0]}, Nbin {key[1]}, Nsweep {key[2]}): "
              + ", ".join(ids), file=log, flush=True)
        for cid, st in zip(ids, run_cases([cases[i] for i in ids], lib=lib, log=log, autotune=autotune,
                                          tune_nwrap=tune_nwrap)):
            stats[cid] = st
    for cid, c in cases.items():
        rows = len((c.run_dir / "density_total").read_text().splitlines())
        if rows != c.params.nbin:
            raise ValueError(f"{c.run_dir / 'density_total'}: expected {c.params.nbin} bins, found {rows}")
    return stats
