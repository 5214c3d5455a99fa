"""Run the reference Fortran solvers built by ``oracle/Makefile.ref``.

TEST INFRASTRUCTURE ONLY (golden-vector generation and the timed CPU
baseline).  Writes the reference's own input files (``paramC_sets.txt``,
``confin.txt``, ``seeds.txt``; grammar of src/number_conserving/src/calc_basic.f90:48-91
and src/number_conserving/src/fields.f90:47-113) into a fresh directory, runs
``mpirun -np R <exe>`` there like src/number_conserving/run_paper.py:113,162,
and parses the per-bin output files.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_DIR = HERE / "_ref"
MPIRUN = os.environ.get("BAFQMC_MPIRUN", "/opt/conda/bin/mpirun")


def executable(pairing: bool) -> Path:
    return REF_DIR / ("bosonDQMC_pairing.out" if pairing else "bosonDQMC_nc.out")


def available(pairing: bool = False) -> bool:
    return executable(pairing).exists() and Path(MPIRUN).exists()


def param_text(p: dict) -> str:
    """Seven list-directed rows (calc_basic.f90:51-59)."""
    first = f"{p['u1']!r} {p['u2']!r} {p['mu']!r}"
    if p.get("pairing"):
        first += f" {p.get('rdelta', 0.0)!r}"
    rows = [
        first,
        f"{p['nlx']} {p['nly']} {p['ltrot']} {p['beta']!r}",
        f"{p.get('nlx_therm', p['nlx'])} {p.get('nly_therm', p['nly'])} {p.get('ltrot_therm', p['ltrot'])}",
        f"{p['nwrap']} {p['nbin']} {p.get('nsweep', 1)} {p['shift_loc']!r}",
        f"{'.true.' if p.get('is_tau') else '.false.'} {p.get('nthermal', 0)}",
        f"{'.true.' if p.get('nwarm', 0) else '.false.'} {p.get('nwarm', 0)} 1.0 1.0",
        f"{p.get('ini_type', 2)} {p.get('ini_ampl', 1.0)!r} {p.get('ini_bias1', 0.0)!r} {p.get('ini_bias2', 0.0)!r}",
    ]
    return "\n".join(rows) + "\n"


def write_inputs(d: Path, p: dict, seeds, fields=None):
    """fields: None (generate from seeds.txt) or array [R, Ltrot, Ns, 2] with explicit values."""
    d.mkdir(parents=True, exist_ok=True)
    (d / "paramC_sets.txt").write_text(param_text(p))
    (d / "seeds.txt").write_text("".join(f"{int(s)}\n" for s in seeds))
    if fields is None:
        (d / "confin.txt").write_text("0\n")
    else:
        lines = []
        for r, block in enumerate(fields):
            lines.append(f"{int(seeds[r])}\n")
            lines.extend(f"{v:.17g}\n" for v in np.asarray(block).ravel())
        (d / "confin.txt").write_text("".join(lines))


def read_rows(path: Path):
    rows = []
    for line in path.read_text().splitlines():
        if line.strip():
            rows.append([float(t.replace("D", "E").replace("d", "e")) for t in line.split()])
    return np.array(rows)


def read_confout(path: Path, nchains: int, nvals: int):
    toks = [t for t in path.read_text().split()]
    out_seeds, out_fields = [], []
    pos = 0
    for _ in range(nchains):
        out_seeds.append(int(toks[pos])); pos += 1
        out_fields.append(np.array([float(t.replace("D", "E")) for t in toks[pos:pos + nvals]]))
        pos += nvals
    return out_seeds, out_fields


def _big_stack():
    """The reference keeps N x N work matrices as automatic (stack) arrays: at
    N = 576 (12x12 pairing) they exceed the usual 8 MiB stack, so the ranks run
    with an unlimited stack, as `ulimit -s unlimited` would give them."""
    import resource
    try:
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    except (ValueError, OSError):
        pass


def run(p: dict, seeds, fields=None, workdir=None, nranks=None, timeout=600):
    """Run the reference solver; return (directory, wall seconds)."""
    nranks = nranks or len(seeds)
    d = Path(workdir) if workdir else Path(tempfile.mkdtemp(prefix="bafq-ref-"))
    write_inputs(d, p, seeds, fields)
    env = os.environ.copy()
    for k in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS"):
        env[k] = "1"
    cmd = [MPIRUN, "-np", str(nranks), str(executable(p.get("pairing", False)))]
    t0 = time.perf_counter()
    res = subprocess.run(cmd, cwd=d, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=timeout, preexec_fn=_big_stack)
    wall = time.perf_counter() - t0
    if res.returncode:
        raise RuntimeError(f"reference run failed ({res.returncode}):\n{res.stdout[-4000:]}")
    return d, wall


def cleanup(d: Path):
    shutil.rmtree(d, ignore_errors=True)
