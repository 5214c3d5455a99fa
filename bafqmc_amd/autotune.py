"""Schedule autotuner for the batched sweep (no counterpart in the reference).

For every model of a batch it times one sweep of a sample of chains under each
execution schedule the library accepts (include/bafqmc.h bafq_schedule: G in
LDS or in device memory, delayed-update width 16/12/8/4 (32 with G in device
memory) or per-site updates)
and keeps the fastest.
Every schedule computes the same Markov chain, so the choice changes time
only.  Optionally it also chooses the stabilisation interval Nwrap among the
divisors of Ltrot near the input's: a longer interval saves QR work but lets
rounding grow between stabilisations, so a candidate is admitted only while
the drift that the sweep measures at every stabilisation (compare_mat,
src/number_conserving/src/stabilization.f90:256-266) stays below `drift_tol`
-- far below the 5.5e-5 at which the reference reports an unstable chain.
Changing Nwrap changes results at the level of that drift only.

Timings are device times of the sweep launch (HIP events).  Results can be
kept in a JSON cache keyed by model, sample size and device.  A sample that
is not the whole batch is timed with one workgroup per chain (BAFQ_PAIRS=0):
the library pairs the workgroups of device-memory chains only when the whole
batch leaves CUs idle, which a small sample of a large or mixed batch would
and the batch would not.
"""
from __future__ import annotations

import contextlib
import dataclasses
import json
import os
from pathlib import Path

import numpy as np

from .analysis import chain_seeds
from .model import ModelSpec, conf_set_batch
from .solver import BatchedSolver, SolverError

SCHEDULES = [(1, 16), (1, 12), (1, 8), (1, 4), (1, 0), (0, 32), (0, 16), (0, 8), (0, 4), (0, 0)]


@dataclasses.dataclass
class Choice:
    schedule: tuple
    nwrap: int
    ms_per_sweep: float
    table: dict          # "(path, delayed, nwrap)" -> ms per sweep (None: rejected)


@contextlib.contextmanager
def _unpaired(active):
    """BAFQ_PAIRS=0 while a handle is created (bafq_create reads it)."""
    if not active:
        yield
        return
    old = os.environ.get("BAFQ_PAIRS")
    os.environ["BAFQ_PAIRS"] = "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["BAFQ_PAIRS"]
        else:
            os.environ["BAFQ_PAIRS"] = old


def _time(lib, p, schedule, chains, sweeps, seed, whole=True):
    phis, seeds = conf_set_batch(p, chain_seeds(seed, chains))
    with _unpaired(not whole):
        s = BatchedSolver([ModelSpec(p)], [0] * chains, lib=lib, schedule=[schedule])
    try:
        s.set_fields(list(phis), seeds)
        if p.nwarm:
            s.warmup(min(p.nwarm, 50))
        s.prepare()
        s.sweep(1, measure=False)
        s.read_bin()
        best = float("inf")
        drift = 0.0
        for _ in range(sweeps):
            s.sweep(1)
            best = min(best, s.last_kernel_ms())
            drift = max(drift, max(b["_stats"]["xmax"] for b in s.read_bin()))
        return best, drift, s.schedule()[0]
    finally:
        s.close()


def _nwrap_candidates(p):
    near = [d for d in range(1, p.ltrot + 1) if p.ltrot % d == 0 and p.nwrap // 2 <= d <= 2 * p.nwrap]
    return sorted(set(near) | {p.nwrap})


def tune_model(lib, params, chains=64, sweeps=2, tune_nwrap=False, drift_tol=1e-7, seed=97, whole=True):
    """Fastest (schedule, Nwrap) for one model; params is a bafqmc_amd.model.Params.
    whole: the chains timed are the whole batch (else timed unpaired)."""
    table = {}
    best = None
    for sched in SCHEDULES:
        try:
            ms, drift, eff = _time(lib, params, sched, chains, sweeps, seed, whole)
        except SolverError:
            table[str((*sched, params.nwrap))] = None
            continue
        table[str((*sched, params.nwrap))] = ms
        if tuple(eff) != sched:
            continue                       # the library substituted another width
        if best is None or ms < best[0]:
            best = (ms, sched, params.nwrap)
    if best is None:
        raise SolverError("no schedule could run this model")
    if tune_nwrap:
        for nw in _nwrap_candidates(params):
            if nw == params.nwrap:
                continue
            q = dataclasses.replace(params, nwrap=nw)
            ms, drift, _ = _time(lib, q, best[1], chains, sweeps, seed, whole)
            ok = drift < drift_tol
            table[str((*best[1], nw))] = ms if ok else None
            if ok and ms < best[0]:
                best = (ms, best[1], nw)
    return Choice(schedule=best[1], nwrap=best[2], ms_per_sweep=best[0], table=table)


def _key(params, chains, device):
    """Cache key: every parameter that changes the sweep's cost or its drift
    (the Hamiltonian, the discretisation, the update width and the initial
    fields), the sample size and the device."""
    p = params
    return json.dumps([device, int(p.pairing), p.nlx, p.nly, p.ltrot, p.nwrap, p.beta, p.u1, p.u2, p.mu, p.rdelta,
                       p.shift_loc, p.ini_type, p.ini_ampl, p.ini_bias1, p.ini_bias2, p.nwarm, chains])


def tune(lib, models, chain_model, sweeps=2, tune_nwrap=False, drift_tol=1e-7, cache: Path | None = None,
         device="gfx950", max_sample=256):
    """Choices per model of a batch: the sample of each model is its chain
    count (at most max_sample), so the timing sees the batch's occupancy."""
    counts = np.bincount(np.asarray(chain_model), minlength=len(models))
    stored = json.loads(Path(cache).read_text()) if cache and Path(cache).exists() else {}
    out = []
    for k, m in enumerate(models):
        p = m.params if isinstance(m, ModelSpec) else m
        n = int(max(1, min(counts[k], max_sample)))
        whole = len(models) == 1 and n == counts[k]
        key = _key(p, n, device) + (":nwrap" if tune_nwrap else "") + ("" if whole else ":sample")
        if key in stored:
            c = stored[key]
            out.append(Choice(tuple(c["schedule"]), c["nwrap"], c["ms_per_sweep"], c["table"]))
            continue
        ch = tune_model(lib, p, n, sweeps, tune_nwrap, drift_tol, whole=whole)
        stored[key] = dataclasses.asdict(ch)
        out.append(ch)
    if cache:
        Path(cache).write_text(json.dumps(stored, indent=1))
    return out
