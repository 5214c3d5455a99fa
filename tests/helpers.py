"""Shared test helpers: oracle/solver runners and comparison utilities."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from oracle import bafq_oracle as O
from bafqmc_amd.model import Params, ModelSpec
from bafqmc_amd.solver import BatchedSolver

GOLDEN = Path(__file__).resolve().parent / "golden"


def fixed_fields(lt, ns):
    t = np.arange(1, lt + 1, dtype=float)[:, None]
    s = np.arange(1, ns + 1, dtype=float)[None, :]
    dens = 0.23 + 0.31 * np.sin(0.71 * t + 0.37 * s) + 0.09 * np.cos(0.53 * t * s)
    rel = -0.19 + 0.28 * np.cos(0.43 * t - 0.61 * s) + 0.13 * np.sin(0.29 * t * s)
    return np.stack((rel, dens), axis=-1)


def oracle_params(d: dict) -> O.Params:
    keys = ("pairing", "u1", "u2", "mu", "nlx", "nly", "ltrot", "beta", "nwrap", "shift_loc", "rdelta")
    return O.Params(**{k: d.get(k, 0.0 if k == "rdelta" else None) for k in keys})


def solver_params(d: dict) -> Params:
    p = Params(pairing=d["pairing"], u1=d["u1"], u2=d["u2"], mu=d["mu"], rdelta=d.get("rdelta", 0.0),
               nlx=d["nlx"], nly=d["nly"], ltrot=d["ltrot"], beta=d["beta"], nwrap=d["nwrap"],
               shift_loc=d["shift_loc"], nbin=d.get("nbin", 1), ini_type=d.get("ini_type", 2),
               ini_ampl=d.get("ini_ampl", 1.0), is_warm=bool(d.get("nwarm", 0)), nwarm=d.get("nwarm", 0))
    p.nlx_therm = d.get("nlx_therm", p.nlx)
    p.nly_therm = d.get("nly_therm", p.nly)
    p.ltrot_therm = d.get("ltrot_therm", p.ltrot)
    return p.validate()


def reduced(g4, ns):
    """{b, c+} block of a 4Ns Nambu matrix (rows/cols b_1..b_Ns, c+_1..c+_Ns)."""
    idx = [*range(ns), *range(3 * ns, 4 * ns)]
    return g4[np.ix_(idx, idx)]


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def compare_bins(mine: dict, ref: dict, keys=None):
    """Worst relative error over observables (complex: scale by |value|)."""
    worst, where = 0.0, None
    for k, v in ref.items():
        if keys is not None and k not in keys:
            continue
        if k not in mine or k.startswith("_"):
            continue
        e = rel_err(mine[k], v)
        if e > worst:
            worst, where = e, k
    return worst, where


def load_golden(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    params = json.loads(str(z["params"]))
    return z, params


def golden_row(z, name, b):
    """Row b of an output file as an observable value (complex for 2 columns)."""
    row = z["out_" + name][b]
    if row.size == 2 and name in ("sf_K", "psf_Gamma", "dw_K", "den_upup_sub11", "den_dodo_sub11", "den_updo"):
        return complex(row[0], row[1])
    if name == "density_site_total":
        return row
    return float(row[0])


def golden_info_value(z, label: str) -> float:
    """Number after `label :` in the reference's info.txt stored with a golden case."""
    for line in str(z["info"]).splitlines():
        if line.strip().startswith(label):
            return float(line.split(":")[1].replace("D", "E"))
    raise KeyError(label)


def run_solver(lib, params: Params, phis, seeds, nbin=1, nsweep=1, prepare=True):
    s = BatchedSolver([ModelSpec(params)], [0] * len(seeds), lib=lib)
    s.set_fields(phis, seeds)
    if prepare:
        s.prepare()
    bins = []
    for _ in range(nbin):
        s.sweep(nsweep)
        bins.append(s.read_bin())
    return s, bins


def chain_average(bins_of_chains, key):
    vals = [b[key] for b in bins_of_chains]
    return sum(vals) / len(vals)


# Observables are sums of products of Green-function entries of magnitude O(1);
# low-density estimators (density = G_ii - 1) inherit absolute, not relative,
# round-off of G.  Tolerance: |a - b| <= OBS_RTOL |b| + OBS_ATOL.
OBS_RTOL = 1e-9
OBS_ATOL = 1e-12
# Per-site densities are single diagonal Green entries minus their O(1) offset,
# with no averaging over sites: they carry the absolute error of one G entry,
# bounded by the 1e-10 relative tolerance on G (max |G| = O(1)).  The pairing
# propagators are non-normal (cosh/sinh blocks): the reference's own wrap drift
# ("Max diff Matrix") is ~1e-6 on its pairing runs against ~1e-12 for NC, and
# the NumPy oracle, a third implementation, differs from the Fortran by up to
# 1.4e-11 absolute on pairing observables; there the absolute bound is 1e-10.
SITE_ATOL = 1e-10
PAIRING_ATOL = 1e-10


def obs_atol(key: str, pairing: bool = False) -> float:
    return SITE_ATOL if (key == "density_site_total" or pairing) else OBS_ATOL


def assert_obs_close(v, ref, label, rtol=OBS_RTOL, atol=OBS_ATOL):
    """Elementwise |v - ref| <= rtol |ref| + atol; complex values use the modulus
    (their imaginary parts are round-off in both codes)."""
    v = np.asarray(v)
    ref = np.asarray(ref)
    err = np.abs(v - ref)
    bound = rtol * np.abs(ref) + atol
    if not np.all(err <= bound):
        raise AssertionError(f"{label}: |diff| {np.max(err):.3e} > bound {np.min(bound):.3e}; "
                             f"got {v!r} want {ref!r}")


# ------------------------------------------------------------- statistics
STAT_OBS = {"density_total": ("density_total", False), "energy_density": ("energy_density", False),
            "sf_K": ("sf_K", True), "dw_K": ("dw_K", True)}


def run_statistics(lib, params: Params, chains: int, therm: int, meas: int, seed: int, read_every: int = 20):
    """`chains` independent chains of one model, seeded from chain_seeds(seed):
    the Gaussian warm-up of the input (Local_sweep_therm), Local_sweep_pre,
    `therm` unmeasured sweeps, then `meas` measured sweeps.  Returns the
    per-chain time averages {observable: [chains]} (real parts), every chain
    counted, and per-chain diagnostics (largest compare_mat drift, whether every
    value stayed finite)."""
    from bafqmc_amd.analysis import chain_seeds
    from bafqmc_amd.model import conf_set_batch
    from bafqmc_amd.solver import COMPLEX_SLOTS, OBS_SLOTS, O_NOBS, STATS
    phis, seeds = conf_set_batch(params, chain_seeds(seed, chains))
    s = BatchedSolver([ModelSpec(params)], [0] * chains, lib=lib)
    try:
        s.set_fields(list(phis), seeds)
        if params.nwarm:
            s.warmup(params.nwarm)
        s.prepare()
        if therm:
            s.sweep(therm, measure=False)
        s.read_bin_raw()
        sums = np.zeros((chains, len(STAT_OBS)))
        nobs = np.zeros(chains)
        drift = np.zeros(chains)
        done = 0
        while done < meas:
            k = min(read_every, meas - done)
            s.sweep(k)
            obs, st = s.read_bin_raw()
            for j, (slot_name, cplx) in enumerate(STAT_OBS.values()):
                sums[:, j] += obs[:, COMPLEX_SLOTS[slot_name] if cplx else OBS_SLOTS[slot_name]]
            nobs += obs[:, O_NOBS]
            drift = np.maximum(drift, st[:, STATS["xmax"]])
            done += k
    finally:
        s.close()
    means = sums / nobs[:, None]
    return ({k: means[:, j] for j, k in enumerate(STAT_OBS)},
            {"drift": drift, "finite": np.all(np.isfinite(means), axis=1)})


def mean_and_error(x):
    """Grand mean and its standard error from the spread of independent chains' means."""
    x = np.asarray(x, dtype=float)
    return float(np.mean(x)), float(np.std(x, ddof=1) / np.sqrt(len(x)))
