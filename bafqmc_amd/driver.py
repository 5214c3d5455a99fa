"""Drop-in replacement of the reference executables' main program.

``run(directory, solver, nchains)`` reads the reference's input files from a run
directory and writes its output files there, like ``mpirun -np R bosonDQMC.out``
does (src/number_conserving/src/main.f90:22-83; src/pairing/src/main.f90),
with the R MPI ranks replaced by R chains of one batched GPU solver:

* inputs: ``paramC_sets.txt``, ``confin.txt``, ``seeds.txt``
  (src/number_conserving/src/calc_basic.f90:48-91, fields.f90:47-113);
* per bin, one appended row per observable file, averaged over chains like the
  per-bin ``MPI_REDUCE(SUM)/ISIZE`` (src/number_conserving/src/fourier_trans.f90:350-457,
  file list :259-345; pairing src/pairing/src/fourier_trans.f90:275-406);
* number-conserving pole diagnostics per chain and bin (pole_diagnostics.f90:91-187);
* at the end ``info.txt`` and ``confout.txt`` (main.f90:59-76, fields.f90:236-278).

With ``is_tau`` each bin after ``Nthermal`` runs the time-displaced pass
between sweep_L and sweep_R (main.f90:49-55, local_sweep.f90:119-123); its
estimators are empty in the reference (src/number_conserving/src/obser_tau.f90:39-54),
so its output is info.txt's six Wrap_tau drift lines (dynamics.f90:60-82).
``tau_mode``: the bosonic block form (default) or the reference's code as
written (``bafqmc_amd.solver.TAU_REFERENCE``, whose drift lines match the
reference's).
"""
from __future__ import annotations

import dataclasses

import math
import sys
import time
from pathlib import Path

import numpy as np

from .model import ModelSpec, Params, read_params, read_fields, write_fields
from .solver import BatchedSolver, NC_OUTPUTS, PAIRING_OUTPUTS, TAU_BOSONIC

TINY = np.finfo(float).tiny


def _fmt(v: float) -> str:
    return f" {v: .16E}"


def _append(path: Path, text: str):
    with open(path, "a") as f:
        f.write(text)


def write_bin(run_dir: Path, per_chain: list, pairing: bool):
    """Chain-averaged row of every observable file (m_write_obs_equal)."""
    names = PAIRING_OUTPUTS if pairing else NC_OUTPUTS
    nc = len(per_chain)
    for name in names:
        vals = [b[name] for b in per_chain]
        if name == "density_site_total":
            v = np.mean(np.stack(vals), axis=0)
            if pairing:
                line = "".join(f" {x:24.16E}" for x in v) + "\n"
            else:
                line = "".join(f"{x:24.16E} " for x in v) + "\n"
        elif isinstance(vals[0], complex):
            z = sum(vals) / nc
            line = _fmt(z.real) + _fmt(z.imag) + "\n"
        else:
            line = _fmt(sum(vals) / nc) + "\n"
        _append(run_dir / name, line)


def pole_diagnostics(g: np.ndarray, phi: np.ndarray):
    """calc_local_diagnostics (pole_diagnostics.f90:91-132): poles z = 1/eig(G), |.|, SVD."""
    ev = np.linalg.eigvals(g)
    sv = np.linalg.svd(g, compute_uv=False)
    z = np.where(np.abs(ev) <= TINY, 1.0 / TINY + 0j, 1.0 / np.where(ev == 0, 1.0, ev))
    pole_distance = max(float(np.min(np.abs(z))), TINY)
    log_weight = -0.5 * float(np.sum(phi * phi)) + 2.0 * float(np.sum(np.log(np.maximum(sv, TINY))))
    return z, (pole_distance, -math.log10(pole_distance), float(np.max(np.abs(ev))), float(np.max(sv)),
               log_weight)


def write_poles(run_dir: Path, zs: list, scalars: list):
    """append_pole_z / append_scalar_column (pole_diagnostics.f90:134-156)."""
    lines = []
    for z in zs:
        lines.append("".join(f" {x.real:24.16E} {x.imag:24.16E}" for x in z) + "\n")
    _append(run_dir / "pole_z", "".join(lines))
    for col, name in enumerate(("pole_distance", "pole_x", "green_spectral_radius", "green_smax", "log_weight")):
        _append(run_dir / name, "".join(f"{s[col]:24.16E}\n" for s in scalars))


def info_header(p: Params, nchains: int, spectrum) -> str:
    """write_info plus the single-particle spectrum: number conserving
    src/number_conserving/src/calc_basic.f90:138-174 and model.f90:27-29, pairing
    (its own labels, the Delta row) src/pairing/src/calc_basic.f90:140-174 and
    src/pairing/src/model.f90:31-33."""
    if p.pairing:
        title = " DQMC for pairing Bose Hubbard on triangular lattice"
        u_rows = [f" Relative-density U1 (>=0)                    : {p.u1}",
                  f" Total-density U2 (<=0)                       : {p.u2}",
                  f" pairing Delta                                  : {p.rdelta}"]
        warm_rows = [f" Thermalize auxiliary field magnitude for U1    : {p.shift_warm1}",
                     f" Thermalize auxiliary field magnitude for U2    : {p.shift_warm2}"]
    else:
        title = " DQMC for boson Hubbard on triangular lattice"
        u_rows = [f" Hubbard U1 (relative-density, repulsive)       : {p.u1}",
                  f" Hubbard U2 (total-density, attractive)         : {p.u2}"]
        warm_rows = [f" Thermalize U1 relative-field magnitude          : {p.shift_warm1}",
                     f" Thermalize U2 total-field magnitude             : {p.shift_warm2}"]
    rows = [
        " =========================",
        title,
        f" Linear lengh Lx                                : {p.nlx}",
        f" Linear lengh Ly                                : {p.nly}",
        f" Hopping t                                      : {1.0}",
        *u_rows,
        f" chemical potential                             : {p.mu}",
        f" Local update auxiliary field magnitude Shift   : {p.shift_loc}",
    ]
    if p.is_warm:
        rows += [f" # Warm                                         : {p.nwarm}", *warm_rows]
    rows += [
        f" Choosing initial distribution type             : {p.ini_type}",
        f" Magnitude of Gaussian distribution             : {p.ini_ampl}",
        f" Beta                                           : {p.beta}",
        f" Trotter number                                 : {p.ltrot}",
        f" =>Dtau                                         : {p.dtau}",
        f" N_Ortho                                        : {p.nwrap}",
        f" # Bins                                         : {p.nbin}",
        f" # Nsweep in one bin                            : {p.nsweep}",
        f" # Cores                                        : {nchains}",
        f" highest energy level of single particle Ham     : {float(np.max(spectrum))}",
        f" lowest  energy level of single particle Ham     : {float(np.min(spectrum))}",
        f" bandwidth            of single particle Ham     : {float(np.max(spectrum) - np.min(spectrum))}",
    ]
    return "\n".join(rows) + "\n"


@dataclasses.dataclass
class Case:
    """One reference run directory inside a batched launch: its parameters,
    initial fields and the counters info.txt reports."""
    run_dir: Path
    pairing: bool
    nchains: int
    params: Params
    phis: list
    seeds: list
    poles: bool
    spec: ModelSpec = None
    therm: list = None
    acc: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(2))
    proposals: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(2))
    xmax: float = 0.0
    xmean: float = 0.0
    nwrap_total: float = 0.0
    tau_xmax: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    tau_xsum: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    tau_ran: bool = False


def load_case(run_dir, solver: str = "number_conserving", nchains: int = 1, poles: bool | None = None) -> Case:
    """Read_input and conf_in of one run directory (calc_basic.f90:48-106,
    fields.f90:35-113)."""
    run_dir = Path(run_dir)
    pairing = solver in ("pairing", "pa", "nambu")
    p = read_params(run_dir / "paramC_sets.txt", pairing)
    phis, seeds = read_fields(run_dir, p, nchains)
    return Case(run_dir, pairing, nchains, p, list(phis), list(seeds),
                (not pairing) if poles is None else poles)


def _warm(p: Params) -> int:
    return p.nwarm if (p.is_warm and p.nwarm > 0) else 0


def run_cases(cases: list, lib=None, log=sys.stdout, autotune: bool = False, tune_nwrap: bool = False,
              tau_mode: int = TAU_BOSONIC):
    """Run several reference run directories as ONE batched solver: every
    case's chains share the launches (one model per case; lattices, Ltrot,
    Hamiltonians and the two solvers may differ), and each case's output files,
    info.txt and confout.txt are written into its own directory exactly as a
    run of that case alone writes them (a chain evolves in a batch bit for bit
    as alone).  The cases must share the warm-up length, Nbin and Nsweep (the
    launches advance every chain together); "Tot CPU time" is the wall time of
    the whole batch.  Returns one dict of run statistics per case."""
    if not cases:
        return []
    keys = {(_warm(c.params), c.params.nbin, c.params.nsweep) for c in cases}
    if len(keys) != 1:
        raise ValueError("cases in one batch need the same Nwarm, Nbin and Nsweep: "
                         f"{sorted(keys)}")
    nwarm, nbin, nsweep = keys.pop()
    t0 = time.perf_counter()
    schedule = None
    if autotune:
        from . import autotune as tuner
        from .native import load_library
        chain_model = np.repeat(np.arange(len(cases)), [c.nchains for c in cases])
        choices = tuner.tune(lib or load_library(), [c.params for c in cases], chain_model, sweeps=2,
                             tune_nwrap=tune_nwrap, max_sample=64)
        schedule = [ch.schedule for ch in choices]
        for c, ch in zip(cases, choices):
            if ch.nwrap != c.params.nwrap:
                c.params = dataclasses.replace(c.params, nwrap=ch.nwrap)
            print(f"bafqmc_amd: {c.run_dir.name}: schedule {ch.schedule}, N_Ortho {c.params.nwrap} "
                  f"({ch.ms_per_sweep:.2f} ms per sweep)", file=log)
    total = sum(c.nchains for c in cases)
    for c in cases:
        c.spec = ModelSpec(c.params)
        (c.run_dir / "info.txt").write_text(info_header(c.params, c.nchains, c.spec.spectrum))
    bounds = np.cumsum([0] + [c.nchains for c in cases])
    chain_model = [k for k, c in enumerate(cases) for _ in range(c.nchains)]
    s = BatchedSolver([c.spec for c in cases], chain_model, lib=lib, schedule=schedule)
    try:
        s.set_fields([f for c in cases for f in c.phis], [sd for c in cases for sd in c.seeds])
        if nwarm:
            # Local_sweep_therm per warm-up sweep (local_sweep.f90:64-78)
            s.warmup(nwarm)
            therm = s.read_bin()
            for k, c in enumerate(cases):
                c.therm = therm[bounds[k]:bounds[k + 1]]
        else:
            for c in cases:
                _append(c.run_dir / "info.txt", " Skipping Bosonic warm-up\n")
        s.prepare()
        s.read_bin()           # clear prepare's counters
        taus = {(c.params.is_tau, c.params.nthermal) for c in cases}
        if len(taus) != 1:
            raise ValueError(f"cases in one batch need the same is_tau and Nthermal: {sorted(taus)}")
        is_tau, nthermal = taus.pop()
        for b in range(nbin):
            # istau_tmp of main.f90:49-52: the time-displaced pass in bins after Nthermal
            if is_tau and b + 1 > nthermal:
                s.sweep_tau(nsweep, measure=True, mode=tau_mode)
                for c in cases:
                    c.tau_ran = True
            else:
                s.sweep(nsweep, measure=True)
            per = s.read_bin()
            fields = s.get_fields()[0] if any(c.poles for c in cases) else None
            for k, c in enumerate(cases):
                mine = per[bounds[k]:bounds[k + 1]]
                write_bin(c.run_dir, mine, c.pairing)
                for d in mine:
                    st = d["_stats"]
                    if st["prop_u1"] > 0:
                        c.acc[0] += st["acc_u1"] / st["prop_u1"]
                    if st["prop_u2"] > 0:
                        c.acc[1] += st["acc_u2"] / st["prop_u2"]
                    c.proposals += (st["prop_u1"], st["prop_u2"])
                    c.xmax = max(c.xmax, st["xmax"])
                    c.xmean += st["xmean_sum"]
                    c.nwrap_total += st["nwrap"]
                    c.tau_xmax = np.maximum(c.tau_xmax, [st["tau_xmax_0t"], st["tau_xmax_t0"], st["tau_xmax_tt"]])
                    c.tau_xsum += [st["tau_xsum_0t"], st["tau_xsum_t0"], st["tau_xsum_tt"]]
                if c.poles:
                    zs, scal = [], []
                    for j in range(bounds[k], bounds[k + 1]):
                        z, sc = pole_diagnostics(s.green(j), fields[j])
                        zs.append(z)
                        scal.append(sc)
                    write_poles(c.run_dir, zs, scal)
        fields, seeds_out = s.get_fields()
    finally:
        s.close()
    wall = time.perf_counter() - t0
    out = []
    for k, c in enumerate(cases):
        write_fields(c.run_dir, fields[bounds[k]:bounds[k + 1]], seeds_out[bounds[k]:bounds[k + 1]])
        _append(c.run_dir / "info.txt", "\n".join(_info_tail(c, wall)) + "\n")
        out.append({"wall_s": wall, "accept": c.acc / max(c.nchains * c.params.nbin, 1), "xmax": c.xmax})
    print(f"bafqmc_amd: {len(cases)} case(s), {total} chains x {nbin} bins in {wall:.2f} s", file=log)
    return out


def _info_tail(c: Case, wall: float):
    """Therm_control_print, Local_control_print and the end of main
    (local_sweep.f90:163-216, main.f90:59-76)."""
    p = c.params
    tail = []
    if c.therm is not None:
        for k, name in ((0, "U1"), (1, "U2")):
            prop = sum(d["_stats"][f"therm_prop_u{k + 1}"] for d in c.therm)
            if prop > 0:
                accept = sum(d["_stats"][f"therm_acc_u{k + 1}"] for d in c.therm) / prop
                tail.append(f" Thermalize {name} Accept Ratio                     : {accept}")
    denom = max(c.nchains * p.nbin, 1)
    for k, name in ((0, "U1"), (1, "U2")):
        if abs((p.u1, p.u2)[k]) <= 1e-10:
            continue
        tail.append(f" Accept_{name}_shift                                : {c.acc[k] / denom}")
        if c.pairing:
            # Local_control_print (src/pairing/src/local_sweep.f90:163-216): the
            # phase of the local ratio det(P)^(-1/2).  In the reduced 2Ns form the
            # ratio is |det P_2Ns|^-1 times positive Gaussian factors, so its
            # argument is zero and no proposal lands on the pi branch.
            tail += [f" Phase_{name}_mean_abs_arg_proposed                : {0.0}",
                     f" Phase_{name}_mean_abs_arg_accepted                : {0.0}",
                     f" Phase_{name}_branch_pi_fraction                   : {0.0}",
                     f" Phase_{name}_local_proposals                      : {float(c.proposals[k])}"]
    if c.tau_ran:
        # Dyn_control_print (dynamics.f90:60-82): the maxima and the sums over
        # ISIZE * Nbin * Nst * Nsweep of the drifts of G(0,tau), G(tau,0),
        # G(tau,tau), under the reference's labels GRtt, GRt0, GR0t in that order
        n = max(c.nchains * p.nbin * p.nst * p.nsweep, 1)
        for k, lab in enumerate(("GRtt", "GRt0", "GR0t")):
            tail += [f" Max diff {lab}                                  : {c.tau_xmax[k]}",
                     f" Mean diff {lab}                                 : {c.tau_xsum[k] / n}"]
    tail += [f" Max diff Matrix                                : {c.xmax}",
             f" Mean diff Matrix                               : {c.xmean / max(c.nwrap_total, 1.0)}",
             f" Tot CPU time                                   : {wall}"]
    return tail


def run(run_dir, solver: str = "number_conserving", nchains: int = 1, lib=None, poles: bool | None = None,
        log=sys.stdout, autotune: bool = False, tune_nwrap: bool = False, tau_mode: int = TAU_BOSONIC):
    """One reference run in ``run_dir``; returns a dict of run statistics.
    autotune: run the fastest execution schedule (bafqmc_amd.autotune; the same
    Markov chain); tune_nwrap: let it also choose N_Ortho under its drift bound
    (info.txt then reports the N_Ortho used)."""
    return run_cases([load_case(run_dir, solver, nchains, poles)], lib=lib, log=log, autotune=autotune,
                     tune_nwrap=tune_nwrap, tau_mode=tau_mode)[0]
