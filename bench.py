#!/usr/bin/env python3
"""Effective S_SF(K) samples/s of the batched BAFQMC sampler on one MI355X,
learned-proposal form vs the plain batched sweep at equal error bars.

Workload (default, `--config nc9x9`): 256 independent number-conserving chains
on a 9x9 triangular lattice (N = 81), t = 1, U1 = 1, U2 = 0, mu = -5,
beta = 8, dtau = 0.1 (Ltrot = 80), Nwrap = 10, shiftLoc = 1.5, fields i.i.d.
N(0,1) from per-chain seeds; proposal network: 4 periodic 3x3x3 convolution
layers, 32 hidden channels, FP32, He-normal initialisation from a seed.
`--config pairing12x12` runs the 12x12 pairing workload (N = 288, Ltrot = 100,
128 chains) the same way.

* plain step: one batched sweep (sweep_L + sweep_R of every chain, 2 Ltrot
  measurements each), then the bin is read;
* learned step: one Hamiltonian global move of every chain (momenta drawn on
  the device, the Gaussian part of log W integrated exactly, the determinant
  part by the network's learned force, exact Metropolis-Hastings with the
  fresh FP64 determinant; bafq_hmc_propose / bafq_global_finish), then the
  estimators on the move's rebuilt Green's function G(beta) (`--measure
  current`; `all` runs a left pass with frozen fields measuring every slice,
  `checkpoints` measures at the stabilisation points; bafq_measure_pass), then
  the bin is read.  Every slice's measurement estimates the same equal-time
  averages; the comparison at equal error bars charges each form its own
  variance per step.

The network is trained before the learned phase on the residual force
d log W / d phi + phi of fields from the plain sampler (the device's gradient
pass).  Each form runs `--steps` K timed steps (the timed region: throughput)
inside a longer run of `--stat-steps` S steps of the same form (the
statistics: tau_int with its window and error, and the standard error of the
mean from the spread of the chains' time averages).  Every chain is counted.

value: plain-sweep-equivalent effective independent S_SF(K) samples/s of the
learned form at equal error bars, i.e. sigma^2 / (SE_learned^2 T_learned)
with sigma^2 the per-step variance of the plain form's S_SF(K) and T the time
the S statistics steps take at the timed rate; config.gain_at_equal_error_bars
is the learned/plain ratio of 1 / (SE^2 T).

Prints one JSON line.  ``--gpus N`` with N > 1 exits: one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
UNSTABLE_DRIFT = 5.5e-5
PEAK_HBM_GBS = 8000.0

CONFIGS = {
    "nc9x9": dict(name="number-conserving 9x9 beta=8 Ltrot=80 U1=1 U2=0 mu=-5, 256 chains, "
                    "4-layer 32-channel FP32 proposal network",
               p=dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=9, nly=9, ltrot=80, beta=8.0, nwrap=10,
                      shift_loc=1.5), chains=256, net=dict(layers=4, hidden=32), slug="nc9x9_256chains"),
    "pairing12x12": dict(name="pairing 12x12 beta=10 Ltrot=100 U1=1 U2=-0.5 Delta=0.3 mu=-5, 128 chains, "
                    "4-layer 32-channel FP32 proposal network",
               p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.3, nlx=12, nly=12, ltrot=100, beta=10.0,
                      nwrap=10, shift_loc=1.0), chains=128, net=dict(layers=4, hidden=32),
               # a strong determinant force: several integrator steps, a longer fit, and measurements at
               # the stabilisation points (a single slice's estimate varies 10x more per step than a sweep's)
               learned=dict(hmc_steps=[4, 8, 16], train_s=60.0, epochs=60, measure="checkpoints"),
               slug="pairing12x12_128chains"),
    "pairing6x6": dict(name="pairing 6x6 beta=8 Ltrot=80 U1=1 U2=-0.5 Delta=0.5 mu=-5, 512 chains",
               p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=6, nly=6, ltrot=80, beta=8.0,
                      nwrap=10, shift_loc=1.0), chains=512, net=dict(layers=4, hidden=32), slug="pairing6x6_512chains"),
    "uneven": dict(name="uneven batch: number-conserving 3x3/6x6/9x9/12x12 x beta 4/8 x U2 0/-0.5, "
                        "U1=1 mu=-5 dtau=0.1, 16 chains each, 256 chains",
                   models=[dict(pairing=False, u1=1.0, u2=u2, mu=-5.0, nlx=L, nly=L, ltrot=int(round(b / 0.1)),
                                beta=b, nwrap=10, shift_loc=1.5)
                           for L in (3, 6, 9, 12) for b in (4.0, 8.0) for u2 in (0.0, -0.5)],
                   chains=256),
    "nc12x12": dict(name="number-conserving 12x12 beta=8 Ltrot=80 U1=1 U2=-0.5 mu=-5, 256 chains",
                 p=dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=12, nly=12, ltrot=80, beta=8.0, nwrap=10,
                        shift_loc=1.5), chains=256, net=dict(layers=4, hidden=32), slug="nc12x12_256chains"),
    "nc3x3": dict(name="number-conserving 3x3 beta=4 Ltrot=40 U1=1 U2=0 mu=-5, 512 chains",
               p=dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=3, nly=3, ltrot=40, beta=4.0, nwrap=10,
                      shift_loc=1.5), chains=512, net=dict(layers=3, hidden=8)),
}

METRIC = "effective S_SF(K) samples/sec, learned vs plain sweep, 1 MI355X at equal σ"


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def nch_of(p):
    return int(abs(p.u1) > 1e-10) + int(abs(p.u2) > 1e-10)


def flops_per_chain_sweep(n, ns, ltrot, nwrap, nch, c_meas=40.0):
    """Algorithmic FP64 flops of one chain-sweep (complex MAC = 8 flops):
    3 kinetic GEMMs per slice step (24 N^3), Ns rank-k Sherman-Morrison updates
    per channel (8 N^3), 56 N^3 per stabilisation, measurement c_meas Ns^2."""
    return (2 * ltrot * (24 + 8 * nch) * n ** 3 + 2 * (ltrot / nwrap) * 56 * n ** 3
            + 2 * ltrot * c_meas * ns ** 2)


def flops_per_measure_pass(n, ns, ltrot, nwrap, c_meas=40.0):
    """One left pass with frozen fields: the kinetic GEMMs and stabilisations of
    sweep_L, no updates, one measurement per slice."""
    return ltrot * 24 * n ** 3 + (ltrot / nwrap) * 56 * n ** 3 + ltrot * c_meas * ns ** 2


def flops_per_checkpoint_pass(n, ns, ltrot, nwrap, c_meas=40.0):
    """Measurement at the stabilisation points only: one B product per slice on
    the U factor (8 N^3) and one Wrap_L stabilisation (56 N^3) per checkpoint,
    one measurement per checkpoint."""
    return ltrot * 8 * n ** 3 + (ltrot / nwrap) * (56 * n ** 3 + c_meas * ns ** 2)


def flops_per_prepare(n, ltrot, nwrap):
    """Local_sweep_pre: one B product per slice on the U factor (8 N^3), stab_UR per
    checkpoint (QR 16/3 + form Q 16/3 + triangular product 4 N^3), one stab_green (41.3 N^3)."""
    return ltrot * 8 * n ** 3 + (ltrot / nwrap) * 14.7 * n ** 3 + 41.3 * n ** 3


def net_flops(desc, batch):
    """FP32 flops of one forward pass of the convolution network."""
    vol = desc["lt"] * desc["ly"] * desc["lx"]
    tot = 0
    for l in range(desc["layers"]):
        cin = desc["channels"] if l == 0 else desc["hidden"]
        cout = desc["channels"] if l == desc["layers"] - 1 else desc["hidden"]
        tot += 2 * 27 * cin * cout * vol
    return tot * batch


def measured_traffic(cfg, kernel, nwrap, chains=None, schedule=None, pairs=None):
    """HBM bytes per launch of a kernel from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated) of this workload at this Nwrap (and,
    where the summary records them, this chain count, schedule and number of
    workgroup pairs), when there is one."""
    if "slug" not in cfg:
        return None
    f = ROOT / "profiles" / f"{kernel}_{cfg['slug']}_hbm.json"
    if not f.exists():
        return None
    rec = json.loads(f.read_text())
    if rec.get("nwrap", nwrap) != nwrap or rec.get("mu", cfg["p"]["mu"]) != cfg["p"]["mu"]:
        return None
    if chains is not None and rec.get("chains", chains) != chains:
        return None
    if schedule is not None and "schedule" in rec and list(rec["schedule"]) != list(schedule):
        return None
    if pairs is not None and rec.get("pairs", 0) != pairs:
        return None
    return rec["traffic_bytes_per_launch"]


def make_params(d):
    from bafqmc_amd.model import Params
    p = Params(pairing=d["pairing"], u1=d["u1"], u2=d["u2"], mu=d["mu"], rdelta=d.get("rdelta", 0.0),
               nlx=d["nlx"], nly=d["nly"], ltrot=d["ltrot"], beta=d["beta"], nwrap=d["nwrap"],
               shift_loc=d["shift_loc"], ini_type=2, ini_ampl=1.0)
    p.nlx_therm, p.nly_therm, p.ltrot_therm = p.nlx, p.nly, p.ltrot
    return p.validate()


def cpu_baseline(cfg, tau, budget_s=12.0):
    """The reference Fortran solver (oracle/_ref) on the host cores, bounded sample."""
    d = cfg["p"]
    cores = max(1, min(16, os.cpu_count() or 1))
    try:
        from oracle import run_ref
        if not run_ref.available(d["pairing"]):
            raise FileNotFoundError("reference binary not built")
        from bafqmc_amd.analysis import seeds_for
        # calibrate with one bin on one rank, then size the bounded run
        pd = dict(d, nbin=1, nsweep=1, ini_type=2, ini_ampl=1.0)
        dd, w1 = run_ref.run(pd, seeds_for(4242, 1))
        run_ref.cleanup(dd)
        nbin = max(1, int(budget_s / max(w1, 1e-3)))
        pd["nbin"] = nbin
        dd, wall = run_ref.run(pd, seeds_for(4242, cores), nranks=cores, timeout=600)
        run_ref.cleanup(dd)
        rate = cores * nbin / wall
        return {"value": rate / (2.0 * tau), "unit": "effective S_SF(K) samples/s",
                "cores": cores, "kind": "reference",
                "sample": (f"reference Fortran (amdflang+MKL sequential) mpirun -np {cores}, one chain per rank, "
                           f"{nbin} bins x 1 sweep of the same workload in {wall:.1f} s = {rate:.3f} chain-sweeps/s; "
                           f"divided by 2*tau_int={2 * tau:.2f} of the plain GPU run (same Markov chain)")}
    except Exception as e:  # noqa: BLE001 - fall back to the NumPy port of the oracle
        from oracle import bafq_oracle as O
        p = O.Params(**{k: d[k] for k in ("pairing", "u1", "u2", "mu", "nlx", "nly", "ltrot", "beta", "nwrap",
                                             "shift_loc")}, rdelta=d.get("rdelta", 0.0))
        rng = np.random.default_rng(1)
        ch = O.Chain(p, rng.normal(size=(p.ltrot, p.nlx * p.nly, 2)), 12345)
        t0 = time.perf_counter()
        ch.pre()
        n = 0
        while time.perf_counter() - t0 < budget_s:
            ch.bin()
            n += 1
        wall = time.perf_counter() - t0
        rate = n / wall
        return {"value": rate / (2.0 * tau), "unit": "effective S_SF(K) samples/s", "cores": 1, "kind": "port",
                "sample": f"NumPy oracle, 1 chain, {n} sweeps in {wall:.1f} s ({rate:.3f} chain-sweeps/s); "
                          f"reference binary unavailable: {e}"}


def roofline(name, flops, kernel_ms, bound="mfma", peak=PEAK_FP64_TFLOPS, traffic=None):
    avg = float(np.mean(kernel_ms))
    achieved = flops / (avg * 1e-3) / 1e12 if avg > 0 else float("nan")
    return {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "kernel_ms": avg, "flops_per_launch": flops}


def choose_schedule(args, p, nchains):
    if args.schedule == "auto":
        from bafqmc_amd import autotune
        from bafqmc_amd.native import load_library
        tuned = autotune.tune_model(load_library(), p, chains=nchains, sweeps=2, tune_nwrap=args.tune_nwrap)
        return [tuned.schedule], tuned
    if args.schedule == "default":
        return None, None
    return [tuple(int(v) for v in args.schedule.split(","))], None


PHASE_SLOTS = {"site": 12, "kinetic": 13, "measure": 14, "wrap": 15, "hs": 16, "wrap_products": 17,
               "wrap_stab_u": 18, "wrap_green": 19}


def run_phase(s, step, K, S, label, phases=None):
    """max(K, S) steps of one form; the first K are the timed region.  Returns
    (per-chain S_SF(K) series [chains, max(K, S)], wall seconds of the K timed
    steps, per-step kernel milliseconds reported by `step`).  phases: a dict
    that receives the device timers of the K timed steps (ms per chain-sweep,
    thread 0 of each chain's workgroup)."""
    n = max(K, S)
    sf = np.zeros((s.nchains, n))
    kms = []
    tim = np.zeros(len(PHASE_SLOTS))
    wall_k = float("nan")
    progress(f"{label}: {K} timed steps, {n} steps for the statistics")
    t0 = time.perf_counter()
    for k in range(n):
        kms.append(step())
        obs, st = s.read_bin_raw()
        if k < K:
            tim += st[:, list(PHASE_SLOTS.values())].sum(axis=0)
        sf[:, k] = obs[:, 20] / obs[:, 32]
        if k + 1 == K:
            wall_k = time.perf_counter() - t0
        if (k + 1) % 50 == 0:
            progress(f"{label} step {k + 1}/{n}")
    if phases is not None:
        phases.update({name: float(v) * 1e3 / (K * s.nchains) for name, v in zip(PHASE_SLOTS, tim)})
    return sf, wall_k, kms


def statistics(sf, wall_k, K, label):
    """Throughput of the timed steps and statistics of the whole series."""
    from bafqmc_amd.analysis import chain_mean_error, tau_int_pooled
    finite = np.all(np.isfinite(sf), axis=1)
    x = sf if finite.all() else sf[finite]
    tau = tau_int_pooled(x)
    err = chain_mean_error(x)
    steps_per_s = K / wall_k
    t_stat = x.shape[1] / steps_per_s                   # time the statistics steps take at the timed rate
    var = float(np.var(x))
    return {"ms_per_step": 1000.0 * wall_k / K, "steps_per_s": steps_per_s,
            "tau_int_sf_k": tau["tau"], "tau_err": tau["err"], "tau_window": tau["window"],
            "stat_steps": tau["length"], "series_shorter_than_50_tau": tau["short"],
            "eff_samples_per_s": len(x) * steps_per_s / (2.0 * tau["tau"]),
            "mean_sf_k": err["mean"], "stderr_sf_k": err["stderr"], "stderr_sf_k_blocked": err["stderr_blocked"],
            "variance_per_step": var, "inverse_variance_per_s": 1.0 / (err["stderr"] ** 2 * t_stat),
            "nonfinite_chains": int((~finite).sum())}


def run_learned(args, cfg, p, spec, nchains):
    """Plain sweeps, then the learned form on the same chains (it continues
    from the plain phase's fields), compared at equal error bars."""
    from bafqmc_amd import global_moves as GM
    from bafqmc_amd.analysis import chain_seeds
    from bafqmc_amd.model import conf_set_batch
    from bafqmc_amd.proposal import ConvNet, collect_forces, train_on_sweeps
    from bafqmc_amd.solver import BatchedSolver
    phis, seeds = conf_set_batch(p, chain_seeds(13579113, nchains))
    schedule, tuned = choose_schedule(args, p, nchains)
    if tuned is not None and tuned.nwrap != p.nwrap:
        import dataclasses
        from bafqmc_amd.model import ModelSpec
        p = dataclasses.replace(p, nwrap=tuned.nwrap)
        spec = ModelSpec(p)
    s = BatchedSolver([spec], [0] * nchains, schedule=schedule)
    s.set_fields(list(phis), seeds)
    s.prepare()
    for _ in range(args.warmup):
        s.sweep(1)
        s.read_bin()
    K, S = args.steps, args.stat_steps

    def plain_step():
        s.sweep(1)
        return s.last_kernel_ms()

    sf0, wall0, sweep_ms = run_phase(s, plain_step, K, S, "plain")
    st_plain = s.stats()
    # ---- network: residual force of fields from the plain sampler
    hidden = args.hidden or cfg["net"]["hidden"]
    layers = args.layers or cfg["net"]["layers"]
    ntrain = max(2, min(12, 4096 // nchains + 1))
    progress(f"collecting the force of {ntrain} sweeps of fields")
    t1 = time.perf_counter()
    x, t = collect_forces(s, p, ntrain)
    net = ConvNet(p.ltrot, p.nly, p.nlx, 2, hidden, layers, max_batch=min(256, nchains), seed=1)
    progress(f"training the network ({net.nparam} parameters) on {len(x)} fields")
    # targets normalised per channel; the move multiplies the output back
    t64 = t.astype(np.float64)
    force_rms = np.sqrt(np.mean(t64 ** 2, axis=(0, 2, 3, 4)))
    fscale = np.where(force_rms > 0, force_rms, 1.0)
    tn = (t64 / fscale[None, :, None, None, None]).astype(np.float32)
    _, r2 = train_on_sweeps(net, x, tn, args.epochs, 1e-3, np.random.default_rng(0), batch=32,
                            budget_s=args.train_s)
    t_train = time.perf_counter() - t1
    force_rms = float(np.sqrt(np.mean(t64 ** 2)))
    progress(f"trained in {t_train:.1f} s, held-out R2 {r2:.4f}, residual force rms {force_rms:.3g}")
    rng = np.random.default_rng(2468)
    # ---- integrator: steps of a quarter rotation in total, the fastest by trial
    fields, sd = s.get_fields()
    trials = {}
    for steps in args.hmc_steps:
        s.set_fields(fields, sd)
        s.prepare()
        acc, tm = 0.0, time.perf_counter()
        for _ in range(3):
            acc += float(GM.hmc_global_device(s, net, 0.5 * math.pi / steps, steps, rng, force_scale=fscale).mean())
            s.measure_pass(1, mode=args.measure)
        trials[steps] = (acc / 3, (time.perf_counter() - tm) / 3)
    s.read_bin()
    hmc_steps = max(trials, key=lambda k: trials[k][0] / trials[k][1])
    angle = 0.5 * math.pi / hmc_steps
    s.set_fields(fields, sd)
    s.prepare()
    s.read_bin()
    accepted = []
    phase_ms = {"move_prepare": [], "measure": []}

    def learned_step():
        a = GM.hmc_global_device(s, net, angle, hmc_steps, rng, force_scale=fscale)
        phase_ms["move_prepare"].append(s.last_kernel_ms())
        accepted.append(float(a.mean()))
        s.measure_pass(1, mode=args.measure)
        ms = s.last_kernel_ms()
        phase_ms["measure"].append(ms)
        return ms

    sf1, wall1, _ = run_phase(s, learned_step, K, S, f"learned ({hmc_steps} step(s) of angle {angle:.3f})")
    st_learned = s.stats()
    # ---- the same move without the network (force 0): what the network adds
    accepted_nn = []

    def gaussian_step():
        accepted_nn.append(float(GM.hmc_global_device(s, None, angle, hmc_steps, rng).mean()))
        s.measure_pass(1, mode=args.measure)
        return s.last_kernel_ms()

    sf2, wall2, _ = run_phase(s, gaussian_step, K, S, "same move without the network")
    conv_ms = net.time_forward(min(256, nchains), 5)
    desc = dict(lt=p.ltrot, ly=p.nly, lx=p.nlx, channels=2, hidden=hidden, layers=layers)
    a0 = statistics(sf0, wall0, K, "plain")
    a1 = statistics(sf1, wall1, K, "learned")
    a2 = statistics(sf2, wall2, K, "no network")
    gain = a1["inverse_variance_per_s"] / a0["inverse_variance_per_s"]
    value = a1["inverse_variance_per_s"] * a0["variance_per_step"]
    nch = nch_of(p)
    kernels = {
        "k_sweep_lds": roofline("k_sweep_lds", flops_per_chain_sweep(p.n, p.ns, p.ltrot, p.nwrap, nch) * nchains,
                                sweep_ms),
        "k_measure_lds": roofline("k_measure_lds", {"all": flops_per_measure_pass(p.n, p.ns, p.ltrot, p.nwrap),
                                                 "checkpoints": flops_per_checkpoint_pass(p.n, p.ns, p.ltrot, p.nwrap),
                                                 "current": 40.0 * p.ns ** 2}[args.measure] * nchains,
                                  phase_ms["measure"]),
        "k_prepare_lds": roofline("k_prepare_lds", flops_per_prepare(p.n, p.ltrot, p.nwrap) * nchains,
                                  phase_ms["move_prepare"]),
        "k_conv_mfma (one forward pass)": roofline("k_conv_mfma", net_flops(desc, min(256, nchains)), [conv_ms],
                                                   peak=PEAK_FP32_TFLOPS),
    }
    if s.schedule()[0][0] != 1:
        for k0, k1 in (("k_sweep_lds", "k_sweep"), ("k_measure_lds", "k_measure"), ("k_prepare_lds", "k_prepare")):
            kernels[k1] = dict(kernels.pop(k0), kernel=k1)
    # per learned step: network passes (one per integrator step and chunk of chains: the output at the
    # current fields is kept from the previous move), G rebuild (prepare), measurement
    share = {"measure": float(np.mean(phase_ms["measure"])), "prepare": float(np.mean(phase_ms["move_prepare"])),
             "conv": conv_ms * hmc_steps * math.ceil(nchains / min(256, nchains))}
    dominant = max(share, key=share.get)
    top = {"measure": kernels.get("k_measure_lds", kernels.get("k_measure")),
           "prepare": kernels.get("k_prepare_lds", kernels.get("k_prepare")),
           "conv": kernels["k_conv_mfma (one forward pass)"]}[dominant]
    top = dict(top)
    top["traffic"] = measured_traffic(cfg, top["kernel"], p.nwrap)
    unstable = int(np.sum(np.maximum(st_plain[:, 4], st_learned[:, 4]) >= UNSTABLE_DRIFT))
    if args.dump:
        np.savez(args.dump, sf_plain=sf0, sf_learned=sf1, wall_plain=wall0, wall_learned=wall1)
    out = {
        "metric": METRIC,
        "value": value, "unit": "plain-sweep-equivalent effective S_SF(K) samples/s at equal error bars",
        "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall1 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (HS fields i.i.d. N(0,1) from per-chain seeds)",
        "config": {"workload": cfg["name"], "chains": nchains, "counted_chains": nchains,
                   "unstable_chains": unstable,
                   "value_is": (f"learned form: one Hamiltonian global move ({hmc_steps} step(s), exact MH with the "
                                f"FP64 determinant) + one measurement pass ({args.measure}) per step; value = "
                                "sigma^2_plain / (SE_learned^2 T_learned)"),
                   "schedule": {"path": s.schedule()[0][0], "delayed": s.schedule()[0][1], "nwrap": p.nwrap,
                                "autotune_ms_per_sweep": tuned.table if tuned else None},
                   "plain": a0,
                   "learned": dict(a1, global_acceptance=float(np.mean(accepted)), hmc_steps=hmc_steps,
                                   angle=angle, hmc_trials={str(k): v for k, v in trials.items()},
                                   step_ms={k: float(np.mean(v)) for k, v in phase_ms.items()},
                                   network={"layers": layers, "hidden": hidden, "params": net.nparam,
                                            "target": "residual force d log W/d phi + (phi - c)",
                                            "heldout_r2": r2, "target_rms": force_rms,
                                            "force_scale": list(map(float, fscale)), "train_s": t_train,
                                            "train_samples": len(x)}),
                   "without_network": dict(a2, global_acceptance=float(np.mean(accepted_nn))),
                   "gain_of_network_at_equal_error_bars": a1["inverse_variance_per_s"] / a2["inverse_variance_per_s"],
                   "residual_force_rms": force_rms,    # the Gaussian part's force -(phi - c) has rms 1
                   "gain_eff_samples": a1["eff_samples_per_s"] / a0["eff_samples_per_s"],
                   "gain_at_equal_error_bars": gain, "target_gain": 10.0,
                   "kernels": kernels, "learned_step_time_share_ms": share},
        "roofline": top,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, a0["tau_int_sf_k"])
    s.close()
    net.close()
    print(json.dumps(out), flush=True)


def run_plain(args, cfg, p, spec, nchains):
    from bafqmc_amd.analysis import chain_seeds
    from bafqmc_amd.model import conf_set_batch
    from bafqmc_amd.solver import BatchedSolver
    phis, seeds = conf_set_batch(p, chain_seeds(13579113, nchains))
    schedule, tuned = choose_schedule(args, p, nchains)
    if tuned is not None and tuned.nwrap != p.nwrap:
        import dataclasses
        from bafqmc_amd.model import ModelSpec
        p = dataclasses.replace(p, nwrap=tuned.nwrap)
        spec = ModelSpec(p)
    s = BatchedSolver([spec], [0] * nchains, schedule=schedule)
    s.set_fields(list(phis), seeds)
    s.prepare()
    for _ in range(args.warmup):
        s.sweep(1)
        s.read_bin()
    phase = np.zeros(8)

    def step():
        s.sweep(1)
        return s.last_kernel_ms()

    phases = {}
    sf, wall, kms = run_phase(s, step, args.steps, args.stat_steps, "plain", phases)
    st = s.stats()
    a = statistics(sf, wall, args.steps, "plain")
    kern = "k_sweep_lds" if s.schedule()[0][0] == 1 else "k_sweep"
    rf = roofline(kern, flops_per_chain_sweep(p.n, p.ns, p.ltrot, p.nwrap, nch_of(p)) * nchains, kms)
    rf["traffic"] = measured_traffic(cfg, kern, p.nwrap, nchains, s.schedule()[0], s.pairs())
    out = {
        "metric": METRIC, "value": a["eff_samples_per_s"], "unit": "effective S_SF(K) samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": a["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (HS fields i.i.d. N(0,1) from per-chain seeds)",
        "config": {"workload": cfg["name"], "chains": nchains, "matrix_n": p.n, "ltrot": p.ltrot,
                   "nwrap": p.nwrap, "value_is": "plain batched sweep (no global moves)", "plain": a,
                   "unstable_chains": int(np.sum(st[:, 4] >= UNSTABLE_DRIFT)), "kernel": kern,
                   "phase_ms_per_chain_sweep": phases,
                   "schedule": {"path": s.schedule()[0][0], "delayed": s.schedule()[0][1],
                                "workgroup_pairs": s.pairs(),
                                "autotune_ms_per_sweep": tuned.table if tuned else None}},
        "roofline": rf,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, a["tau_int_sf_k"])
    s.close()
    print(json.dumps(out), flush=True)


def run_uneven(args, cfg, nchains):
    """Plain sweeps of a batch mixing lattices, Ltrot and U2 in one handle: the
    LDS path (N <= 96) and the device-memory path (12x12, N = 144) run as two
    persistent kernels in one HIP graph; schedules from the autotuner."""
    from bafqmc_amd import autotune
    from bafqmc_amd.analysis import chain_seeds
    from bafqmc_amd.model import ModelSpec, conf_set_batch
    from bafqmc_amd.native import load_library
    from bafqmc_amd.solver import BatchedSolver
    params = [make_params(d) for d in cfg["models"]]
    chain_model = np.arange(nchains) % len(params)
    t0 = time.perf_counter()
    choices = autotune.tune(load_library(), params, chain_model, sweeps=1, max_sample=16) \
        if args.schedule == "auto" else None
    t_tune = time.perf_counter() - t0
    s = BatchedSolver([ModelSpec(p) for p in params], chain_model,
                      schedule=[c.schedule for c in choices] if choices else None)
    phis, seeds = [], []
    for k, p in enumerate(params):
        idx = np.flatnonzero(chain_model == k)
        f, sd = conf_set_batch(p, chain_seeds(24680 + 97 * k, len(idx)))
        phis.append(list(f))
        seeds.append(list(sd))
    order = [0] * len(params)
    all_f, all_s = [], []
    for c in range(nchains):
        k = chain_model[c]
        all_f.append(phis[k][order[k]])
        all_s.append(seeds[k][order[k]])
        order[k] += 1
    s.set_fields(all_f, all_s)
    s.prepare()
    for _ in range(args.warmup):
        s.sweep(1)
        s.read_bin()

    def step():
        s.sweep(1)
        return s.last_kernel_ms()

    sf, wall, kms = run_phase(s, step, args.steps, args.stat_steps, "uneven batch")
    st = s.stats()
    a = statistics(sf, wall, args.steps, "uneven")
    flops = sum(flops_per_chain_sweep(params[m].n, params[m].ns, params[m].ltrot, params[m].nwrap,
                                      nch_of(params[m])) for m in chain_model)
    out = {
        "metric": METRIC, "value": a["eff_samples_per_s"], "unit": "effective S_SF(K) samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": a["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (HS fields i.i.d. N(0,1) from per-chain seeds)",
        "config": {"workload": cfg["name"], "chains": nchains, "value_is": "plain batched sweep, uneven batch",
                   "plain": a, "unstable_chains": int(np.sum(st[:, 4] >= UNSTABLE_DRIFT)),
                   "sweeps_per_s": nchains * a["steps_per_s"], "autotune_s": t_tune,
                   "schedules": {f"{p.nlx}x{p.nly} b{p.beta:g} U2{p.u2:g}": list(sc)
                                 for p, sc in zip(params, s.schedule())},
                   "paths": {"lds_chains": int(sum(1 for c in range(nchains) if s.schedule()[chain_model[c]][0] == 1)),
                             "hbm_chains": int(sum(1 for c in range(nchains) if s.schedule()[chain_model[c]][0] == 0))}},
        "roofline": roofline("k_sweep_lds + k_sweep (one graph)", flops, kms),
    }
    s.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps of each form")
    ap.add_argument("--warmup", type=int, default=20, help="untimed plain sweeps before the plain phase")
    ap.add_argument("--stat-steps", type=int, default=400,
                    help="steps of each form for the statistics (tau_int, error bars); the timed steps are "
                         "the first --steps of them")
    ap.add_argument("--config", default="nc9x9", choices=sorted(CONFIGS))
    ap.add_argument("--chains", type=int, default=0)
    ap.add_argument("--mu", type=float, default=None, help="chemical potential (default: the workload's)")
    ap.add_argument("--nwrap", type=int, default=None,
                    help="stabilisation interval (default: the workload's, or the autotuner's choice)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--schedule", default="auto",
                    help="'auto' (autotuner, before the timed region), 'default' (library choice) or 'PATH,DELAYED'")
    ap.add_argument("--no-tune-nwrap", dest="tune_nwrap", action="store_false",
                    help="keep the workload's Nwrap (default: the autotuner also chooses Nwrap among the "
                         "divisors of Ltrot near it, admitting only intervals whose measured drift stays "
                         "below 1e-7)")
    ap.add_argument("--hmc-steps", type=lambda v: [int(x) for x in v.split(",")], default=None,
                    help="integrator step counts tried, a quarter rotation in total (default 1,2; the 12x12 "
                         "pairing workload 4,8,16)")
    ap.add_argument("--measure", default=None, choices=["all", "checkpoints", "current"],
                    help="learned step's measurement: every slice, the stabilisation points only, or the "
                         "rebuilt G(beta) of the move once")
    ap.add_argument("--layers", type=int, default=0, help="convolution layers (default: the workload's)")
    ap.add_argument("--hidden", type=int, default=0, help="hidden channels (default: the workload's)")
    ap.add_argument("--epochs", type=int, default=None, help="training epochs (default 20; 12x12 pairing 60)")
    ap.add_argument("--train-s", type=float, default=None,
                    help="training time budget in seconds (default 20; 12x12 pairing 60)")
    ap.add_argument("--dump", default="", help="learned mode: write both phases' per-chain S_SF(K) series")
    ap.add_argument("--mode", default="learned", choices=["plain", "learned"])
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(chains shard across GPUs by construction, but no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 4, "copy_probe_bytes_written": b.numel() * 4}))
        return
    from bafqmc_amd.model import ModelSpec
    cfg = CONFIGS[args.config]
    learned_defaults = dict(dict(hmc_steps=[1, 2], train_s=20.0, epochs=20, measure="current"), **cfg.get("learned", {}))
    for k, v in learned_defaults.items():
        if getattr(args, k) is None:
            setattr(args, k, v)
    nchains = args.chains or cfg["chains"]
    if nchains != cfg["chains"]:
        cfg = dict(cfg, name=cfg["name"].replace(f"{cfg['chains']} chains", f"{nchains} chains"))
    if "models" in cfg:
        run_uneven(args, cfg, nchains)
        return
    if args.mu is not None:
        cfg = dict(cfg, p=dict(cfg["p"], mu=args.mu), name=cfg["name"].replace(f"mu={cfg['p']['mu']:g}", f"mu={args.mu:g}"))
    if args.nwrap is not None:
        cfg = dict(cfg, p=dict(cfg["p"], nwrap=args.nwrap))
    p = make_params(cfg["p"])
    spec = ModelSpec(p)
    if args.mode == "learned":
        run_learned(args, cfg, p, spec, nchains)
    else:
        run_plain(args, cfg, p, spec, nchains)


if __name__ == "__main__":
    main()
