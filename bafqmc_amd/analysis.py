"""Statistics of bin series: blocking errors and integrated autocorrelation time.

``block_statistics`` restates the reference's blocking estimator
(src/number_conserving/benchmarks/campaign_analysis.py:215-232).
``tau_int`` is the integrated autocorrelation time with Sokal's automatic
window (stop at the first W >= c * tau(W), c = 5), estimated from many
independent chains at once: the autocovariance is averaged over chains, each
chain centred on its own mean.
"""
from __future__ import annotations

import math

import numpy as np


def seeds_for(seed: int, count: int):
    """Per-chain seeds as the reference's campaign writer makes them
    (src/pairing/benchmarks/campaign_pairing_delta.py:633-643)."""
    value = int(seed) % 100000000
    if value <= 0:
        value += 1000003
    out = []
    for _ in range(count):
        value = (1664525 * value + 1013904223) % 100000000
        if value <= 0:
            value += 1
        out.append(value)
    return out


def chain_seeds(seed: int, count: int):
    """`count` distinct positive seeds: the campaign writer's sequence while it
    yields new values (its LCG modulo 10^8 with a multiplier divisible by 5
    repeats after a few hundred terms), continued with the solver's own
    generator (ranf, seed <- 48828125 seed mod 2^31) started from the last value."""
    out, seen = [], set()
    for v in seeds_for(seed, count):
        if v in seen:
            break
        seen.add(v)
        out.append(v)
    x = out[-1] if out else (int(seed) % 2147483647) or 1
    while len(out) < count:
        x = (48828125 * x) % 2147483648
        if x > 0 and x not in seen:
            seen.add(x)
            out.append(x)
    return out


def block_statistics(samples, block_size: int, skip: int = 0):
    x = np.asarray(samples, dtype=float)[skip:]
    nb = len(x) // block_size
    if nb < 2:
        raise ValueError("need at least two blocks")
    means = x[: nb * block_size].reshape(nb, block_size).mean(axis=1)
    return {"actual": float(means.mean()), "stderr": float(means.std(ddof=1) / math.sqrt(nb)), "blocks": nb}


def tau_int_pooled(series, c: float = 6.0):
    """Integrated autocorrelation time of many independent chains of one
    stationary process: the autocovariance is averaged over chains about the
    grand mean, Sokal's window stops at the first W >= c tau(W), and the
    standard error is the Madras-Sokal estimate tau sqrt(2 (2 W + 1) / N) with
    N the number of samples.  Returns a dict with tau, its error, the window,
    the number of samples per chain and `short` (series shorter than 50 tau)."""
    x = np.asarray(series, dtype=float)
    if x.ndim == 1:
        x = x[None, :]
    m = float(np.mean(x))
    x = x - m
    nb = x.shape[1]
    var = float(np.mean(x * x))
    out = {"tau": 0.5, "err": 0.0, "window": 0, "length": int(nb), "short": True}
    if var <= 0.0 or nb < 4:
        return out
    tau, w = 0.5, 0
    for w in range(1, nb // 2):
        tau += float(np.mean(x[:, :-w] * x[:, w:])) / var
        if w >= c * tau:
            break
    tau = max(tau, 0.5)
    out.update(tau=tau, window=int(w), err=tau * math.sqrt(2.0 * (2 * w + 1) / x.size),
               short=bool(nb < 50 * tau))
    return out


def chain_mean_error(series):
    """Standard error of the grand mean from the spread of the chains' time
    averages (independent chains; valid when each series is much longer than
    tau), and the blocked standard error of the pooled series
    (campaign_analysis.py:215-232 blocking, blocks of 1/10 of a chain) as a
    cross-check."""
    x = np.asarray(series, dtype=float)
    means = x.mean(axis=1)
    se = float(np.std(means, ddof=1) / math.sqrt(len(means)))
    bs = max(1, x.shape[1] // 10)
    blocked = [block_statistics(row, bs) for row in x] if x.shape[1] >= 2 * bs else []
    se_blk = float(np.sqrt(np.mean([b["stderr"] ** 2 for b in blocked])) / math.sqrt(len(x))) if blocked else float("nan")
    return {"mean": float(means.mean()), "stderr": se, "stderr_blocked": se_blk}


def tau_int(series, c: float = 5.0):
    """series: [chains, bins] real.  Returns (tau_int, window, acf)."""
    x = np.asarray(series, dtype=float)
    if x.ndim == 1:
        x = x[None, :]
    x = x - x.mean(axis=1, keepdims=True)
    nb = x.shape[1]
    var = float(np.mean(x * x))
    if var <= 0.0 or nb < 4:
        return 0.5, 0, np.ones(1)
    acf = [1.0]
    for t in range(1, nb // 2):
        acf.append(float(np.mean(x[:, :-t] * x[:, t:])) / var)
    acf = np.array(acf)
    tau = 0.5
    for w in range(1, len(acf)):
        tau += acf[w]
        if w >= c * tau:
            return max(tau, 0.5), w, acf
    return max(tau, 0.5), len(acf) - 1, acf
