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
