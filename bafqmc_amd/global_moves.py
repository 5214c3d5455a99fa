"""Global field moves accepted by exact Metropolis-Hastings.

The weight of a field configuration phi (time slices x sites x 2 channels) is
the product of the Gaussian HS measure and the bosonic determinant factor:

* number conserving (two flavours, the second the complex conjugate of the
  first): W = exp(-sum phi^2 / 2) |det G|^2, G = (1 - B(beta, 0))^-1 -- the
  weight whose logarithm src/number_conserving/src/pole_diagnostics.f90:130-131
  reports as log_weight;
* Nambu pairing: W = exp(-sum phi^2 / 2) exp(-alpha_2 sum phi_2) |det G_2Ns|,
  the 4Ns Nambu form's det(1 - B_4Ns)^(-1/2) = |det(1 - B_2Ns)|^-1 of the
  reduced form, with the normal-ordering constant of the U2 channel
  (src/pairing/src/operator_Hubbard.f90:100-137).

Single-site moves of this weight reproduce the local ratios of LocalU_metro
exactly (tests/test_global_moves.py).  The reference declares a global update
(src/number_conserving/src/global_update.f90) but never calls it; the moves
here are new, and the local sweep stays the sampler that they interleave with.

A move: propose phi' for every chain, rebuild G(phi') from scratch (the
stabilised B-product of bafq_prepare gives log|det G'|), accept with
min(1, W'/W q(phi|phi') / q(phi'|phi)), and restore phi for rejected chains.
metropolis_global takes proposals made on the host (any proposal density);
metropolis_global_device draws the learned proposal on the device and
restores rejected chains from a device copy of their state.  The acceptance
uniforms (and host proposals) come from a host generator, the device's normal
numbers from a counter-based generator: the chains' own LCG streams are never
touched.
"""
from __future__ import annotations

import math

import numpy as np

LOGDET = 7          # stats slot of log|det G| (csrc/sweep.hpp S_LOGDET)


def log_weight(params, phi, logdet):
    """log W of one chain's fields given log|det G| (see the module docstring)."""
    lw = -0.5 * float(np.sum(phi * phi))
    if params.pairing:
        dtau = params.beta / params.ltrot
        a2 = math.sqrt(-2.0 * params.u2 * dtau) if params.u2 < -1e-10 else 0.0
        return lw - a2 * float(np.sum(phi[:, :, 1])) + logdet
    return lw + 2.0 * logdet


def log_weights(solver, fields=None):
    """log W of every chain's current fields (G must be current: after
    prepare or a sweep)."""
    if fields is None:
        fields, _ = solver.get_fields()
    st = solver.stats()
    shapes = {np.shape(f) for f in fields}
    if len(shapes) == 1 and len({id(solver.params(c)) for c in range(solver.nchains)}) == 1:
        # one model: all chains at once
        F = np.stack(fields)
        p = solver.params(0)
        lw = -0.5 * np.sum(F * F, axis=(1, 2, 3))
        if p.pairing:
            dtau = p.beta / p.ltrot
            a2 = math.sqrt(-2.0 * p.u2 * dtau) if p.u2 < -1e-10 else 0.0
            return lw - a2 * np.sum(F[:, :, :, 1], axis=(1, 2)) + st[:, LOGDET]
        return lw + 2.0 * st[:, LOGDET]
    return np.array([log_weight(solver.params(c), fields[c], st[c, LOGDET]) for c in range(solver.nchains)])


def metropolis_global(solver, proposals, rng, log_q_ratio=None):
    """One exact global Metropolis-Hastings step for every chain.

    proposals: per-chain proposed fields phi'; log_q_ratio: per-chain
    log q(phi|phi') - log q(phi'|phi) (None for a symmetric proposal).
    Returns the boolean acceptance per chain.  Leaves G current."""
    fields, seeds = solver.get_fields()
    lw0 = log_weights(solver, fields)
    solver.set_fields(proposals, seeds)
    solver.prepare()
    lw1 = log_weights(solver, proposals)
    log_a = lw1 - lw0
    if log_q_ratio is not None:
        log_a = log_a + np.asarray(log_q_ratio)
    u = rng.random(solver.nchains)
    accept = np.log(np.maximum(u, 1e-300)) < log_a
    if not np.all(accept):
        final = [proposals[c] if accept[c] else fields[c] for c in range(solver.nchains)]
        solver.set_fields(final, seeds)
        solver.prepare()
    return accept


def uniform_shift_proposals(solver, rng, width):
    """phi' = phi + delta on every slice and site, one delta ~ U(-width/2, width/2)
    per chain and active channel (symmetric)."""
    fields, _ = solver.get_fields()
    out = []
    for c, f in enumerate(fields):
        p = solver.params(c)
        g = f.copy()
        for nf, active in ((0, abs(p.u1) > 1e-10), (1, abs(p.u2) > 1e-10)):
            if active:
                g[:, :, nf] += width * (rng.random() - 0.5)
        out.append(g)
    return out


def _accept_device(solver, out, st0, rng):
    """Exact acceptance of a device move from its per-chain sums (the out
    layout of bafq_global_propose / bafq_hmc_propose), then bafq_global_finish."""
    import ctypes as C
    st1 = solver.stats()
    p = solver.params(0)
    if p.pairing:
        dtau = p.beta / p.ltrot
        a2 = math.sqrt(-2.0 * p.u2 * dtau) if p.u2 < -1e-10 else 0.0
        lw0 = -0.5 * out[:, 1] - a2 * out[:, 3] + st0[:, LOGDET]
        lw1 = -0.5 * out[:, 2] - a2 * out[:, 4] + st1[:, LOGDET]
    else:
        lw0 = -0.5 * out[:, 1] + 2.0 * st0[:, LOGDET]
        lw1 = -0.5 * out[:, 2] + 2.0 * st1[:, LOGDET]
    u = rng.random(solver.nchains)
    log_a = lw1 - lw0 + out[:, 0]
    accept = np.log(np.maximum(u, 1e-300)) < log_a
    flags = np.ascontiguousarray(accept.astype(np.int32))
    solver._check(solver.lib.bafq_global_finish(solver.h, flags.ctypes.data_as(C.POINTER(C.c_int32))),
                  "bafq_global_finish")
    return accept, log_a


def metropolis_global_device(solver, net, lam, sigma, rng, seed=None):
    """One exact global Metropolis-Hastings step for every chain with the
    proposal of bafqmc_amd.proposal.LearnedProposal (phi' = phi + lam f(phi) +
    sigma xi, sigma per channel, already scaled), drawn, evaluated and undone
    on the device (bafq_global_propose / bafq_global_finish): no fields cross
    to the host.  net: a proposal.ConvNet or None (no drift).  Returns the
    boolean acceptance per chain; G is current afterwards."""
    import ctypes as C
    lib = solver.lib
    st0 = solver.stats()
    out = np.zeros((solver.nchains, 5))
    sig = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, dtype=np.float64), (2,)))
    if seed is None:
        seed = int(rng.integers(0, 2 ** 63))
    solver._check(lib.bafq_global_propose(solver.h, net.h if net is not None else None, float(lam),
                                          sig.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(seed),
                                          out.ctypes.data_as(C.POINTER(C.c_double))), "bafq_global_propose")
    return _accept_device(solver, out, st0, rng)[0]


def gaussian_centers(params):
    """Centre of the Gaussian part of log W per channel: the pairing weight's
    normal-ordering factor exp(-alpha_2 sum phi_2) shifts the U2 channel's
    Gaussian to -alpha_2 (src/pairing/src/operator_Hubbard.f90:100-137); every
    other centre is 0."""
    if params.pairing and params.u2 < -1e-10:
        return np.array([0.0, -math.sqrt(-2.0 * params.u2 * params.beta / params.ltrot)])
    return np.zeros(2)


def hmc_global_device(solver, net, angle, steps, rng, seed=None, return_log_a=False, force_scale=1.0):
    """One exact global Metropolis-Hastings step for every chain with the split
    Hamiltonian move of bafq_hmc_propose: momenta p ~ N(0, 1), `steps` steps of
    a half kick by the network's force (force_scale per channel times its
    output), a rotation of (phi - c, p) by `angle` and a half kick, then the
    exact acceptance with the true weight and the kinetic energies.  Returns
    the boolean acceptance per chain."""
    import ctypes as C
    lib = solver.lib
    st0 = solver.stats()
    out = np.zeros((solver.nchains, 5))
    cen = np.ascontiguousarray(gaussian_centers(solver.params(0)), dtype=np.float64)
    fsc = np.ascontiguousarray(np.broadcast_to(np.asarray(force_scale, dtype=np.float64), (2,)))
    if seed is None:
        seed = int(rng.integers(0, 2 ** 63))
    solver._check(lib.bafq_hmc_propose(solver.h, net.h if net is not None else None, float(angle), int(steps),
                                       cen.ctypes.data_as(C.POINTER(C.c_double)),
                                       fsc.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint64(seed),
                                       out.ctypes.data_as(C.POINTER(C.c_double))), "bafq_hmc_propose")
    accept, log_a = _accept_device(solver, out, st0, rng)
    return (accept, log_a) if return_log_a else accept


def grad_log_weights(solver):
    """d log W / d phi of every chain's current fields (bafq_grad_log_weight),
    one array [Ltrot, Ns, 2] per chain."""
    import ctypes as C
    out = np.zeros(int(solver.field_offsets[-1]))
    solver._check(solver.lib.bafq_grad_log_weight(solver.h, out.ctypes.data_as(C.POINTER(C.c_double))),
                  "bafq_grad_log_weight")
    res = []
    for c in range(solver.nchains):
        p = solver.params(c)
        res.append(out[solver.field_offsets[c]:solver.field_offsets[c + 1]].reshape(p.ltrot, p.nlx * p.nly, 2))
    return res
