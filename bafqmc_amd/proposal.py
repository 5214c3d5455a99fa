"""Learned global proposals: a periodic 3D convolutional network over the HS
field, trained on the GPU, proposing whole-field moves that are accepted by
exact Metropolis-Hastings (bafqmc_amd.global_moves).  No counterpart in the
reference.

Proposal (a Langevin-type move with a learned drift):
    phi' = phi + f(phi) + sigma xi,  xi ~ N(0, 1) per field,
    log q(phi'|phi) = -|phi' - phi - f(phi)|^2 / (2 sigma^2) + const,
so the acceptance min(1, W(phi')/W(phi) q(phi|phi')/q(phi'|phi)) needs f at
both phi and phi'.  The network f is trained by mean squared error to predict
the displacement of the fields over one plain sweep from the fields before it,
i.e. the mean drift of the local sampler; with f = 0 the move is the plain
Gaussian random walk it is compared with.  Whatever f is, the stationary
distribution is unchanged.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .native import NetDesc, load_library

c_float_p = C.POINTER(C.c_float)


def fptr(a):
    return a.ctypes.data_as(c_float_p)


class ConvNet:
    """Handle of the device network (include/bafqmc.h bafq_net_*)."""

    def __init__(self, lt, ly, lx, channels=2, hidden=8, layers=3, max_batch=64, seed=0, lib=None):
        self.lib = lib or load_library()
        self.desc = NetDesc(lt, ly, lx, channels, hidden, layers)
        self.shape = (channels, lt, ly, lx)
        self.max_batch = max_batch
        h = C.c_void_p()
        rc = self.lib.bafq_net_create(C.byref(self.desc), max_batch, C.byref(h))
        self.h = h
        if rc:
            msg = self.lib.bafq_net_last_error(h).decode()
            self.lib.bafq_net_destroy(h)
            self.h = None
            raise RuntimeError(msg)
        self.nparam = int(self.lib.bafq_net_param_count(self.h))
        self.set_params(self.he_init(seed))

    def he_init(self, seed):
        """He-normal weights (std sqrt(2 / fan_in)), zero biases; the last layer
        is scaled by 0.1 so an untrained drift starts small."""
        rng = np.random.default_rng(seed)
        c, hdim, L = self.desc.channels, self.desc.hidden, self.desc.layers
        out = []
        for l in range(L):
            cin = c if l == 0 else hdim
            cout = c if l == L - 1 else hdim
            w = rng.normal(size=cout * cin * 27) * np.sqrt(2.0 / (cin * 27))
            if l == L - 1:
                w *= 0.1
            out += [w, np.zeros(cout)]
        return np.concatenate(out).astype(np.float32)

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: {self.lib.bafq_net_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.bafq_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._check(self.lib.bafq_net_set_params(self.h, fptr(p)), "bafq_net_set_params")

    def get_params(self):
        p = np.zeros(self.nparam, np.float32)
        self._check(self.lib.bafq_net_get_params(self.h, fptr(p)), "bafq_net_get_params")
        return p

    def _batch(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape[1:] == self.shape and 1 <= x.shape[0] <= self.max_batch
        return x

    def forward(self, x):
        x = self._batch(x)
        y = np.zeros_like(x)
        self._check(self.lib.bafq_net_forward(self.h, x.shape[0], fptr(x), fptr(y)), "bafq_net_forward")
        return y

    def time_forward(self, batch, reps=5):
        """Device milliseconds of one forward pass over `batch` resident inputs."""
        ms = C.c_double()
        self._check(self.lib.bafq_net_time_forward(self.h, int(batch), int(reps), C.byref(ms)),
                    "bafq_net_time_forward")
        return ms.value

    def gradient(self, x, target):
        x = self._batch(x)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_double()
        g = np.zeros(self.nparam, np.float32)
        self._check(self.lib.bafq_net_gradient(self.h, x.shape[0], fptr(x), fptr(t), C.byref(loss), fptr(g)),
                    "bafq_net_gradient")
        return loss.value, g

    def train(self, x, target, steps, lr):
        x = self._batch(x)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.bafq_net_train(self.h, x.shape[0], fptr(x), fptr(t), int(steps), float(lr),
                                            C.byref(loss)), "bafq_net_train")
        return loss.value


def to_net(phis, params):
    """per-chain fields [Ltrot, Ns, 2] (site = x + Lx y) -> float32 [B, 2, Ltrot, Ly, Lx]."""
    a = np.stack([np.asarray(f) for f in phis])
    b = a.reshape(a.shape[0], params.ltrot, params.nly, params.nlx, 2)
    return np.ascontiguousarray(np.transpose(b, (0, 4, 1, 2, 3)), dtype=np.float32)


def from_net(y, params):
    b = np.transpose(np.asarray(y, dtype=np.float64), (0, 2, 3, 4, 1))
    return [b[i].reshape(params.ltrot, params.nly * params.nlx, 2) for i in range(b.shape[0])]


def active_channels(params):
    """Channels whose HS field couples (U1 and U2 nonzero); the others are left unmoved."""
    return np.array([abs(params.u1) > 1e-10, abs(params.u2) > 1e-10], dtype=np.float64)


def train_on_sweeps(net, x, t, epochs, lr, rng, holdout=0.25, batch=None, budget_s=None):
    """Fit the network to targets t of fields x (both in the network layout)
    by minibatch Adam (at most `epochs` epochs, or until `budget_s` seconds
    have passed); returns the per-channel residual variance on the held-out
    part and the held-out R^2 (variance explained)."""
    import time
    batch = min(batch or net.max_batch, net.max_batch)
    n = len(x)
    ntr = n - max(batch, int(n * holdout))
    t0 = time.perf_counter()
    for ep in range(epochs):
        idx = rng.permutation(ntr)
        for b in range(0, ntr - batch + 1, batch):
            j = idx[b:b + batch]
            net.train(x[j], t[j], 1, lr if ep < (3 * epochs) // 4 else lr / 3)
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
    res = []
    for i in range(ntr, n, batch):
        res.append(net.forward(x[i:i + batch]) - t[i:i + batch])
    r = np.concatenate(res).astype(np.float64)
    var_c = r.var(axis=(0, 2, 3, 4))
    r2 = 1.0 - float(np.mean(r * r)) / float(t[ntr:].astype(np.float64).var())
    return var_c, r2


class LearnedProposal:
    """phi' = phi + lam f(phi) + sqrt(lam) sigma xi, f the network (0 when net is
    None), sigma per channel.  With f the mean one-sweep displacement and
    sigma^2 its residual variance, lam = 1 imitates one local sweep and
    lam > 1 extrapolates it (a Langevin step of lam times the sweep's
    diffusion).  Channels whose field does not couple are not moved."""

    def __init__(self, net, params, sigma, lam=1.0):
        self.net, self.params = net, params
        self.lam = float(lam)
        self.mask = active_channels(params)
        self.sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (2,)) * np.sqrt(self.lam) * self.mask

    def drift(self, F):
        """lam f(F) * mask for stacked fields F [B, Ltrot, Ns, 2]."""
        if self.net is None:
            return np.zeros_like(F)
        out = np.empty_like(F)
        for i in range(0, len(F), self.net.max_batch):
            part = F[i:i + self.net.max_batch]
            out[i:i + len(part)] = np.stack(from_net(self.net.forward(to_net(part, self.params)), self.params))
        return self.lam * out * self.mask

    def propose(self, phis, rng):
        """(proposals, log q(phi|phi') - log q(phi'|phi)) for every chain; all
        chains at once as one array (the noise is drawn in FP32 and used exactly
        as drawn in both the move and its density)."""
        F = np.stack([np.asarray(f, dtype=np.float64) for f in phis])
        f0 = self.drift(F)
        xi = rng.standard_normal(size=F.shape, dtype=np.float32)
        P = F + f0 + self.sigma * xi
        f1 = self.drift(P)
        on = self.sigma > 0
        w = np.where(on, 1.0 / (2.0 * np.where(on, self.sigma, 1.0) ** 2), 0.0)
        lq = np.sum(w * ((P - F - f0) ** 2 - (F - P - f1) ** 2), axis=(1, 2, 3))
        return list(P), lq


def collect_forces(solver, params, sweeps, max_grad=50.0):
    """Training pairs of the Hamiltonian moves' network from the plain sampler:
    the fields phi of every chain after each of `sweeps` local sweeps and the
    residual force d log W / d phi + (phi - c) (the part of the gradient beyond
    the Gaussian centred at c, bafqmc_amd.global_moves.gaussian_centers) on the
    channels that couple.  Chains whose gradient exceeds max_grad (near a pole
    of (1 - B)^-1) are left out.  Returns (x, target) in the network layout."""
    from .global_moves import gaussian_centers, grad_log_weights
    cen = gaussian_centers(params)
    mask = active_channels(params)
    xs, ts = [], []
    for _ in range(sweeps):
        fields, _ = solver.get_fields()
        for f, g in zip(fields, grad_log_weights(solver)):
            if np.all(np.isfinite(g)) and np.max(np.abs(g)) < max_grad:
                xs.append(f)
                ts.append((g + (f - cen)) * mask)
        solver.sweep(1, measure=False)
    solver.read_bin()
    return to_net(xs, params), to_net(ts, params)
