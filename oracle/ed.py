"""Finite-Fock exact diagonalisation of the two-flavour boson model on small
triangular clusters (the statistical oracle of the sweep).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of sampled
observables, never by the product path.

Hamiltonian (the reference's conventions, tests/physics/test_ed_physics.py:33-90,
docs/algorithm.md:160-199):

    H = sum_ij h_ij (b+_i b_j + c+_i c_j)
        + sum_i [U1 (n_bi - n_ci)^2 + U2 (n_bi + n_ci)^2]
        + Delta sum_i (b_i c_i + c+_i b+_i)  -  mu N

with h the bond-multiplicity hopping matrix of the periodic triangular
cluster (forward bonds (x+1,y), (x,y+1), (x-1,y+1) and their conjugates,
src/number_conserving/src/lattice.f90:95-106, non_interact.f90:33-71).  The
basis holds every occupation of the 2 Ns modes with at most ``nmax`` bosons
per mode and at most ``ncut`` in total (the reference's paired cutoff
nmax=3, ncut=4 gives 7297 states on 3x3, docs/algorithm.md:197-199).
Q = N_b - N_c commutes with every term, so H is diagonalised block by block.

Observables follow src/number_conserving/src/obser_equal.f90:150-192 and
docs/algorithm.md:175-186: rho = <N>/Ns, energy_density = <H + mu N>/Ns,
S_SF(K) = Ns^-2 sum_ij e^{iK(r_i-r_j)} <b+_i b_j + c+_i c_j>,
S_DW(K) = Ns^-2 sum_ij e^{iK(r_i-r_j)} <n_i n_j>, pair_equal =
Ns^-1 sum_i <b+_i c+_i + b_i c_i>, K = (4 pi/3, 0).

Pinned by tests/test_ed_oracle.py: the reference's stored ED results on the
same finite space (benchmarks/paper/data/pairing/ed/*.json, collected in
tests/golden/ed_3x3.json), the free Bogoliubov closed form at U1 = U2 = 0
(tests/physics/test_gaussian_solvers.py:44-77), and the number-conserving
shell-converged ED results within their stated truncation tail.
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def triangular_cluster(lx: int, ly: int):
    """Hopping matrix (t = 1) and cos(K.(r_i - r_j)) of the periodic cluster."""
    ns = lx * ly
    inv = lambda x, y: (y % ly) * lx + (x % lx)
    h = np.zeros((ns, ns))
    for y in range(ly):
        for x in range(lx):
            i = inv(x, y)
            for j in (inv(x + 1, y), inv(x, y + 1), inv(x - 1, y + 1)):
                h[i, j] += 1.0
                h[j, i] += 1.0
    # K.(x a1 + y a2) with K = (4 pi / 3, 0), a1 = (1, 0), a2 = (1/2, sqrt 3 / 2)
    kr = np.array([4.0 * math.pi / 3.0 * x + 2.0 * math.pi / 3.0 * y for y in range(ly) for x in range(lx)])
    ck = np.cos(kr[:, None] - kr[None, :])
    return h, ck


def fock_basis(nmodes: int, nmax: int, ncut: int) -> np.ndarray:
    """All occupation vectors with n_m <= nmax and sum n <= ncut (rows)."""
    out = []

    def rec(prefix, left, m):
        if m == nmodes:
            out.append(prefix)
            return
        for k in range(min(nmax, left) + 1):
            rec(prefix + (k,), left - k, m + 1)

    rec((), ncut, 0)
    return np.array(out, dtype=np.int64)


class FockED:
    """Thermal averages of the 3x3 (or any small) cluster in a finite Fock space."""

    def __init__(self, lx: int, ly: int, nmax: int = 3, ncut: int = 4):
        self.lx, self.ly = lx, ly
        self.ns = ns = lx * ly
        self.h, self.ck = triangular_cluster(lx, ly)
        self.nmax, self.ncut = nmax, ncut
        self.occ = fock_basis(2 * ns, nmax, ncut)            # [state, mode]; modes b_0.., c_0..
        base = nmax + 1
        self.weights = base ** np.arange(2 * ns, dtype=np.int64)
        self.codes = self.occ @ self.weights
        order = np.argsort(self.codes)
        self.occ, self.codes = self.occ[order], self.codes[order]
        nb, nc = self.occ[:, :ns], self.occ[:, ns:]
        self.q = nb.sum(1) - nc.sum(1)
        # N_b and N_c are separately conserved when Delta = 0: smaller blocks
        self.nbnc = nb.sum(1) * (2 * ns * nmax + 1) + nc.sum(1)

    # -------------------------------------------------------------- operators
    def _lookup(self, codes):
        pos = np.searchsorted(self.codes, codes)
        pos = np.minimum(pos, len(self.codes) - 1)
        ok = self.codes[pos] == codes
        return pos, ok

    def _hop_terms(self, coef):
        """(rows, cols, values) of sum_ij coef_ij (b+_i b_j + c+_i c_j), i != j."""
        rows, cols, vals = [], [], []
        ns = self.ns
        for flavour in (0, 1):
            off = flavour * ns
            for i, j in itertools.product(range(ns), range(ns)):
                if i == j or coef[i, j] == 0.0:
                    continue
                mi, mj = off + i, off + j
                src = np.flatnonzero((self.occ[:, mj] > 0) & (self.occ[:, mi] < self.nmax))
                if src.size == 0:
                    continue
                dst_codes = self.codes[src] + self.weights[mi] - self.weights[mj]
                dst, ok = self._lookup(dst_codes)
                src, dst = src[ok], dst[ok]
                amp = coef[i, j] * np.sqrt(self.occ[src, mj] * (self.occ[src, mi] + 1.0))
                rows.append(dst); cols.append(src); vals.append(amp)
        return rows, cols, vals

    def _pair_terms(self, delta):
        """(rows, cols, values) of delta sum_i (b_i c_i + h.c.)."""
        rows, cols, vals = [], [], []
        ns = self.ns
        for i in range(ns):
            mb, mc = i, ns + i
            src = np.flatnonzero((self.occ[:, mb] > 0) & (self.occ[:, mc] > 0))
            dst, ok = self._lookup(self.codes[src] - self.weights[mb] - self.weights[mc])
            src, dst = src[ok], dst[ok]
            amp = delta * np.sqrt(self.occ[src, mb] * self.occ[src, mc] * 1.0)
            rows += [dst, src]; cols += [src, dst]; vals += [amp, amp]
        return rows, cols, vals

    def _dense_block(self, idx, terms, diag):
        pos = -np.ones(len(self.codes), dtype=np.int64)
        pos[idx] = np.arange(idx.size)
        m = np.zeros((idx.size, idx.size))
        for r, c, v in zip(*terms):
            sel = (pos[r] >= 0) & (pos[c] >= 0)
            np.add.at(m, (pos[r[sel]], pos[c[sel]]), v[sel])
        m[np.arange(idx.size), np.arange(idx.size)] += diag[idx]
        return m

    # -------------------------------------------------------------- averages
    def thermal(self, u1: float, u2: float, mu: float, delta: float, beta: float):
        """Grand-canonical averages of the benchmark observables."""
        ns = self.ns
        nb, nc = self.occ[:, :ns].astype(float), self.occ[:, ns:].astype(float)
        ntot = nb.sum(1) + nc.sum(1)
        n_site = nb + nc
        hdiag = (nb + nc) @ np.diag(self.h)
        inter = (u1 * (nb - nc) ** 2 + u2 * (nb + nc) ** 2).sum(1)
        diag = hdiag + inter - mu * ntot
        kin = self._hop_terms(self.h)
        pair = self._pair_terms(delta) if delta != 0.0 else ([], [], [])
        ham_terms = [kin[0] + pair[0], kin[1] + pair[1], kin[2] + pair[2]]
        sf = self._hop_terms(self.ck / ns ** 2)
        sf_diag = (nb + nc) @ np.diag(self.ck) / ns ** 2
        dw_diag = np.einsum("si,ij,sj->s", n_site, self.ck, n_site) / ns ** 2
        pr_unit = self._pair_terms(1.0 / ns)
        kin_diag = hdiag
        blocks = []
        sector = self.q if delta != 0.0 else self.nbnc
        for q in np.unique(sector):
            idx = np.flatnonzero(sector == q)
            hq = self._dense_block(idx, ham_terms, diag)
            e, v = np.linalg.eigh(hq)
            blocks.append((idx, e, v))
        e0 = min(b[1].min() for b in blocks)
        z = 0.0
        acc = dict.fromkeys(("n", "h", "kin", "inter", "sf", "dw", "pair"), 0.0)
        for idx, e, v in blocks:
            w = np.exp(-beta * (e - e0))
            z += w.sum()
            p2 = v * v                                         # |<s|n>|^2
            diag_avg = lambda d: float(np.sum(w * (d[idx] @ p2)))
            acc["n"] += diag_avg(ntot)
            acc["inter"] += diag_avg(inter)
            kin_op = self._dense_block(idx, kin, kin_diag)
            acc["kin"] += float(np.sum(w * np.einsum("sn,sn->n", v, kin_op @ v)))
            sf_op = self._dense_block(idx, sf, sf_diag)
            acc["sf"] += float(np.sum(w * np.einsum("sn,sn->n", v, sf_op @ v)))
            acc["dw"] += diag_avg(dw_diag)
            if delta != 0.0:
                pr_op = self._dense_block(idx, pr_unit, np.zeros(len(self.codes)))
                acc["pair"] += float(np.sum(w * np.einsum("sn,sn->n", v, pr_op @ v)))
        avg = {k: val / z for k, val in acc.items()}
        energy = avg["kin"] + avg["inter"] + delta * ns * avg["pair"]
        return {
            "density_total": avg["n"] / ns,
            "energy_density": energy / ns,
            "kinetic": avg["kin"] / ns,
            "interaction_energy_density": avg["inter"] / ns,
            "sf_K": avg["sf"],
            "dw_K": avg["dw"],
            "pair_equal": avg["pair"],
            "basis_size": int(len(self.codes)),
        }
