"""NumPy restatement of the reference BAFQMC sweep (the oracle).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP solver in
``bafqmc_amd``; it is imported only by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, never by the product path.

It restates, in the reference's own matrix conventions (no reduced forms),
both solvers of the reference:

* number conserving (NC): ``N = Ns`` complex Green matrix of the ``b`` flavour,
  the ``c`` flavour is its complex conjugate
  (src/number_conserving/src/obser_equal.f90:113-117);
* Nambu pairing: full ``N = 4 Ns`` Nambu basis ``(b, c, b+, c+)``
  (src/pairing/src/lattice.f90:26-35, src/pairing/src/non_interact.f90:33-93).

Each function cites the reference lines it follows.  Pinning (see
``tests/test_oracle_*.py``): RNG checkpoints (src/common/test_numerics.f90:45-56),
closed-form Gaussian/Bogoliubov limit (tests/physics/test_gaussian_solvers.py:44-77),
independent dense fixed-field products (tests/physics/test_fixed_field_solver.py:55-119),
brute-force determinant ratios, and golden outputs of the reference Fortran
binaries built from source by ``oracle/Makefile.ref``.  The time-displaced
pass (``Chain.tau_pass``, ``stab_green_big``) is pinned in
``tests/test_time_displaced.py``: with the reference's conventions it
reproduces the Fortran runs' Wrap_tau drift lines, with the bosonic ones the
closed form from explicit products (``dense_green_tau``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

ZERO = 1.0e-10          # src/number_conserving/src/calc_basic.f90:5
MODULUS = 2 ** 31
MULT = 48828125
NAUX = 2


# ---------------------------------------------------------------------------
# RNG (src/common/random.f90:3-11) and field initialisation
# (src/number_conserving/src/fields.f90:115-154, 280-288)
# ---------------------------------------------------------------------------
class Ranf:
    """seed <- 48828125*seed mod 2^31, u = seed/2^31 (src/common/random.f90:3-11)."""

    def __init__(self, seed: int):
        self.seed = int(seed)

    def __call__(self) -> float:
        self.seed = (MULT * self.seed) % MODULUS
        return self.seed / MODULUS

    def gaussian(self) -> float:
        """Box-Muller, fields.f90:280-288."""
        x1 = self()
        x2 = self()
        return math.sqrt(-2.0 * math.log(x1)) * math.cos(2.0 * math.pi * x2)


def conf_set(rng: Ranf, ini_type: int, ini_ampl: float, ini_bias, nsite: int, ltrot: int):
    """Initial fields phi[nt, site, ns] (Fortran phi(ns, ii, nt)); fields.f90:115-154."""
    phi = np.zeros((ltrot, nsite, NAUX))
    for nt in range(ltrot):
        for ii in range(nsite):
            for ns in range(NAUX):
                if ini_type == 1:
                    phi[nt, ii, ns] = ini_bias[ns] + ini_ampl * (rng() - 0.5)
                elif ini_type == 2:
                    phi[nt, ii, ns] = ini_bias[ns] + ini_ampl * rng.gaussian()
                elif ini_type == 3:
                    phi[nt, ii, ns] = ini_bias[ns] + ini_ampl * abs(rng.gaussian())
                else:
                    raise ValueError(f"incorrect initype input {ini_type}")
    return phi


def conf_transfer(phi_therm, nlx_t, nly_t, nlx, nly, ltrot):
    """Tile a thermalisation-lattice field onto the full lattice; fields.f90:156-234."""
    lt_t = phi_therm.shape[0]
    phi = np.zeros((ltrot, nlx * nly, NAUX))
    for nt in range(lt_t):
        for ny in range(nly):
            for nx in range(nlx):
                sx = nx if nx < nlx_t else nx - nlx_t
                sy = ny if ny < nly_t else ny - nly_t
                if sx < nlx_t and sy < nly_t:
                    phi[nt, ny * nlx + nx] = phi_therm[nt, sy * nlx_t + sx]
    for nt in range(lt_t, ltrot):
        phi[nt] = phi[nt - lt_t]
    return phi


# ---------------------------------------------------------------------------
# Lattice (src/number_conserving/src/lattice.f90:21-125)
# ---------------------------------------------------------------------------
@dataclass
class Lattice:
    nlx: int
    nly: int
    lq: int = 0
    imj: np.ndarray = field(default=None, repr=False)
    bonds: np.ndarray = field(default=None, repr=False)
    xk: np.ndarray = field(default=None, repr=False)
    aimj: np.ndarray = field(default=None, repr=False)
    k_dot_r: np.ndarray = field(default=None, repr=False)
    index_k: int = -1

    def __post_init__(self):
        lx, ly = self.nlx, self.nly
        self.lq = lq = lx * ly
        cell = [(n % lx, n // lx) for n in range(lq)]          # x fastest (:40-48)
        inv = lambda x, y: (y % ly) * lx + (x % lx)
        self.imj = np.array([[inv(cell[i][0] - cell[j][0], cell[i][1] - cell[j][1])
                              for j in range(lq)] for i in range(lq)])   # (:63-74)
        self.bonds = np.array([[inv(x + 1, y), inv(x, y + 1), inv(x - 1, y + 1)]
                               for (x, y) in cell])              # (:95-106)
        a1 = np.array([1.0, 0.0]); a2 = np.array([0.5, math.sqrt(3.0) / 2.0])
        b1 = np.array([2.0 * math.pi, -2.0 * math.pi / math.sqrt(3.0)])
        b2 = np.array([0.0, 4.0 * math.pi / math.sqrt(3.0)])
        self.aimj = np.array([x * a1 + y * a2 for (x, y) in cell])
        self.xk = np.array([x * b1 / lx + y * b2 / ly for (x, y) in cell])
        self.k_dot_r = self.xk @ self.aimj.T                      # (:87-91)
        if lx % 3 == 0 and ly % 3 == 0:                           # obser_equal.f90:123-125
            self.index_k = inv(2 * lx // 3, ly // 3)

    def hopping(self, t: float = 1.0):
        """Bond-multiplicity hopping matrix (non_interact.f90:33-71)."""
        h = np.zeros((self.lq, self.lq))
        for i in range(self.lq):
            for j in self.bonds[i]:
                h[i, j] += t
                h[j, i] += t
        return h


# ---------------------------------------------------------------------------
# Model parameters (calc_basic.f90:48-106)
# ---------------------------------------------------------------------------
@dataclass
class Params:
    pairing: bool
    u1: float
    u2: float
    mu: float
    nlx: int
    nly: int
    ltrot: int
    beta: float
    nwrap: int
    shift_loc: float
    rdelta: float = 0.0
    t: float = 1.0

    @property
    def dtau(self):
        return self.beta / self.ltrot

    @property
    def nst(self):
        if self.ltrot % self.nwrap:
            raise ValueError("Ltrot is not a multiple of Nwrap")
        return self.ltrot // self.nwrap


def hs_alpha(u: float, dtau: float, channel: int) -> complex:
    """alpha = sqrt(-2 U dtau): real for U2<0, imaginary for U1>0.

    NC: operator_Hubbard.f90:36-43; pairing: operator_Hubbard.f90:62-88.
    """
    if channel == 1 and u < -ZERO:
        raise ValueError("U1 must be non-negative")
    if channel == 2 and u > ZERO:
        raise ValueError("U2 must be non-positive")
    if u < -ZERO:
        return complex(math.sqrt(-2.0 * u * dtau), 0.0)
    if u > ZERO:
        return complex(0.0, math.sqrt(2.0 * u * dtau))
    return 0j


# ---------------------------------------------------------------------------
# Kinetic propagators
# ---------------------------------------------------------------------------
def kinetic_nc(latt: Lattice, p: Params):
    """e^{-dtau H}, e^{+dtau H}, H = hopping - mu (non_interact.f90:33-109, ZHEEV)."""
    h = latt.hopping(p.t) - p.mu * np.eye(latt.lq)
    w, u = np.linalg.eigh(h)
    ep = (u * np.exp(-p.dtau * w)) @ u.conj().T
    em = (u * np.exp(p.dtau * w)) @ u.conj().T
    return ep.astype(complex), em.astype(complex), w


def hamt_pairing(latt: Lattice, p: Params):
    """4Ns Nambu commutator matrix in (b, c, b+, c+) order; pairing non_interact.f90:33-93."""
    ns = latt.lq
    n = 4 * ns
    a = np.zeros((n, n), complex)
    idx = lambda s, sec: sec * ns + s
    z = complex(p.t, 0.0)
    for i in range(ns):
        for j in latt.bonds[i]:
            i1, i2, i3, i4 = (idx(i, k) for k in range(4))
            j1, j2, j3, j4 = (idx(j, k) for k in range(4))
            a[i1, j1] += z; a[i2, j2] += z
            a[j3, i3] -= z; a[j4, i4] -= z
            a[j1, i1] += z.conjugate(); a[j2, i2] += z.conjugate()
            a[i3, j3] -= z.conjugate(); a[i4, j4] -= z.conjugate()
    for i in range(ns):
        i1, i2, i3, i4 = (idx(i, k) for k in range(4))
        a[i1, i4] += p.rdelta; a[i2, i3] += p.rdelta
        a[i3, i2] -= p.rdelta; a[i4, i1] -= p.rdelta
        a[i1, i1] -= p.mu; a[i2, i2] -= p.mu
        a[i3, i3] += p.mu; a[i4, i4] += p.mu
    return a


def kinetic_pairing(latt: Lattice, p: Params):
    """General (non-Hermitian) eigen-decomposition exponential; non_interact.f90:135-178."""
    a = hamt_pairing(latt, p)
    w, v = np.linalg.eig(a)
    vinv = np.linalg.inv(v)
    ep = (v * np.exp(-p.dtau * w)) @ vinv
    em = (v * np.exp(p.dtau * w)) @ vinv
    return ep, em, w


# ---------------------------------------------------------------------------
# Stabilisation (src/number_conserving/src/stabilization.f90:41-190)
# ---------------------------------------------------------------------------
def qdrp(mat):
    """M P = Q D V with V rows normalised by |R_ii| (QDRP_decompose, :41-67; ZGEQP3)."""
    q, r, perm = sla.qr(mat, pivoting=True)
    d = np.abs(np.diag(r))
    v = np.triu(r) / d[:, None]
    return q, d, v, perm


class Prop:
    """Propagator state + wrap list (process_matrix.f90:6-33, 36-46, 94-102)."""

    def __init__(self, n: int, nst: int):
        eye = np.eye(n, dtype=complex)
        self.ur, self.vr, self.dr = eye.copy(), eye.copy(), np.ones(n)
        self.ul, self.vl, self.dl = eye.copy(), eye.copy(), np.ones(n)
        self.g = eye.copy()
        self.xmax = 0.0
        self.xmean = 0.0
        z = np.zeros((nst + 1, n, n), complex)
        self.urlist, self.vrlist = z.copy(), z.copy()
        self.ullist, self.vllist = z.copy(), z.copy()
        self.drlist = np.zeros((nst + 1, n))
        self.dllist = np.zeros((nst + 1, n))
        self.logdet_g = 0.0


def stab_ur(pr: Prop):
    """stabilization.f90:69-93."""
    m = pr.ur * pr.dr[None, :]
    q, d, v, perm = qdrp(m)
    pr.vr = v @ pr.vr[perm, :]
    pr.ur = q
    pr.dr = d


def stab_ul(pr: Prop):
    """stabilization.f90:95-119."""
    m = pr.ul.conj().T * pr.dl[None, :]
    q, d, v, perm = qdrp(m)
    pr.vl = pr.vl[:, perm] @ v.conj().T
    pr.ul = q.conj().T
    pr.dl = d


def stab_green(pr: Prop, nt: int, ltrot: int):
    """G = (1 - UR DR VR VL DL UL)^{-1}; stabilization.f90:121-190.

    Also records log|det G| = -sum log D (|det V| = 1, UL, UR, Q unitary).
    """
    vrvl = pr.vr @ pr.vl
    inv_ulur = pr.ur.conj().T @ pr.ul.conj().T
    m = inv_ulur - pr.dr[:, None] * vrvl * pr.dl[None, :]
    half = ltrot // 2
    if nt > half:
        m = m.conj().T
    q, d, v, perm = qdrp(m)
    if nt < half + 1:
        tmp = (q.conj().T @ pr.ur.conj().T) / d[:, None]
        tmp = sla.solve_triangular(v, tmp, lower=False)
        ptmp = np.empty_like(tmp)
        ptmp[perm, :] = tmp                       # ZLAPMR backward
        g = pr.ul.conj().T @ ptmp
    else:
        tmp = (pr.ul.conj().T @ q) / d[None, :]
        # X V^H = tmp  ->  V X^H = tmp^H
        x = sla.solve_triangular(v, tmp.conj().T, lower=False).conj().T
        px = np.empty_like(x)
        px[:, perm] = x                           # ZLAPMT backward
        g = px @ pr.ur.conj().T
    pr.logdet_g = -float(np.sum(np.log(d)))
    return g


def compare_mat(a, b):
    """stabilization.f90:256-266."""
    return float(np.sum(np.abs(a - b)))


# ---------------------------------------------------------------------------
# Time-displaced Green's functions (stabilization.f90:192-254)
# ---------------------------------------------------------------------------
# Sign of the B(beta, tau) block of the 2N x 2N matrix.  The reference writes
# mat_big(nr, nr + Ndim) = +DUL(nr) (stabilization.f90:206-208), which inverts
# [[1, B(beta,tau)], [-B(tau,0), 1]] -- the fermionic (1 + B)^{-1} form; its own
# propagation of the time-displaced functions (dynamics.f90:44-57 from
# process_matrix.f90:78-85) is the bosonic one, G(0) = (1 - B)^{-1}.  With
# -DUL the inverse is [[G(0,0), G(0,tau)], [G(tau,0), G(tau,tau)]] of the
# bosonic chain.
REFERENCE_TAU_SIGN = 1.0
BOSONIC_TAU_SIGN = -1.0


def udv(a):
    """A = U diag(D) V by Householder QR without pivoting, D = |R_ii|, rows of
    R divided by D (src/common/mymats.f90:79-113, ncon = 0)."""
    q, r = np.linalg.qr(a)
    d = np.abs(np.diag(r))
    v = np.triu(r) / d[:, None]
    return q, d, v


def stab_green_big(pr: Prop, sign: float = BOSONIC_TAU_SIGN):
    """(G00, G0t, Gt0, Gtt) from the UDV factors of B(tau,0) = UR DR VR and
    B(beta,tau) = VL DL UL (stab_green_big, stabilization.f90:192-254)."""
    n = pr.ur.shape[0]
    inv = np.linalg.inv
    big = np.zeros((2 * n, 2 * n), complex)
    big[:n, :n] = inv(pr.vr @ pr.vl)
    big[n:, n:] = inv(pr.ul @ pr.ur)
    big[np.arange(n), n + np.arange(n)] = sign * pr.dl
    big[n + np.arange(n), np.arange(n)] = -pr.dr
    bu, bd, bv = udv(big)
    left = sla.block_diag(inv(pr.vr), inv(pr.ul)) @ inv(bv) / bd[None, :]
    right = inv(bu) @ sla.block_diag(inv(pr.vl), inv(pr.ur))
    tot = left @ right
    return tot[:n, :n], tot[:n, n:], tot[n:, :n], tot[n:, n:]


# ---------------------------------------------------------------------------
# The chain: both solvers share the sweep orchestration
# (local_sweep.f90:55-161; identical in src/pairing/src/local_sweep.f90)
# ---------------------------------------------------------------------------
class Chain:
    """One Markov chain of the reference solver (one MPI rank)."""

    def __init__(self, p: Params, phi: np.ndarray, seed: int):
        self.p = p
        self.latt = Lattice(p.nlx, p.nly)
        self.ns = self.latt.lq
        self.nsec = 4 if p.pairing else 1
        self.n = self.ns * self.nsec
        if p.pairing:
            self.ep, self.em, self.spec = kinetic_pairing(self.latt, p)
        else:
            self.ep, self.em, self.spec = kinetic_nc(self.latt, p)
        self.alpha = [hs_alpha(p.u1, p.dtau, 1), hs_alpha(p.u2, p.dtau, 2)]
        self.active = [abs(p.u1) > ZERO, abs(p.u2) > ZERO]
        self.phi = np.array(phi, dtype=float).copy()
        assert self.phi.shape == (p.ltrot, self.ns, NAUX)
        self.rng = Ranf(seed)
        self.prop = Prop(self.n, p.nst)
        self.acc = np.zeros((2, 2))        # [channel, (proposed, accepted)] per bin
        self.acc_ratio_sum = np.zeros(2)   # sum over bins of per-bin ratios (Acc_calc_ratio)
        self.phase = np.zeros((2, 5))      # pairing phase counters per bin
        self.tau_xmax = np.zeros(3)        # Wrap_tau drifts of G0t, Gt0, Gtt (PropGreen Xmaxm, Xmeanm)
        self.tau_xmean = np.zeros(3)
        self.obs = None
        self.nobs = 0

    # -- HS factors ---------------------------------------------------------
    def _sector_signs(self, nf: int):
        """Per-sector exponent signs (pairing operator_Hubbard.f90:167-199)."""
        if not self.p.pairing:
            return (1,)
        if nf == 1:      # relative density U1: (+, -, -, +)
            return (1, -1, -1, 1)
        return (1, 1, -1, -1)   # total density U2

    def _rows(self, i):
        return [sec * self.ns + i for sec in range(self.nsec)]

    def scale_rows(self, mat, nf, nt, i, nflag, fld=None):
        """opU_mmult_R: row(s) of site i times exp(+-alpha phi); operator_Hubbard.f90:70-85.
        fld: the field channel of phi (default: the operator's own, nf)."""
        a = self.alpha[nf - 1] * self.phi[nt, i, (fld or nf) - 1] * nflag
        for r, s in zip(self._rows(i), self._sector_signs(nf)):
            mat[r, :] *= np.exp(s * a)

    def scale_cols(self, mat, nf, nt, i, nflag, fld=None):
        """opU_mmult_L: column(s) of site i; operator_Hubbard.f90:87-102."""
        a = self.alpha[nf - 1] * self.phi[nt, i, (fld or nf) - 1] * nflag
        for r, s in zip(self._rows(i), self._sector_signs(nf)):
            mat[:, r] *= np.exp(s * a)

    # -- local Metropolis ---------------------------------------------------
    def metro(self, nf: int, i: int, nt: int, therm: bool = False, shift=None):
        """LocalU_metro (NC localU.f90:31-77; pairing localU.f90:33-113)."""
        p = self.p
        g = self.prop.g
        shift = p.shift_loc if shift is None else shift
        phi_old = self.phi[nt, i, nf - 1]
        xflip = self.rng()
        phi_new = phi_old + (xflip - 0.5) * abs(shift)
        al = self.alpha[nf - 1]
        r_gauss = math.exp(-0.5 * phi_new * phi_new) / math.exp(-0.5 * phi_old * phi_old)
        if therm:
            u = self.rng()
            ok = abs(r_gauss) > u
            if ok:
                self.phi[nt, i, nf - 1] = phi_new
            return ok
        if not p.pairing:
            delta = np.exp(al * phi_new) / np.exp(al * phi_old) - 1.0
            r_det = 1.0 / (1.0 + delta * (1.0 - g[i, i]))
            ratio_abs = abs(r_gauss * r_det * np.conj(r_det))
            u = self.rng()
            ok = ratio_abs > u
            self.acc[nf - 1, 0] += 1
            if ok:
                self.acc[nf - 1, 1] += 1
                col = g[:, i].copy()
                row = -g[i, :].copy()
                row[i] += 1.0
                g -= (r_det * delta) * np.outer(col, row)
                self.phi[nt, i, nf - 1] = phi_new
            return ok
        # pairing: 4x4 P block, det^{-1/2}
        dphi = phi_new - phi_old
        ep = np.exp(al * dphi)
        em = np.exp(-al * dphi)
        signs = self._sector_signs(nf)
        dvec = np.array([(ep if s > 0 else em) - 1.0 for s in signs])
        r_const = 1.0
        if nf == 2 and self.active[1]:
            r_const = np.exp(al * phi_old) / np.exp(al * phi_new)
        isec = self._rows(i)
        gss = g[np.ix_(isec, isec)]
        pb = np.eye(4) + dvec[:, None] * (np.eye(4) - gss)
        pinv = np.linalg.inv(pb)
        det = np.linalg.det(pb)
        ratio_det = complex(det) ** (-0.5)
        ratio_total = r_gauss * r_const * ratio_det
        phase = math.atan2(ratio_total.imag, ratio_total.real)
        ratio_abs = abs(ratio_total)
        u = self.rng()
        ok = ratio_abs > u
        self.acc[nf - 1, 0] += 1
        ph = self.phase[nf - 1]
        ph[0] += 1; ph[1] += abs(phase)
        if abs(phase) > 0.9 * math.pi:
            ph[4] += 1
        if ok:
            self.acc[nf - 1, 1] += 1
            ph[2] += 1; ph[3] += abs(phase)
            left = g[:, isec].copy()
            mid = pinv * dvec[None, :]
            right = -g[isec, :].copy()
            for k, r in enumerate(isec):
                right[k, r] += 1.0
            g -= left @ mid @ right
            self.phi[nt, i, nf - 1] = phi_new
        return ok

    def prop_L(self, nf, nt):
        """LocalU_prop_L (NC localU.f90:79-95)."""
        for i in range(self.ns - 1, -1, -1):
            self.metro(nf, i, nt)
            self.scale_cols(self.prop.g, nf, nt, i, 1)
            self.scale_rows(self.prop.g, nf, nt, i, -1)
        for i in range(self.ns - 1, -1, -1):
            self.scale_cols(self.prop.ul, nf, nt, i, 1)

    def prop_R(self, nf, nt):
        """LocalU_prop_R (NC localU.f90:97-113)."""
        for i in range(self.ns):
            self.scale_rows(self.prop.g, nf, nt, i, 1)
            self.scale_cols(self.prop.g, nf, nt, i, -1)
            self.metro(nf, i, nt)
        for i in range(self.ns):
            self.scale_rows(self.prop.ur, nf, nt, i, 1)

    # -- wraps (stabilization.f90:268-366) -----------------------------------
    def wrap_pre(self, nt):
        pr, p = self.prop, self.p
        st = nt // p.nwrap
        if nt != 0:
            stab_ur(pr)
        pr.urlist[st], pr.vrlist[st], pr.drlist[st] = pr.ur, pr.vr, pr.dr
        if nt == p.ltrot:
            pr.g = stab_green(pr, nt, p.ltrot)

    def wrap_L(self, nt):
        pr, p = self.prop, self.p
        st = nt // p.nwrap
        pr.ur, pr.vr, pr.dr = pr.urlist[st].copy(), pr.vrlist[st].copy(), pr.drlist[st].copy()
        if nt == 0:
            pr.urlist[:] = 0; pr.vrlist[:] = 0; pr.drlist[:] = 0
        if nt != p.ltrot:
            stab_ul(pr)
            g = stab_green(pr, nt, p.ltrot)
            dif = compare_mat(g, pr.g)
            pr.xmax = max(pr.xmax, dif)
            pr.xmean += dif
            pr.g = g
        pr.ullist[st], pr.vllist[st], pr.dllist[st] = pr.ul, pr.vl, pr.dl

    def wrap_R(self, nt):
        pr, p = self.prop, self.p
        st = nt // p.nwrap
        pr.ul, pr.vl, pr.dl = pr.ullist[st].copy(), pr.vllist[st].copy(), pr.dllist[st].copy()
        if nt == p.ltrot:
            pr.ullist[:] = 0; pr.vllist[:] = 0; pr.dllist[:] = 0
        if nt != 0:
            stab_ur(pr)
            g = stab_green(pr, nt, p.ltrot)
            dif = compare_mat(g, pr.g)
            pr.xmax = max(pr.xmax, dif)
            pr.xmean += dif
            pr.g = g
        pr.urlist[st], pr.vrlist[st], pr.drlist[st] = pr.ur, pr.vr, pr.dr

    # -- sweep phases (local_sweep.f90:64-161) -------------------------------
    def warmup(self, shift_warm):
        """Local_sweep_therm: Gaussian-weight-only Metropolis (local_sweep.f90:64-78)."""
        for nt in range(self.p.ltrot):
            for nf in (1, 2):
                if self.active[nf - 1]:
                    for i in range(self.ns):
                        self.metro(nf, i, nt, therm=True, shift=self.p.shift_loc)

    def pre(self):
        """Local_sweep_pre (local_sweep.f90:80-94): UR = B(beta,0) and G(beta)."""
        pr, p = self.prop, self.p
        self.wrap_pre(0)
        for nt in range(1, p.ltrot + 1):
            for nf in (1, 2):
                if self.active[nf - 1]:
                    for i in range(self.ns):
                        self.scale_rows(pr.ur, nf, nt - 1, i, 1)
            pr.ur = self.ep @ pr.ur
            if nt % p.nwrap == 0:
                self.wrap_pre(nt)

    def prop_T_L(self):
        pr = self.prop
        pr.g = self.em @ (pr.g @ self.ep)
        pr.ul = pr.ul @ self.ep

    def prop_T_R(self):
        pr = self.prop
        pr.g = (self.ep @ pr.g) @ self.em
        pr.ur = self.ep @ pr.ur

    def sweep_L(self):
        p = self.p
        for nt in range(p.ltrot, 0, -1):
            if nt % p.nwrap == 0:
                self.wrap_L(nt)
            self.measure()
            self.prop_T_L()
            if self.active[1]:
                self.prop_L(2, nt - 1)
            if self.active[0]:
                self.prop_L(1, nt - 1)
        self.wrap_L(0)

    def sweep_R(self):
        p = self.p
        self.wrap_R(0)
        for nt in range(1, p.ltrot + 1):
            if self.active[0]:
                self.prop_R(1, nt - 1)
            if self.active[1]:
                self.prop_R(2, nt - 1)
            self.prop_T_R()
            if nt % p.nwrap == 0:
                self.wrap_R(nt)
            self.measure()

    # -- time-displaced path (dynamics.f90:38-57, stabilization.f90:368-404) ---
    def tau_pass(self, sign: float = BOSONIC_TAU_SIGN, record=None, u2_field: int = 2):
        """Dyn_reset + Dyn_sweep_R with the fields frozen, run between sweep_L
        and sweep_R (local_sweep.f90:119-123): G(tau,0), G(0,tau), G(tau,tau)
        propagated slice by slice from G(0) (PropGr_reset, process_matrix.f90:78-85;
        propgrU_R, propgrT_R, multiply.f90:48-64,88-97) and recomputed at every
        stabilisation point (Wrap_tau) with the drift of each against its
        propagated value.  ``record``: a list that receives (G00, Gt0, G0t, Gtt)
        at tau = 0 .. Ltrot - 1, the points where Obs_tau_calc sees them.
        ``u2_field``: the field channel the U2 operator is propagated with; the
        reference passes nf = 1 for both operators (dynamics.f90:50-52), so its
        U2 factor carries the U1 channel's field, while its checkpoint lists
        (sweep_L) carry phi_2; ``REFERENCE_TAU_SIGN`` with ``u2_field = 1``
        reproduces the reference's pass, the defaults give the bosonic chain's
        time-displaced functions."""
        pr, p = self.prop, self.p
        n = self.n
        pd = Prop(n, p.nst)                       # Prop_d = Prop (Prop_assign)
        pd.ur, pd.vr, pd.dr = pr.ur.copy(), pr.vr.copy(), pr.dr.copy()
        pd.ul, pd.vl, pd.dl = pr.ul.copy(), pr.vl.copy(), pr.dl.copy()
        g00, gt0, gtt = pr.g.copy(), pr.g.copy(), pr.g.copy()
        g0t = pr.g - np.eye(n)
        for nt in range(1, p.ltrot + 1):
            if record is not None:
                record.append((g00.copy(), gt0.copy(), g0t.copy(), gtt.copy()))
            for nf in (1, 2):
                if not self.active[nf - 1]:
                    continue
                fld = u2_field if nf == 2 else 1
                for i in range(self.ns):
                    self.scale_rows(gt0, nf, nt - 1, i, 1, fld)
                    self.scale_cols(g0t, nf, nt - 1, i, -1, fld)
                    self.scale_rows(gtt, nf, nt - 1, i, 1, fld)
                    self.scale_cols(gtt, nf, nt - 1, i, -1, fld)
                for i in range(self.ns):
                    self.scale_rows(pd.ur, nf, nt - 1, i, 1, fld)
            gt0 = self.ep @ gt0
            g0t = g0t @ self.em
            gtt = (self.ep @ gtt) @ self.em
            pd.ur = self.ep @ pd.ur
            if nt % p.nwrap == 0:
                st = nt // p.nwrap
                pd.ul, pd.vl, pd.dl = pr.ullist[st].copy(), pr.vllist[st].copy(), pr.dllist[st].copy()
                stab_ur(pd)
                n00, n0t, nt0, ntt = stab_green_big(pd, sign)
                for k, (new, old) in enumerate(((n0t, g0t), (nt0, gt0), (ntt, gtt))):
                    dif = compare_mat(new, old)
                    self.tau_xmax[k] = max(self.tau_xmax[k], dif)
                    self.tau_xmean[k] += dif
                g00, g0t, gt0, gtt = n00, n0t, nt0, ntt

    def bin(self, nsweep: int = 1, tau: bool = False, tau_sign: float = BOSONIC_TAU_SIGN, record=None,
            tau_u2_field: int = 2):
        """Local_sweep with is_beta=.true. (local_sweep.f90:136-161), then Obs ave;
        tau: the time-displaced pass before each sweep_R (toggle = .true.)."""
        self.obs = None
        self.nobs = 0
        self.acc[:] = 0
        self.phase[:] = 0
        for _ in range(nsweep):
            self.sweep_L()
            if tau:
                self.tau_pass(tau_sign, record, tau_u2_field)
            self.sweep_R()
        out = {k: v / self.nobs for k, v in self.obs.items()}
        for c in range(2):
            if self.active[c] and self.acc[c, 0] > 0:
                self.acc_ratio_sum[c] += self.acc[c, 1] / self.acc[c, 0]
        return out

    # -- measurement ---------------------------------------------------------
    def measure(self):
        sample = measure_pairing(self) if self.p.pairing else measure_nc(self)
        if self.obs is None:
            self.obs = {k: np.array(v, dtype=complex if np.iscomplexobj(v) else float) * 0
                        for k, v in sample.items()}
        for k, v in sample.items():
            self.obs[k] = self.obs[k] + v
        self.nobs += 1

    def log_weight(self):
        """-1/2 sum phi^2 + 2 log|det G| (pole_diagnostics.f90:130-131, fields.f90:28-33)."""
        s = np.linalg.svd(self.prop.g, compute_uv=False)
        return -0.5 * float(np.sum(self.phi ** 2)) + 2.0 * float(np.sum(np.log(s)))


def measure_nc(ch: Chain):
    """One equal-time sample; src/number_conserving/src/obser_equal.f90:95-198."""
    latt, p = ch.latt, ch.p
    lq = latt.lq
    g = ch.prop.g
    grup = g
    grupc = g.T - np.eye(lq)
    grdo = g.conj()
    grdoc = grdo.T - np.eye(lq)
    du = np.diag(grupc)
    dd = np.diag(grdoc)
    o = {}
    o["density_up"] = float(np.sum(du.real)) / lq
    o["density_do"] = float(np.sum(dd.real)) / lq
    o["density_total"] = float(np.sum((du + dd).real)) / lq
    o["density_site_total"] = (du + dd).real.copy()
    o["doubleOcc"] = float(np.sum((du * dd).real)) / lq
    o["squareOcc"] = float(np.sum((du * du + dd * dd).real)) / lq
    o["num_up"] = float(np.sum(du.real))
    o["num_do"] = float(np.sum(dd.real))
    n2u = float(np.sum((2.0 * du * du + du).real))
    n2d = float(np.sum((2.0 * dd * dd + dd).real))
    dbl = float(np.sum((du * dd).real)) / lq
    o["numsquare_up"] = float(np.sum((du[:, None] * du[None, :] + grupc * grup).real))
    o["numsquare_do"] = float(np.sum((dd[:, None] * dd[None, :] + grdoc * grdo).real))
    norm = 1.0 / (lq * lq)
    o["psf_Gamma"] = complex(np.sum(grupc * grdoc) * norm)
    if latt.index_k >= 0:
        ph = np.exp(1j * latt.k_dot_r[latt.index_k, latt.imj])
        o["sf_K"] = complex(np.sum(ph * (grupc + grdoc)) * norm)
        dw = (du[:, None] * du[None, :] + grupc * grup + dd[:, None] * dd[None, :] + grdoc * grdo
              + du[:, None] * dd[None, :] + dd[:, None] * du[None, :])
        o["dw_K"] = complex(np.sum(ph * dw) * norm)
    else:
        o["sf_K"] = 0j
        o["dw_K"] = 0j
    cu = du[:, None] * du[None, :] + grupc * grup
    cd = dd[:, None] * dd[None, :] + grdoc * grdo
    cud = du[:, None] * dd[None, :]
    # den_* files report the k=Gamma Fourier component of the imj-binned sums
    # (fourier_trans.f90:330-345 with FFT_RtoK at nk = inv_cell_list(1,1)).
    o["den_upup_sub11"] = complex(np.sum(cu)) / lq / lq
    o["den_dodo_sub11"] = complex(np.sum(cd)) / lq / lq
    o["den_updo"] = complex(np.sum(cud)) / lq / lq
    kin = 0.0
    for i in range(lq):
        for j in latt.bonds[i]:
            kin += p.t * (grupc[i, j] + grupc[j, i] + grdoc[i, j] + grdoc[j, i]).real / lq
    inter = (p.u1 + p.u2) * (n2u + n2d) / lq + 2.0 * (p.u2 - p.u1) * dbl
    o["kinetic"] = kin
    o["onsite_n2_up"] = n2u
    o["onsite_n2_do"] = n2d
    o["interaction_energy_density"] = inter
    o["energy_density"] = kin + inter
    return o


def measure_pairing(ch: Chain):
    """One equal-time sample; src/pairing/src/obser_equal.f90:112-271."""
    latt, p = ch.latt, ch.p
    ns = latt.lq
    g = ch.prop.g
    blk = lambda r, c: g[r * ns:(r + 1) * ns, c * ns:(c + 1) * ns]
    # sector order (b, c, b+, c+); columns carry (b+, c+, -b, -c) (:135-153)
    G = {}
    names = ("b", "c", "bdag", "cdag")
    colname = ("bdag", "cdag", "b", "c")
    for r in range(4):
        for c in range(4):
            sign = 1.0 if c < 2 else -1.0
            G[names[r], colname[c]] = sign * blk(r, c)
    gbb_d = G["b", "bdag"]; gb_c = G["b", "c"]; gb_cd = G["b", "cdag"]; gb_b = G["b", "b"]
    gc_b = G["c", "b"]; gc_bd = G["c", "bdag"]; gc_cd = G["c", "cdag"]; gc_c = G["c", "c"]
    gbd_b = G["bdag", "b"]; gbd_bd = G["bdag", "bdag"]; gbd_cd = G["bdag", "cdag"]; gbd_c = G["bdag", "c"]
    gcd_c = G["cdag", "c"]; gcd_cd = G["cdag", "cdag"]; gcd_bd = G["cdag", "bdag"]; gcd_b = G["cdag", "b"]
    nb = np.diag(gbd_b)
    nc = np.diag(gcd_c)
    Dbb1 = nb * nb + np.diag(gbd_b) * np.diag(gbb_d) + np.diag(gbd_bd) * np.diag(gb_b)
    Dcc1 = nc * nc + np.diag(gcd_c) * np.diag(gc_cd) + np.diag(gcd_cd) * np.diag(gc_c)
    Dbc1 = nb * nc + np.diag(gbd_cd) * np.diag(gb_c) + np.diag(gbd_c) * np.diag(gb_cd)
    o = {}
    o["density_up"] = float(np.sum(nb.real)) / ns
    o["density_do"] = float(np.sum(nc.real)) / ns
    o["density"] = float(np.sum((nb + nc).real)) / ns
    o["num_up"] = float(np.sum(nb.real))
    o["num_do"] = float(np.sum(nc.real))
    o["doubleOcc"] = float(np.sum(Dbc1.real)) / ns
    o["squareOcc"] = float(np.sum((nb * nb + nc * nc).real)) / ns
    o["local_numsquare"] = float(np.sum((Dbb1 + Dcc1).real)) / ns
    pair = float(np.sum((np.diag(gb_c) + np.diag(gbd_cd)).real)) / ns
    o["pair_equal"] = pair
    o["density_site_total"] = (nb + nc).real.copy()
    dens = float(np.sum((nb + nc).real)) / ns
    n2u = float(np.sum(Dbb1.real))
    n2d = float(np.sum(Dcc1.real))
    dbl = float(np.sum(Dbc1.real)) / ns
    Dbb = nb[:, None] * nb[None, :] + gbd_b * gbb_d + gbd_bd * gb_b
    Dcc = nc[:, None] * nc[None, :] + gcd_c * gc_cd + gcd_cd * gc_c
    Dbc = nb[:, None] * nc[None, :] + gbd_cd * gb_c + gbd_c * gb_cd
    Dcb = nc[:, None] * nb[None, :] + gcd_bd * gc_b + gcd_b * gc_bd
    o["numsquare_up"] = float(np.sum(Dbb.real))
    o["numsquare_do"] = float(np.sum(Dcc.real))
    norm = 1.0 / (ns * ns)
    o["den_upup_sub11"] = complex(np.sum(Dbb)) / ns / ns
    o["den_dodo_sub11"] = complex(np.sum(Dcc)) / ns / ns
    o["den_updo"] = complex(np.sum(Dbc)) / ns / ns
    if latt.index_k >= 0:
        ph = np.exp(1j * latt.k_dot_r[latt.index_k, latt.imj])
        o["sf_K"] = complex(np.sum(ph * (gbd_b + gcd_c)) * norm)
        o["dw_K"] = complex(np.sum(ph * (Dbb + Dcc + Dbc + Dcb)) * norm)
    else:
        o["sf_K"] = 0j
        o["dw_K"] = 0j
    o["psf_Gamma"] = complex(np.sum(np.diag(gbd_cd)[:, None] * np.diag(gc_b)[None, :]
                                    + gbd_c * gcd_b + gbd_b * gcd_c) * norm)
    kin = 0.0
    for i in range(ns):
        for j in latt.bonds[i]:
            kin += p.t * (gbd_b[i, j] + gbd_b[j, i] + gcd_c[i, j] + gcd_c[j, i]).real / ns
    inter = (p.u1 + p.u2) * (n2u + n2d) / ns + 2.0 * (p.u2 - p.u1) * dbl
    pe = p.rdelta * pair
    che = -p.mu * dens
    o["kinetic"] = kin
    o["onsite_n2_up"] = n2u
    o["onsite_n2_do"] = n2d
    o["density_total"] = dens
    o["interaction_energy_density"] = inter
    o["pairing_energy_density"] = pe
    o["chemical_energy_density"] = che
    o["energy_density"] = kin + inter + pe
    o["grand_energy_density"] = kin + inter + pe + che
    return o


# ---------------------------------------------------------------------------
# Independent dense references (no wrap / update code)
# ---------------------------------------------------------------------------
def slice_matrices(ch: Chain):
    """B_l = e^{-dtau K} diag(e^{V_l}) for l = 1..Ltrot (multiply.f90:7-16,66-70)."""
    out = []
    for nt in range(ch.p.ltrot):
        v = np.zeros(ch.n, complex)
        for nf in (1, 2):
            if not ch.active[nf - 1]:
                continue
            for i in range(ch.ns):
                for r, s in zip(ch._rows(i), ch._sector_signs(nf)):
                    v[r] += s * ch.alpha[nf - 1] * ch.phi[nt, i, nf - 1]
        out.append(ch.ep * np.exp(v)[None, :])
    return out


def dense_green(ch: Chain, tau: int):
    """G(tau) = (1 - B_tau ... B_1 B_L ... B_{tau+1})^{-1}, built from scratch."""
    bs = slice_matrices(ch)
    lt = ch.p.ltrot
    prod = np.eye(ch.n, dtype=complex)
    for idx in (*range(tau, lt), *range(tau)):
        prod = bs[idx] @ prod
    return np.linalg.inv(np.eye(ch.n) - prod), prod


def dense_green_tau(ch: Chain, tau: int):
    """(G(0,0), G(tau,0), G(0,tau), G(tau,tau)) of the bosonic chain from
    explicit products: G(0) = (1 - B(beta,0))^{-1}, G(tau,0) = B(tau,0) G(0),
    G(0,tau) = (G(0) - 1) B(tau,0)^{-1} = G(0) B(beta,tau) (no inverse of B(tau,0),
    whose pairing spectrum spans e^{+-tau |E|}), G(tau,tau) = (1 - B(tau,0) B(beta,tau))^{-1}."""
    bs = slice_matrices(ch)
    eye = np.eye(ch.n)
    b_t0 = eye.astype(complex)
    for idx in range(tau):
        b_t0 = bs[idx] @ b_t0
    b_bt = eye.astype(complex)
    for idx in range(tau, ch.p.ltrot):
        b_bt = bs[idx] @ b_bt
    g0 = np.linalg.inv(eye - b_bt @ b_t0)
    return g0, b_t0 @ g0, g0 @ b_bt, np.linalg.inv(eye - b_t0 @ b_bt)


def gaussian_closed_form(delta: float, beta: float, mu: float, length: int = 3):
    """Bogoliubov thermal sums at U1=U2=0 (tests/physics/test_gaussian_solvers.py:44-77)."""
    nsite = length ** 2
    occ, ano, eps = {}, {}, {}
    for x in range(length):
        for y in range(length):
            k1, k2 = 2 * math.pi * x / length, 2 * math.pi * y / length
            e = 2 * (math.cos(k1) + math.cos(k2) + math.cos(k1 - k2))
            xi = e - mu
            om = math.sqrt(xi * xi - delta * delta)
            coth = 1 / math.tanh(beta * om / 2)
            occ[x, y] = (xi / om * coth - 1) / 2
            ano[x, y] = -delta * coth / (2 * om)
            eps[x, y] = e
    kx, ky = (2 * length // 3) % length, (length // 3) % length
    pair = 2 * math.fsum(ano.values()) / nsite
    energy = 2 * math.fsum(eps[k] * occ[k] for k in occ) + nsite * delta * pair
    dsum = math.fsum(occ[x, y] * (1 + occ[(x + kx) % length, (y + ky) % length])
                     + ano[x, y] * ano[(x + kx) % length, (y + ky) % length] for x, y in occ)
    return {
        "density_total": 2 * math.fsum(occ.values()) / nsite,
        "energy_density": energy / nsite,
        "sf_K": 2 * occ[kx, ky] / nsite,
        "dw_K": 2 * dsum / nsite ** 2,
        "pair_equal": pair,
    }
