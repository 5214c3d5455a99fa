"""Host-side model setup: input files, lattice, fields and kinetic propagators.

Mirrors the reference's set-up code that runs once per chain on the CPU:
``read_input``/``Params_set`` (src/number_conserving/src/calc_basic.f90:48-106,
src/pairing/src/calc_basic.f90:48-108), ``Lattice_make``
(src/number_conserving/src/lattice.f90:21-125), ``conf_in``/``conf_set``/
``conf_transfer``/``conf_out`` (src/number_conserving/src/fields.f90:35-288) and
``opT_set`` (src/number_conserving/src/non_interact.f90:74-109,
src/pairing/src/non_interact.f90:121-178).  The kinetic propagators are built in
the solver's reduced basis (N = Ns for the number-conserving solver, N = 2 Ns
with rows b_1..b_Ns, c+_1..c+_Ns for the pairing solver).
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field, asdict
from pathlib import Path

import numpy as np

ZERO = 1.0e-10
NAUX = 2
MODULUS = 2 ** 31
MULT = 48828125


# ---------------------------------------------------------------- parameters
@dataclass
class Params:
    """The seven list-directed rows of ``paramC_sets.txt`` (calc_basic.f90:51-59)."""
    pairing: bool = False
    u1: float = 0.0
    u2: float = 0.0
    mu: float = 0.0
    rdelta: float = 0.0
    nlx: int = 3
    nly: int = 3
    ltrot: int = 40
    beta: float = 4.0
    nlx_therm: int = 3
    nly_therm: int = 3
    ltrot_therm: int = 40
    nwrap: int = 10
    nbin: int = 1
    nsweep: int = 1
    shift_loc: float = 1.0
    is_tau: bool = False
    nthermal: int = 0
    is_warm: bool = False
    nwarm: int = 0
    shift_warm1: float = 1.0
    shift_warm2: float = 1.0
    ini_type: int = 2
    ini_ampl: float = 1.0
    ini_bias1: float = 0.0
    ini_bias2: float = 0.0

    @property
    def dtau(self) -> float:
        return self.beta / self.ltrot

    @property
    def ns(self) -> int:
        return self.nlx * self.nly

    @property
    def n(self) -> int:
        return 2 * self.ns if self.pairing else self.ns

    @property
    def nst(self) -> int:
        """Stabilisation points per sweep direction (Params_set, calc_basic.f90:100-104)."""
        return self.ltrot // self.nwrap

    def validate(self):
        """Checks of read_input/Params_set (calc_basic.f90:61-62,100-104)."""
        if self.u1 < -ZERO:
            raise ValueError("U1 must be non-negative (relative-density channel)")
        if self.u2 > ZERO:
            raise ValueError("U2 must be non-positive (total-density channel)")
        if self.ltrot % self.nwrap:
            raise ValueError("Ltrot is not a multiple of Nwrap")
        if self.nlx_therm > self.nlx or self.nly_therm > self.nly or self.ltrot_therm > self.ltrot:
            raise ValueError("thermalisation lattice larger than the lattice")
        return self

    def to_text(self) -> str:
        first = f"{self.u1!r} {self.u2!r} {self.mu!r}" + (f" {self.rdelta!r}" if self.pairing else "")
        b = lambda v: ".true." if v else ".false."
        rows = [
            first,
            f"{self.nlx} {self.nly} {self.ltrot} {self.beta!r}",
            f"{self.nlx_therm} {self.nly_therm} {self.ltrot_therm}",
            f"{self.nwrap} {self.nbin} {self.nsweep} {self.shift_loc!r}",
            f"{b(self.is_tau)} {self.nthermal}",
            f"{b(self.is_warm)} {self.nwarm} {self.shift_warm1!r} {self.shift_warm2!r}",
            f"{self.ini_type} {self.ini_ampl!r} {self.ini_bias1!r} {self.ini_bias2!r}",
        ]
        legend = ["U1 U2 mu" + (" RDelta" if self.pairing else ""), "Nlx Nly Ltrot Beta",
                  "NlxTherm NlyTherm LtrotTherm", "Nwrap Nbin Nsweep shiftLoc", "is_tau Nthermal",
                  "is_warm Nwarm shiftWarm1 shiftWarm2", "iniType iniAmpl iniBias1 iniBias2"]
        return "\n".join(rows) + "\n\n" + "\n".join(legend) + "\n"


def _tokens(line: str):
    return [t for t in re.split(r"[\s,]+", line.strip()) if t]


def _num(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "E"))


def _logical(tok: str) -> bool:
    t = tok.strip().lower().lstrip(".")
    if t.startswith("t"):
        return True
    if t.startswith("f"):
        return False
    raise ValueError(f"not a Fortran logical: {tok!r}")


def read_params(path, pairing: bool) -> Params:
    """Parse ``paramC_sets.txt``: each Fortran READ consumes one record."""
    lines = [l for l in Path(path).read_text().splitlines()]
    recs = [_tokens(l) for l in lines]
    recs = [r for r in recs if r][:7]
    if len(recs) < 7:
        raise ValueError("paramC_sets.txt needs seven parameter rows")
    p = Params(pairing=pairing)
    r = recs[0]
    p.u1, p.u2, p.mu = _num(r[0]), _num(r[1]), _num(r[2])
    if pairing:
        p.rdelta = _num(r[3])
    r = recs[1]
    p.nlx, p.nly, p.ltrot, p.beta = int(r[0]), int(r[1]), int(r[2]), _num(r[3])
    r = recs[2]
    p.nlx_therm, p.nly_therm, p.ltrot_therm = int(r[0]), int(r[1]), int(r[2])
    r = recs[3]
    p.nwrap, p.nbin, p.nsweep, p.shift_loc = int(r[0]), int(r[1]), int(r[2]), _num(r[3])
    r = recs[4]
    p.is_tau, p.nthermal = _logical(r[0]), int(r[1])
    r = recs[5]
    p.is_warm, p.nwarm, p.shift_warm1, p.shift_warm2 = _logical(r[0]), int(r[1]), _num(r[2]), _num(r[3])
    r = recs[6]
    p.ini_type, p.ini_ampl, p.ini_bias1, p.ini_bias2 = int(r[0]), _num(r[1]), _num(r[2]), _num(r[3])
    return p.validate()


# ---------------------------------------------------------------- lattice
class Lattice:
    """Triangular lattice tables (lattice.f90:21-125): x fastest, forward bonds
    (x+1,y), (x,y+1), (x-1,y+1), momentum K=(4pi/3,0) at cell (2Lx/3, Ly/3)."""

    def __init__(self, nlx: int, nly: int):
        self.nlx, self.nly = nlx, nly
        self.ns = nlx * nly
        inv = lambda x, y: (y % nly) * nlx + (x % nlx)
        self.bonds = np.array([[inv(x + 1, y), inv(x, y + 1), inv(x - 1, y + 1)]
                               for y in range(nly) for x in range(nlx)], dtype=np.int64)

    def hopping(self, t: float = 1.0) -> np.ndarray:
        """Bond-multiplicity hopping matrix; non_interact.f90:33-71 with RT = 1."""
        h = np.zeros((self.ns, self.ns))
        for i in range(self.ns):
            for j in self.bonds[i]:
                h[i, j] += t
                h[j, i] += t
        return h


def kinetic_propagators(p: Params):
    """exp(-dtau A), exp(+dtau A) in the reduced basis, plus the spectrum of A.

    NC: A = H = hopping - mu (Hermitian, real symmetric here).
    Pairing: A_1 = [[h, Delta], [-Delta, -h]], h = hopping - mu, the {b, c+}
    block of the reference's 4Ns commutator matrix (src/pairing/src/non_interact.f90:33-93).
    In the eigenbasis of h each mode is a 2x2 block M = [[xi, D], [-D, -xi]] with
    M^2 = (xi^2 - D^2) I, so exp(-tau M) = cosh(w tau) I - sinh(w tau)/w M exactly.
    """
    latt = Lattice(p.nlx, p.nly)
    h = latt.hopping() - p.mu * np.eye(latt.ns)
    xi, w = np.linalg.eigh(h)
    tau = p.dtau
    if not p.pairing:
        ep = (w * np.exp(-tau * xi)) @ w.T
        em = (w * np.exp(tau * xi)) @ w.T
        return ep, em, xi

    d = float(p.rdelta)
    om2 = xi * xi - d * d

    def blocks(sgn):
        # exp(-sgn tau M) = C I - sgn S M, C = cosh(w tau), S = sinh(w tau)/w
        c = np.empty_like(xi)
        s = np.empty_like(xi)
        for k, o2 in enumerate(om2):
            if o2 > 0:
                om = math.sqrt(o2)
                c[k] = math.cosh(om * tau)
                s[k] = math.sinh(om * tau) / om
            elif o2 < 0:
                om = math.sqrt(-o2)
                c[k] = math.cos(om * tau)
                s[k] = math.sin(om * tau) / om
            else:
                c[k] = 1.0
                s[k] = tau
        bb = c - sgn * s * xi
        bc = -sgn * s * d
        cb = sgn * s * d
        cc = c + sgn * s * xi
        tr = lambda v: (w * v) @ w.T
        return np.block([[tr(bb), tr(bc)], [tr(cb), tr(cc)]])

    spec = np.concatenate([np.sqrt(np.abs(om2)) * np.sign(om2), -np.sqrt(np.abs(om2)) * np.sign(om2)])
    return blocks(+1.0), blocks(-1.0), spec


# ---------------------------------------------------------------- fields
class LcgBatch:
    """The reference RNG (src/common/random.f90:3-11) advanced for many chains at once."""

    def __init__(self, seeds):
        self.s = np.asarray(seeds, dtype=np.int64).copy()

    def ranf(self):
        self.s = (MULT * self.s) % MODULUS
        return self.s.astype(np.float64) / MODULUS

    def gaussian(self):
        """Box-Muller, fields.f90:280-288."""
        x1 = self.ranf()
        x2 = self.ranf()
        return np.sqrt(-2.0 * np.log(x1)) * np.cos(2.0 * math.pi * x2)


def conf_set_batch(p: Params, seeds):
    """Initial fields for many chains (fields.f90:115-154); returns (phi[c, nt, site, ns], seeds')."""
    rng = LcgBatch(seeds)
    nc = len(seeds)
    nst = p.nlx_therm * p.nly_therm
    phi = np.zeros((nc, p.ltrot_therm, nst, NAUX))
    bias = (p.ini_bias1, p.ini_bias2)
    for nt in range(p.ltrot_therm):
        for ii in range(nst):
            for ns in range(NAUX):
                if p.ini_type == 1:
                    x = rng.ranf()
                    phi[:, nt, ii, ns] = bias[ns] + p.ini_ampl * (x - 0.5)
                elif p.ini_type == 2:
                    phi[:, nt, ii, ns] = bias[ns] + p.ini_ampl * rng.gaussian()
                elif p.ini_type == 3:
                    phi[:, nt, ii, ns] = bias[ns] + p.ini_ampl * np.abs(rng.gaussian())
                else:
                    raise ValueError(f"incorrect initype input {p.ini_type}")
    return phi, rng.s.astype(np.int64)


def conf_transfer(p: Params, phi_therm):
    """Tile thermalisation-lattice fields onto the lattice (fields.f90:156-234)."""
    nc = phi_therm.shape[0]
    phi = np.zeros((nc, p.ltrot, p.ns, NAUX))
    lxt, lyt = p.nlx_therm, p.nly_therm
    for ny in range(p.nly):
        for nx in range(p.nlx):
            sx = nx if nx < lxt else nx - lxt
            sy = ny if ny < lyt else ny - lyt
            if sx < lxt and sy < lyt:
                phi[:, :p.ltrot_therm, ny * p.nlx + nx] = phi_therm[:, :, sy * lxt + sx]
    for nt in range(p.ltrot_therm, p.ltrot):
        phi[:, nt] = phi[:, nt - p.ltrot_therm]
    return phi


def read_fields(run_dir, p: Params, nchains: int):
    """conf_in (fields.f90:35-113): confin.txt seed 0 -> generate from seeds.txt,
    otherwise explicit per-chain fields.  Returns (phi[c, nt, site, ns], seeds)."""
    run_dir = Path(run_dir)
    toks = (run_dir / "confin.txt").read_text().split()
    first = int(float(toks[0].replace("d", "e").replace("D", "E")))
    nst = p.nlx_therm * p.nly_therm
    nvals = p.ltrot_therm * nst * NAUX
    if first == 0:
        seed_toks = (run_dir / "seeds.txt").read_text().split()
        if len(seed_toks) < nchains:
            raise ValueError(f"seeds.txt holds {len(seed_toks)} seeds for {nchains} chains")
        seeds = [int(float(t)) for t in seed_toks[:nchains]]
        phi_t, seeds_out = conf_set_batch(p, seeds)
    else:
        seeds_out = []
        blocks = []
        pos = 1
        for c in range(nchains):
            if c == 0:
                seeds_out.append(first)
            else:
                seeds_out.append(int(float(toks[pos])))
                pos += 1
            vals = np.array([_num(t) for t in toks[pos:pos + nvals]])
            if vals.size != nvals:
                raise ValueError("confin.txt holds fewer fields than the chains need")
            pos += nvals
            blocks.append(vals.reshape(p.ltrot_therm, nst, NAUX))
        phi_t = np.stack(blocks)
        seeds_out = np.array(seeds_out, dtype=np.int64)
    return conf_transfer(p, phi_t), np.asarray(seeds_out, dtype=np.int64)


def write_fields(run_dir, phi, seeds):
    """conf_out (fields.f90:236-278): per chain the seed then one field per line."""
    lines = []
    for c in range(len(seeds)):
        lines.append(f" {int(seeds[c])}\n")
        lines.extend(f" {v: .16E}\n" for v in np.asarray(phi[c]).ravel())
    (Path(run_dir) / "confout.txt").write_text("".join(lines))


# ---------------------------------------------------------------- model spec
@dataclass
class ModelSpec:
    """One Hamiltonian/discretisation for the C ABI (``bafq_model_desc``)."""
    params: Params
    expk_p: np.ndarray = field(default=None, repr=False)
    expk_m: np.ndarray = field(default=None, repr=False)
    spectrum: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        if self.expk_p is None:
            self.expk_p, self.expk_m, self.spectrum = kinetic_propagators(self.params)
        self.expk_p = np.ascontiguousarray(self.expk_p, dtype=np.float64)
        self.expk_m = np.ascontiguousarray(self.expk_m, dtype=np.float64)
