"""Batched solver: many independent Markov chains advanced by one kernel launch.

One ``BatchedSolver`` owns one C-ABI handle.  Each chain follows the reference
chain of one MPI rank exactly (same RNG stream, same proposal order, same
estimators); the batch dimension replaces the reference's MPI ranks
(src/number_conserving/src/main.f90:22-83).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .model import ModelSpec, NAUX
from .native import ModelDesc, NativeLibrary, Schedule, load_library, dptr, iptr

# observable slots of the per-chain accumulator block (csrc/sweep.hpp ObsIndex)
OBS_SLOTS = {
    "density_up": 0, "density_do": 1, "density_total": 2, "density": 3, "kinetic": 4,
    "doubleOcc": 5, "squareOcc": 6, "local_numsquare": 7, "pair_equal": 8, "num_up": 9,
    "num_do": 10, "numsquare_up": 11, "numsquare_do": 12, "onsite_n2_up": 13, "onsite_n2_do": 14,
    "interaction_energy_density": 15, "pairing_energy_density": 16, "chemical_energy_density": 17,
    "grand_energy_density": 18, "energy_density": 19,
}
COMPLEX_SLOTS = {"sf_K": 20, "psf_Gamma": 22, "dw_K": 24, "den_upup_sub11": 26,
                 "den_dodo_sub11": 28, "den_updo": 30}
O_NOBS = 32
O_SITE0 = 40
STATS = {"prop_u1": 0, "acc_u1": 1, "prop_u2": 2, "acc_u2": 3, "xmax": 4, "xmean_sum": 5,
         "nwrap": 6, "logdet": 7, "therm_prop_u1": 8, "therm_acc_u1": 9, "therm_prop_u2": 10,
         "therm_acc_u2": 11, "t_site": 12, "t_kinetic": 13, "t_measure": 14, "t_wrap": 15, "t_hs": 16,
         "t_wrap_products": 17, "t_wrap_stab_u": 18, "t_wrap_green": 19,
         # Wrap_tau drifts of G(0,tau), G(tau,0), G(tau,tau): running maxima, sums, count
         "tau_xmax_0t": 20, "tau_xmax_t0": 21, "tau_xmax_tt": 22, "tau_xsum_0t": 23, "tau_xsum_t0": 24,
         "tau_xsum_tt": 25, "tau_nwrap": 26}
TAU_BOSONIC, TAU_REFERENCE = 1, 2

NC_OUTPUTS = ["density_up", "density_do", "density_total", "density_site_total", "kinetic", "doubleOcc",
              "squareOcc", "num_up", "num_do", "numsquare_up", "numsquare_do", "onsite_n2_up",
              "onsite_n2_do", "interaction_energy_density", "energy_density", "sf_K", "psf_Gamma", "dw_K",
              "den_upup_sub11", "den_dodo_sub11", "den_updo"]
PAIRING_OUTPUTS = ["density_up", "density_do", "density", "density_total", "density_site_total", "num_up",
                   "num_do", "kinetic", "doubleOcc", "squareOcc", "local_numsquare", "numsquare_up",
                   "numsquare_do", "onsite_n2_up", "onsite_n2_do", "pair_equal", "interaction_energy_density",
                   "pairing_energy_density", "chemical_energy_density", "grand_energy_density",
                   "energy_density", "sf_K", "dw_K", "psf_Gamma", "den_upup_sub11", "den_dodo_sub11",
                   "den_updo"]


class SolverError(RuntimeError):
    pass


class BatchedSolver:
    """A batch of Markov chains (one per entry of chain_model) on one GPU.

    schedule: None (the library chooses), one (path, delayed) pair for every
    model, or a list of pairs per model -- see bafq_schedule in
    include/bafqmc.h; bafqmc_amd.autotune chooses one by timing.
    default_schedule applies when schedule is None."""

    default_schedule = None

    def __init__(self, models, chain_model, lib: NativeLibrary | None = None, schedule=None):
        self.lib = lib or load_library()
        self.models = [m if isinstance(m, ModelSpec) else ModelSpec(m) for m in models]
        self.chain_model = np.ascontiguousarray(chain_model, dtype=np.int32)
        self.nchains = len(self.chain_model)
        descs = (ModelDesc * len(self.models))()
        self._keep = []
        for k, m in enumerate(self.models):
            p = m.params
            descs[k].kind = 1 if p.pairing else 0
            descs[k].nlx, descs[k].nly = p.nlx, p.nly
            descs[k].ltrot, descs[k].nwrap = p.ltrot, p.nwrap
            descs[k].beta, descs[k].u1, descs[k].u2 = p.beta, p.u1, p.u2
            descs[k].mu, descs[k].rdelta, descs[k].shift_loc = p.mu, p.rdelta, p.shift_loc
            self._keep += [m.expk_p, m.expk_m]
            descs[k].expk_p = dptr(m.expk_p)
            descs[k].expk_m = dptr(m.expk_m)
        schedule = self.default_schedule if schedule is None else schedule
        sched = None
        if schedule is not None:
            if len(schedule) == 2 and all(isinstance(v, (int, np.integer)) for v in schedule):
                schedule = [tuple(schedule)] * len(self.models)
            sched = (Schedule * len(self.models))()
            for k, (path, delayed) in enumerate(schedule):
                sched[k].path, sched[k].delayed = int(path), int(delayed)
        h = C.c_void_p()
        rc = self.lib.bafq_create_scheduled(descs, len(self.models), iptr(self.chain_model), self.nchains,
                                            sched, C.byref(h))
        self.h = h
        if rc:
            msg = self.lib.bafq_last_error(h).decode() if h else "bafq_create failed"
            if h:
                self.lib.bafq_destroy(h)
            self.h = None
            raise SolverError(msg)
        n = C.c_int32(); os_ = C.c_int32(); ss = C.c_int32(); nf = C.c_int32()
        self.n = []
        self.nfields = []
        for c in range(self.nchains):
            self.lib.bafq_dims(self.h, c, C.byref(n), C.byref(os_), C.byref(ss), C.byref(nf))
            self.n.append(n.value)
            self.nfields.append(nf.value)
        self.obs_stride = os_.value
        self.stats_stride = ss.value
        self.field_offsets = np.concatenate([[0], np.cumsum(self.nfields)])

    def schedule(self):
        """(path, delayed) in effect per model."""
        out = []
        sc = Schedule()
        for k in range(len(self.models)):
            self._check(self.lib.bafq_get_schedule(self.h, k, C.byref(sc)), "bafq_get_schedule")
            out.append((sc.path, sc.delayed))
        return out

    def pairs(self):
        """Pairs of workgroups sharing a chain in the device-memory sweep kernel (0: none)."""
        return int(self.lib.bafq_pairs(self.h))

    # ------------------------------------------------------------------ utils
    def _check(self, rc, what):
        if rc:
            raise SolverError(f"{what}: {self.lib.bafq_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.bafq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, chain):
        return self.models[self.chain_model[chain]].params

    # ------------------------------------------------------------------ state
    def set_fields(self, phis, seeds):
        """phis: list (or array) of per-chain [Ltrot, Ns, 2] fields."""
        flat = np.concatenate([np.ascontiguousarray(p, dtype=np.float64).ravel() for p in phis])
        if flat.size != self.field_offsets[-1]:
            raise SolverError("field array size does not match the chains")
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        self._check(self.lib.bafq_set_fields(self.h, dptr(flat), iptr(s)), "bafq_set_fields")

    def get_fields(self):
        flat = np.zeros(int(self.field_offsets[-1]))
        seeds = np.zeros(self.nchains, np.int32)
        self._check(self.lib.bafq_get_fields(self.h, dptr(flat), iptr(seeds)), "bafq_get_fields")
        out = []
        for c in range(self.nchains):
            p = self.params(c)
            out.append(flat[self.field_offsets[c]:self.field_offsets[c + 1]].reshape(p.ltrot, p.ns, NAUX))
        return out, seeds.astype(np.int64)

    def warmup(self, n):
        self._check(self.lib.bafq_warmup(self.h, int(n)), "bafq_warmup")

    def prepare(self):
        self._check(self.lib.bafq_prepare(self.h), "bafq_prepare")

    def sweep(self, n=1, measure=True):
        self._check(self.lib.bafq_sweep(self.h, int(n), 1 if measure else 0), "bafq_sweep")

    def sweep_tau(self, n=1, measure=True, mode=TAU_BOSONIC, record=False):
        """Sweeps with the time-displaced pass before every sweep_R (Local_sweep
        with toggle = .true., local_sweep.f90:119-123)."""
        self._check(self.lib.bafq_sweep_tau(self.h, int(n), 1 if measure else 0, int(mode), 1 if record else 0),
                    "bafq_sweep_tau")

    def green_tau(self, chain):
        """[tau][G00, Gt0, G0t, Gtt] of the last recorded pass, Ltrot x 4 x N x N."""
        n, lt = self.n[chain], self.params(chain).ltrot
        g = np.zeros((lt, 4, n, n), np.complex128)
        self._check(self.lib.bafq_get_green_tau(self.h, int(chain), dptr(g.view(np.float64))), "bafq_get_green_tau")
        return g

    MEASURE_MODES = {"all": 0, "checkpoints": 1, "current": 2}

    def measure_pass(self, n=1, mode="all"):
        """Measure with frozen fields (bafq_measure_pass): n left passes measuring
        every slice ("all") or only the stabilisation points ("checkpoints"),
        or the current Green's function once ("current")."""
        self._check(self.lib.bafq_measure_pass(self.h, int(n), self.MEASURE_MODES[mode]), "bafq_measure_pass")

    def last_kernel_ms(self):
        v = C.c_double()
        self._check(self.lib.bafq_last_kernel_ms(self.h, C.byref(v)), "bafq_last_kernel_ms")
        return v.value

    def green(self, chain):
        n = self.n[chain]
        g = np.zeros((n, n), np.complex128)
        self._check(self.lib.bafq_get_green(self.h, int(chain), dptr(g.view(np.float64))), "bafq_get_green")
        return g

    def stats(self):
        """Per-chain counters without resetting them (slot 7: log|det G|)."""
        st = np.zeros((self.nchains, self.stats_stride))
        self._check(self.lib.bafq_get_stats(self.h, dptr(st)), "bafq_get_stats")
        return st

    def read_bin_raw(self):
        obs = np.zeros((self.nchains, self.obs_stride))
        st = np.zeros((self.nchains, self.stats_stride))
        self._check(self.lib.bafq_read_bin(self.h, dptr(obs), dptr(st)), "bafq_read_bin")
        return obs, st

    def read_bin(self):
        """Per-chain bin averages (Obs_equal%ave, obser_equal.f90:66-93) and counters."""
        obs, st = self.read_bin_raw()
        out = []
        for c in range(self.nchains):
            p = self.params(c)
            nobs = obs[c, O_NOBS]
            z = 1.0 / nobs if nobs > 0 else 0.0
            d = {k: obs[c, i] * z for k, i in OBS_SLOTS.items()}
            for k, i in COMPLEX_SLOTS.items():
                d[k] = complex(obs[c, i], obs[c, i + 1]) * z
            d["density_site_total"] = obs[c, O_SITE0:O_SITE0 + p.ns] * z
            d["_nobs"] = nobs
            d["_stats"] = {k: st[c, i] for k, i in STATS.items()}
            out.append(d)
        return out
