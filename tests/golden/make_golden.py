"""Write golden vectors for the sweep from the reference Fortran solvers.

Runs the binaries that ``oracle/Makefile.ref`` builds from the reference's own
sources (src/number_conserving/src, src/pairing/src) on fixed inputs and stores
inputs and outputs as ``tests/golden/<case>.npz`` (data only: parameters,
seeds, fields, per-bin output rows, final fields and seeds).

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import run_ref  # noqa: E402

OUT = Path(__file__).resolve().parent

NC_FILES = ["density_up", "density_do", "density_total", "density_site_total", "kinetic", "doubleOcc",
            "squareOcc", "num_up", "num_do", "numsquare_up", "numsquare_do", "onsite_n2_up", "onsite_n2_do",
            "interaction_energy_density", "energy_density", "sf_K", "psf_Gamma", "dw_K", "den_upup_sub11",
            "den_dodo_sub11", "den_updo", "log_weight", "pole_distance", "pole_x", "green_spectral_radius",
            "green_smax", "pole_z"]
PA_FILES = ["density_up", "density_do", "density", "density_total", "density_site_total", "num_up", "num_do",
            "kinetic", "doubleOcc", "squareOcc", "local_numsquare", "numsquare_up", "numsquare_do",
            "onsite_n2_up", "onsite_n2_do", "pair_equal", "interaction_energy_density", "pairing_energy_density",
            "chemical_energy_density", "grand_energy_density", "energy_density", "sf_K", "dw_K", "psf_Gamma",
            "den_upup_sub11", "den_dodo_sub11", "den_updo"]


def fixed_fields(lt, ns):
    """Nonuniform explicit fields in the pattern of tests/physics/test_fixed_field_solver.py:41-52."""
    t = np.arange(1, lt + 1, dtype=float)[:, None]
    s = np.arange(1, ns + 1, dtype=float)[None, :]
    dens = 0.23 + 0.31 * np.sin(0.71 * t + 0.37 * s) + 0.09 * np.cos(0.53 * t * s)
    rel = -0.19 + 0.28 * np.cos(0.43 * t - 0.61 * s) + 0.13 * np.sin(0.29 * t * s)
    return np.stack((rel, dens), axis=-1)


CASES = {
    "nc_fixed_3x3_w1": dict(p=dict(pairing=False, u1=0.7, u2=-0.1, mu=-5.0, nlx=3, nly=3, ltrot=6, beta=0.6,
                                   nwrap=1, nbin=3, shift_loc=0.0), seeds=[1357911], fixed=True),
    "nc_fixed_3x3_w2": dict(p=dict(pairing=False, u1=0.7, u2=-0.1, mu=-5.0, nlx=3, nly=3, ltrot=6, beta=0.6,
                                   nwrap=2, nbin=3, shift_loc=0.0), seeds=[1357911], fixed=True),
    "nc_fixed_3x3_w6": dict(p=dict(pairing=False, u1=0.7, u2=-0.1, mu=-5.0, nlx=3, nly=3, ltrot=6, beta=0.6,
                                   nwrap=6, nbin=3, shift_loc=0.0), seeds=[1357911], fixed=True),
    "pairing_fixed_3x3_w2": dict(p=dict(pairing=True, u1=0.7, u2=-0.1, mu=-5.0, rdelta=0.3, nlx=3, nly=3, ltrot=6,
                                        beta=0.6, nwrap=2, nbin=3, shift_loc=0.0), seeds=[1357911], fixed=True),
    # the 3x3 beta = 4 benchmark Hamiltonian, two chains generated from seeds.txt (one per MPI rank)
    "nc_traj_3x3_c1": dict(p=dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=3, nly=3, ltrot=40, beta=4.0,
                                  nwrap=10, nbin=5, shift_loc=1.5, ini_type=2, ini_ampl=1.0),
                           seeds=[13579113, 24681357]),
    "nc_traj_3x3_u2": dict(p=dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=3, nly=3, ltrot=20, beta=2.0,
                                  nwrap=5, nbin=4, shift_loc=1.5, ini_type=2, ini_ampl=1.0), seeds=[97531]),
    "pairing_traj_3x3": dict(p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=3, nly=3, ltrot=20,
                                    beta=2.0, nwrap=10, nbin=4, shift_loc=1.0, ini_type=2, ini_ampl=1.0),
                             seeds=[11223344, 55667788]),
    # a pairing 6x6 chain, beta=8, Ltrot=80 (one chain, one bin)
    "pairing_traj_6x6_c2": dict(p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=6, nly=6, ltrot=80,
                                       beta=8.0, nwrap=10, nbin=1, shift_loc=1.0, ini_type=2, ini_ampl=1.0),
                                seeds=[13579113]),
    # Gaussian warm-up sweeps before the local sweeps (Local_sweep_therm), both solvers
    "nc_warm_3x3": dict(p=dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=3, nly=3, ltrot=20, beta=2.0,
                               nwrap=5, nbin=2, shift_loc=1.5, ini_type=2, ini_ampl=1.0, nwarm=3),
                        seeds=[86420, 13579]),
    "pairing_warm_3x3": dict(p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=3, nly=3, ltrot=20,
                                    beta=2.0, nwrap=10, nbin=2, shift_loc=1.0, ini_type=2, ini_ampl=1.0, nwarm=3),
                             seeds=[424242]),
    # the learned workload's lattice: NC 9x9, beta=8, Ltrot=80 (one chain, one bin)
    "nc_traj_9x9_c4": dict(p=dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=9, nly=9, ltrot=80, beta=8.0,
                                  nwrap=10, nbin=1, shift_loc=1.5, ini_type=2, ini_ampl=1.0), seeds=[24681357]),
    # thermalisation lattice smaller than the lattice: fields generated on 3x3 x 20
    # slices and tiled onto 6x6 x 40 by conf_transfer (fields.f90:156-234)
    "nc_transfer_6x6": dict(p=dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=6, nly=6, ltrot=40, beta=4.0,
                                   nlx_therm=3, nly_therm=3, ltrot_therm=20, nwrap=10, nbin=2, shift_loc=1.5,
                                   ini_type=2, ini_ampl=1.0), seeds=[424243, 8642]),
    "pairing_transfer_6x6": dict(p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=6, nly=6, ltrot=40,
                                        beta=4.0, nlx_therm=3, nly_therm=3, ltrot_therm=20, nwrap=10, nbin=2,
                                        shift_loc=1.0, ini_type=2, ini_ampl=1.0), seeds=[97531]),
    # 12x12 lattices on the device-memory path: NC (N = 144) and the 12x12 pairing
    # chain (reference N = 576, reduced N = 288), one chain, one bin
    "nc_traj_12x12": dict(p=dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=12, nly=12, ltrot=80, beta=8.0,
                                 nwrap=10, nbin=1, shift_loc=1.5, ini_type=2, ini_ampl=1.0), seeds=[13579113]),
    "pairing_traj_12x12_c5": dict(p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.3, nlx=12, nly=12,
                                         ltrot=100, beta=10.0, nwrap=10, nbin=1, shift_loc=1.0, ini_type=2,
                                         ini_ampl=1.0), seeds=[24681357]),
    # pairing solver with the U1 channel alone (U2 = 0): the inputs of the
    # reference's 3x3 paper points at Delta = 0 and 0.3 (dtau = 0.01, Ltrot = 400,
    # small initial fields, 500 Gaussian warm-up sweeps), two chains, three bins
    "pairing_traj_3x3_d0": dict(p=dict(pairing=True, u1=1.0, u2=0.0, mu=-5.0, rdelta=0.0, nlx=3, nly=3, ltrot=400,
                                       beta=4.0, nwrap=10, nbin=3, shift_loc=1.0, ini_type=2, ini_ampl=0.1,
                                       nwarm=500), seeds=[31415926, 27182818]),
    "pairing_traj_3x3_d03": dict(p=dict(pairing=True, u1=1.0, u2=0.0, mu=-5.0, rdelta=0.3, nlx=3, nly=3, ltrot=400,
                                        beta=4.0, nwrap=10, nbin=3, shift_loc=1.0, ini_type=2, ini_ampl=0.1,
                                        nwarm=500), seeds=[16180339]),
    # pairing with both channels at U2 < 0 has no finite trace (docs/algorithm.md:152-156):
    # the chain runs to a pole of (1 - B)^-1 within a few dozen sweeps, 3x3, beta = 1
    "pairing_blowup_3x3": dict(p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=3, nly=3, ltrot=10,
                                      beta=1.0, nwrap=10, nbin=20, shift_loc=1.0, ini_type=2, ini_ampl=1.0),
                               seeds=[13579113]),
    # the time-displaced pass (is_tau) between sweep_L and sweep_R of every bin
    # after Nthermal; its only output is info.txt's six Wrap_tau drift lines
    "nc_tau_3x3": dict(p=dict(pairing=False, u1=1.0, u2=-0.5, mu=-5.0, nlx=3, nly=3, ltrot=20, beta=2.0,
                              nwrap=5, nbin=3, shift_loc=1.5, ini_type=2, ini_ampl=1.0, is_tau=True, nthermal=1),
                       seeds=[97531, 13579]),
    "nc_tau_6x6_u1": dict(p=dict(pairing=False, u1=1.0, u2=0.0, mu=-5.0, nlx=6, nly=6, ltrot=40, beta=4.0,
                                 nwrap=10, nbin=2, shift_loc=1.5, ini_type=2, ini_ampl=1.0, is_tau=True,
                                 nthermal=0), seeds=[24681357]),
    "pairing_tau_3x3": dict(p=dict(pairing=True, u1=1.0, u2=-0.5, mu=-5.0, rdelta=0.5, nlx=3, nly=3, ltrot=20,
                                   beta=2.0, nwrap=5, nbin=2, shift_loc=1.0, ini_type=2, ini_ampl=1.0,
                                   is_tau=True, nthermal=0), seeds=[11223344]),
}


def make(name, spec):
    p = dict(spec["p"])
    seeds = spec["seeds"]
    ns = p["nlx"] * p["nly"]
    fields = None
    if spec.get("fixed"):
        fields = [fixed_fields(p["ltrot"], ns) for _ in seeds]
    d, wall = run_ref.run(p, seeds, fields=fields)
    files = PA_FILES if p["pairing"] else NC_FILES
    rec = {"params": np.array(json.dumps(p)), "seeds": np.array(seeds, dtype=np.int64),
           "wall_s": np.array(wall)}
    if fields is not None:
        rec["fields_in"] = np.stack(fields)
    for f in files:
        if (d / f).exists():
            rec["out_" + f] = run_ref.read_rows(d / f)
    s_out, f_out = run_ref.read_confout(d / "confout.txt", len(seeds), p["ltrot"] * ns * 2)
    rec["seeds_out"] = np.array(s_out, dtype=np.int64)
    rec["fields_out"] = np.stack(f_out).reshape(len(seeds), p["ltrot"], ns, 2)
    info = (d / "info.txt").read_text()
    rec["info"] = np.array(info)
    np.savez_compressed(OUT / f"{name}.npz", **rec)
    run_ref.cleanup(d)
    print(f"{name}: {wall:.2f} s")


if __name__ == "__main__":
    only = sys.argv[1:]
    for name, spec in CASES.items():
        if not only or name in only:
            make(name, spec)
