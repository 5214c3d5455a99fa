"""Pinning of the finite-Fock ED oracle (oracle/ed.py).

* The reference's stored pairing ED results (benchmarks/paper/data/pairing/ed/*.json,
  collected in tests/golden/ed_3x3.json) were computed on the same finite space
  (nmax = 3 per mode, ncut = 4 in total, 7297 states): they must agree to rounding.
* The reference's number-conserving ED results sum particle-number shells up to
  8 or 10 bosons; the oracle's space stops at 4, so they agree to the weight of
  the shells beyond (relative 1e-7 at U1 = 1, mu = -3.5, beta = 4; 1e-4 at
  U2 = -0.5, mu = -7, beta = 1, where attraction favours doubly occupied sites).
* At U1 = U2 = 0 the model is Gaussian: the Bogoliubov closed form
  (oracle.bafq_oracle.gaussian_closed_form, restated from
  tests/physics/test_gaussian_solvers.py:44-77) holds up to the truncation
  weight of the dilute gas.
"""
from __future__ import annotations

import json
from pathlib import Path

import pytest

from oracle import bafq_oracle as O
from oracle.ed import FockED

FIXTURE = json.loads((Path(__file__).resolve().parent / "golden" / "ed_3x3.json").read_text())
KEYS = {"density_total": "density_total", "energy_density": "energy_density", "S_SF_K": "sf_K", "S_DW_K": "dw_K",
        "pair_equal": "pair_equal"}


@pytest.fixture(scope="module")
def ed():
    return FockED(3, 3, nmax=3, ncut=4)


def compare(ed, name, rtol):
    rec = FIXTURE[name]
    p = rec["params"]
    got = ed.thermal(p["u1"], p["u2"], p["mu"], p["delta"], p["beta"])
    for k, mine in KEYS.items():
        if k not in rec["observables"]:
            continue
        want = rec["observables"][k]
        assert abs(got[mine] - want) <= rtol * abs(want) + 1e-15, (name, k, got[mine], want)


def test_basis_is_the_references(ed):
    assert len(ed.codes) == FIXTURE["triangle_pairing_delta_3x3_D0p3_mum5_b4"]["cutoff"]["basis_size"] == 7297


@pytest.mark.parametrize("name", ["triangle_pairing_delta_3x3_D0_mum5_b4", "triangle_pairing_delta_3x3_D0p1_mum5_b4",
                                  "triangle_pairing_delta_3x3_D0p3_mum5_b4"])
def test_pairing_ed_matches_reference(ed, name):
    compare(ed, name, 1e-10)


@pytest.mark.parametrize("name,rtol", [("triangle_3x3_U1_1_mu-3.5_b4", 1e-6),
                                       ("triangle_3x3_U2_-0.5_mu-7_b1", 2e-4),
                                       ("triangle_3x3_U2_0_mu-7_b1", 1e-6)])
def test_number_conserving_ed_within_shell_tail(ed, name, rtol):
    compare(ed, name, rtol)


@pytest.mark.parametrize("delta,rtol", [(0.0, 1e-8), (0.3, 5e-3)])
def test_free_gas_closed_form(ed, delta, rtol):
    """beta = 4, mu = -5: the dilute gas (rho 3e-6 at Delta = 0) fits the 4-boson
    space to 6e-11; the pair field at Delta = 0.3 populates it (rho 3e-3) and
    the space's cutoff costs 0.2 %."""
    got = ed.thermal(0.0, 0.0, -5.0, delta, 4.0)
    want = O.gaussian_closed_form(delta, 4.0, -5.0)
    for k in ("density_total", "energy_density", "sf_K", "dw_K"):
        assert abs(got[k] - want[k]) <= rtol * abs(want[k]), (k, got[k], want[k])
    assert abs(got["pair_equal"] - want["pair_equal"]) <= rtol * abs(want["pair_equal"]) + 1e-15
