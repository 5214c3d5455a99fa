"""The C ABI: both libraries load and export every function include/bafqmc.h declares
(no device call is made here)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from bafqmc_amd import build

HEADER = Path(__file__).resolve().parents[1] / "include" / "bafqmc.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(bafq_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("bafq_create", "bafq_destroy", "bafq_set_fields", "bafq_get_fields", "bafq_warmup",
                     "bafq_prepare", "bafq_sweep", "bafq_read_bin", "bafq_get_green", "bafq_last_error"):
        assert required in names


@pytest.mark.parametrize("which", ["hip", "emu"])
def test_library_exports_every_declared_symbol(which):
    path = build.HIP_LIB if which == "hip" else build.EMU_LIB
    if not path.exists():
        (build.build_hip if which == "hip" else build.build_emu)()
    lib = ctypes.CDLL(str(path))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_rejects_invalid_models(emu_lib):
    """Input validation of read_input/Params_set (calc_basic.f90:61-62,100-104) at the ABI."""
    import numpy as np
    from bafqmc_amd.model import Params, ModelSpec
    from bafqmc_amd.solver import BatchedSolver, SolverError
    p = Params(nlx=3, nly=3, ltrot=6, beta=0.6, nwrap=2, u1=1.0, u2=-0.1, mu=-5.0)
    spec = ModelSpec(p)
    spec.params = Params(nlx=3, nly=3, ltrot=6, beta=0.6, nwrap=4, u1=1.0, u2=-0.1, mu=-5.0)
    with pytest.raises(SolverError, match="multiple of Nwrap"):
        BatchedSolver([spec], [0], lib=emu_lib)
    spec.params = Params(nlx=3, nly=3, ltrot=6, beta=0.6, nwrap=2, u1=1.0, u2=0.3, mu=-5.0)
    with pytest.raises(SolverError, match="U2 must be non-positive"):
        BatchedSolver([spec], [0], lib=emu_lib)
    s = BatchedSolver([ModelSpec(p)], [0], lib=emu_lib)
    with pytest.raises(SolverError, match="field array size"):
        s.set_fields([np.zeros((5, 9, 2))], [1])
    with pytest.raises(SolverError, match="seeds must be positive"):
        s.set_fields([np.zeros((6, 9, 2))], [0])
    s.close()
