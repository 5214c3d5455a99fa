import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def hip_lib():
    """The product library; on a GPU box its absence is an error, never a skip."""
    from bafqmc_amd.native import load_library
    return load_library(emulator=False)


@pytest.fixture(scope="session")
def emu_lib():
    """The same device code compiled for the CPU (fibers per workgroup), with
    the GPU's workgroup sizes."""
    from bafqmc_amd.build import EMU_LIB, build_emu
    if not EMU_LIB.exists():
        build_emu()
    from bafqmc_amd.native import load_library
    return load_library(emulator=True)
