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
    # small workgroups keep the thread-per-lane emulator fast (small N only)
    os.environ.setdefault("BAFQ_BLOCK", "64")
    os.environ.setdefault("BAFQ_LDS_BLOCK", "128")
    from bafqmc_amd.build import EMU_LIB, build_emu
    if not EMU_LIB.exists():
        build_emu()
    from bafqmc_amd.native import load_library
    return load_library(emulator=True)
