"""Build the native libraries in-tree.

* ``_lib/libbafqmc.so``      -- gfx950 HIP kernels + C ABI (hipcc).
* ``_lib/libbafqmc_emu.so``  -- the same kernels on the CPU emulator back end
  (g++), used by the CPU test suite only.

Run ``python -m bafqmc_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIBDIR = HERE / "_lib"
HIP_LIB = LIBDIR / "libbafqmc.so"
EMU_LIB = LIBDIR / "libbafqmc_emu.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BAFQ_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build_hip(out: Path = HIP_LIB):
    LIBDIR.mkdir(exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-I", CSRC, CSRC / "bafqmc.hip", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_emu(out: Path = EMU_LIB):
    LIBDIR.mkdir(exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    tmp = out.with_suffix(".so.tmp")
    _run([cxx, "-std=c++20", "-O2", "-g", "-DBAFQ_EMU", "-fPIC", "-shared", "-pthread",
          "-I", CSRC, CSRC / "emu.cpp", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    build_emu()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "emu" in what:
        build_emu()
    if "hip" in what:
        build_hip()
    if "all" in what:
        build_all()
