"""MI355X solver for the BAFQMC Markov-chain sweep (number-conserving and Nambu pairing).

The sweep runs in hand-written gfx950 HIP kernels (``bafqmc_amd/csrc``) behind the
C ABI declared in ``include/bafqmc.h``; this package is the Python host that keeps
the reference's input files, driver loop and per-bin output layout.
"""
from .model import ModelSpec, Params, read_params, kinetic_propagators
from .native import NativeLibrary, load_library
from .solver import BatchedSolver

__all__ = ["ModelSpec", "Params", "read_params", "kinetic_propagators",
           "NativeLibrary", "load_library", "BatchedSolver"]
