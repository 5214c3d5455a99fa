"""ctypes binding of the C ABI declared in ``include/bafqmc.h``.

The HIP library ``_lib/libbafqmc.so`` is the product.  The emulator build
``_lib/libbafqmc_emu.so`` runs the same kernels on CPU threads and is selected
only by an explicit ``load_library(emulator=True)`` from the CPU test suite.
There is no fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .build import HIP_LIB, EMU_LIB

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)


class ModelDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nlx", C.c_int32), ("nly", C.c_int32),
        ("ltrot", C.c_int32), ("nwrap", C.c_int32),
        ("beta", C.c_double), ("u1", C.c_double), ("u2", C.c_double), ("mu", C.c_double),
        ("rdelta", C.c_double), ("shift_loc", C.c_double),
        ("expk_p", c_double_p), ("expk_m", c_double_p),
    ]


class Schedule(C.Structure):
    """bafq_schedule: path -1 choose / 0 device memory / 1 LDS; delayed -1
    widest / 0 per-site / 4, 8, 12, 16."""
    _fields_ = [("path", C.c_int32), ("delayed", C.c_int32)]


class NetDesc(C.Structure):
    _fields_ = [("lt", C.c_int32), ("ly", C.c_int32), ("lx", C.c_int32),
                ("channels", C.c_int32), ("hidden", C.c_int32), ("layers", C.c_int32)]


c_float_p = C.POINTER(C.c_float)

SYMBOLS = {
    "bafq_create": (C.c_int, [C.POINTER(ModelDesc), C.c_int, c_int32_p, C.c_int, C.POINTER(C.c_void_p)]),
    "bafq_create_scheduled": (C.c_int, [C.POINTER(ModelDesc), C.c_int, c_int32_p, C.c_int, C.POINTER(Schedule),
                                        C.POINTER(C.c_void_p)]),
    "bafq_get_schedule": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Schedule)]),
    "bafq_pairs": (C.c_int, [C.c_void_p]),
    "bafq_destroy": (None, [C.c_void_p]),
    "bafq_last_error": (C.c_char_p, [C.c_void_p]),
    "bafq_dims": (C.c_int, [C.c_void_p, C.c_int, c_int32_p, c_int32_p, c_int32_p, c_int32_p]),
    "bafq_set_fields": (C.c_int, [C.c_void_p, c_double_p, c_int32_p]),
    "bafq_get_fields": (C.c_int, [C.c_void_p, c_double_p, c_int32_p]),
    "bafq_warmup": (C.c_int, [C.c_void_p, C.c_int]),
    "bafq_prepare": (C.c_int, [C.c_void_p]),
    "bafq_sweep": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "bafq_sweep_tau": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bafq_get_green_tau": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "bafq_read_bin": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "bafq_get_stats": (C.c_int, [C.c_void_p, c_double_p]),
    "bafq_net_create": (C.c_int, [C.POINTER(NetDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "bafq_net_destroy": (None, [C.c_void_p]),
    "bafq_net_last_error": (C.c_char_p, [C.c_void_p]),
    "bafq_net_param_count": (C.c_int64, [C.c_void_p]),
    "bafq_net_set_params": (C.c_int, [C.c_void_p, c_float_p]),
    "bafq_net_get_params": (C.c_int, [C.c_void_p, c_float_p]),
    "bafq_net_forward": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p]),
    "bafq_net_forward_noalloc": (C.c_int, [C.POINTER(NetDesc), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "bafq_net_gradient": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p, c_double_p, c_float_p]),
    "bafq_net_time_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p]),
    "bafq_net_train": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p, C.c_int, C.c_float, c_double_p]),
    "bafq_global_propose": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, c_double_p, C.c_uint64, c_double_p]),
    "bafq_global_finish": (C.c_int, [C.c_void_p, c_int32_p]),
    "bafq_hmc_propose": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int, c_double_p, c_double_p, C.c_uint64,
                                   c_double_p]),
    "bafq_measure_pass": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "bafq_grad_log_weight": (C.c_int, [C.c_void_p, c_double_p]),
    "bafq_get_green": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "bafq_last_kernel_ms": (C.c_int, [C.c_void_p, c_double_p]),
    "bafq_debug_gemm": (C.c_int, [C.c_int, c_double_p, C.c_int, c_double_p, C.c_int, c_double_p]),
    "bafq_debug_qrp": (C.c_int, [C.c_int, c_double_p, c_double_p, c_int32_p, c_double_p, C.c_int]),
    "bafq_debug_mfma_probe": (C.c_int, [c_double_p]),
    "bafq_debug_mfma32_probe": (C.c_int, [C.POINTER(C.c_float)]),
    "bafq_debug_lane_ops": (C.c_int, [c_double_p, c_int32_p, c_double_p]),
}


def dptr(a: np.ndarray):
    return a.ctypes.data_as(c_double_p)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(c_int32_p)


class NativeLibrary:
    def __init__(self, path: Path):
        path = Path(path)
        if not path.exists():
            raise FileNotFoundError(
                f"native library {path} is missing: build it with `python -m bafqmc_amd.build`")
        self.path = path
        self.emulator = path.name == EMU_LIB.name
        self.lib = C.CDLL(str(path))
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args

    def __getattr__(self, name):
        return getattr(self.lib, name)

    def debug_gemm(self, a, b, op_a=0, op_b=0):
        n = a.shape[0]
        a = np.ascontiguousarray(a, dtype=np.complex128)
        b = np.ascontiguousarray(b, dtype=np.complex128)
        c = np.zeros((n, n), np.complex128)
        rc = self.lib.bafq_debug_gemm(n, dptr(a.view(np.float64)), op_a, dptr(b.view(np.float64)), op_b,
                                      dptr(c.view(np.float64)))
        if rc:
            raise RuntimeError("bafq_debug_gemm failed")
        return c

    def debug_mfma_probe(self):
        out = np.zeros((64, 64, 4))
        if self.lib.bafq_debug_mfma_probe(dptr(out)):
            raise RuntimeError("bafq_debug_mfma_probe failed")
        return out

    def debug_mfma32_probe(self):
        out = np.zeros((64, 64, 4), np.float32)
        if self.lib.bafq_debug_mfma32_probe(out.ctypes.data_as(C.POINTER(C.c_float))):
            raise RuntimeError("bafq_debug_mfma32_probe failed")
        return out

    def debug_lane_ops(self, v, idx):
        v = np.ascontiguousarray(v, dtype=np.float64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        assert v.shape == (64,) and idx.shape == (64,)
        out = np.zeros(320)
        if self.lib.bafq_debug_lane_ops(dptr(v), iptr(idx), dptr(out)):
            raise RuntimeError("bafq_debug_lane_ops failed")
        return out

    def debug_qrp(self, a, variant=0):
        n = a.shape[0]
        a = np.ascontiguousarray(a, dtype=np.complex128).copy()
        tau = np.zeros(n, np.complex128)
        perm = np.zeros(n, np.int32)
        d = np.zeros(n)
        rc = self.lib.bafq_debug_qrp(n, dptr(a.view(np.float64)), dptr(tau.view(np.float64)), iptr(perm), dptr(d),
                                     variant)
        if rc:
            raise RuntimeError("bafq_debug_qrp failed")
        return a, tau, perm, d


_CACHE = {}


def load_library(emulator: bool = False) -> NativeLibrary:
    """Load the HIP library (default) or, when asked explicitly, the CPU emulator."""
    path = EMU_LIB if emulator else HIP_LIB
    if path not in _CACHE:
        _CACHE[path] = NativeLibrary(path)
    return _CACHE[path]
