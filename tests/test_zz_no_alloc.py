"""bafq_net_forward_noalloc: the network's forward pass with its hidden
activations in the library's static device scratch, called through the C ABI
with caller-owned device buffers; its output must equal bafq_net_forward's
(the same convolution kernels run, so bit for bit) on consecutive calls, the
second with a smaller batch, including a batch the scratch splits into chunks."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from bafqmc_amd.native import NetDesc, load_library
from bafqmc_amd.proposal import ConvNet

SCRATCH_FLOATS = 16 << 20


def desc_of(net):
    return NetDesc(net.desc.lt, net.desc.ly, net.desc.lx, net.desc.channels, net.desc.hidden, net.desc.layers)


def test_noalloc_rejects_bad_arguments_emulator():
    lib = load_library(emulator=True)
    d = NetDesc(8, 3, 3, 2, 4, 3)
    buf = np.zeros(4096, dtype=np.float32)
    p = buf.ctypes.data
    assert lib.bafq_net_forward_noalloc(C.byref(d), p, 0, p, p) == 1            # batch < 1
    assert lib.bafq_net_forward_noalloc(C.byref(d), None, 1, p, p) == 1         # null parameters
    bad = NetDesc(8, 3, 3, 2, 4, 1)                                             # one layer
    assert lib.bafq_net_forward_noalloc(C.byref(bad), p, 1, p, p) == 1
    big = NetDesc(1024, 64, 64, 2, 32, 4)                                       # 2*32*4 Mi > 16 Mi floats
    assert 2 * 32 * 1024 * 64 * 64 > SCRATCH_FLOATS
    assert lib.bafq_net_forward_noalloc(C.byref(big), p, 1, p, p) == 2


def test_noalloc_matches_forward_emulator():
    """On the emulator device pointers are host pointers."""
    lib = load_library(emulator=True)
    net = ConvNet(8, 3, 3, 2, 6, 3, max_batch=4, seed=3, lib=lib)
    params = net.get_params().astype(np.float32)
    rng = np.random.default_rng(11)
    d = desc_of(net)
    for batch in (4, 2):
        x = rng.normal(size=(batch, 2, 8, 3, 3)).astype(np.float32)
        want = net.forward(x)
        y = np.full_like(x, np.nan)
        rc = lib.bafq_net_forward_noalloc(C.byref(d), params.ctypes.data, batch, x.ctypes.data, y.ctypes.data)
        assert rc == 0
        np.testing.assert_array_equal(y, want)
    net.close()


def hip_runtime():
    """The HIP runtime libbafqmc.so links (torch bundles another one, which must
    not be mixed into this process); device memory for the caller's buffers."""
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMalloc.restype = C.c_int
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipFree.restype = C.c_int
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipMemcpy.restype = C.c_int
    rt.hipDeviceSynchronize.restype = C.c_int
    return rt


H2D, D2H = 1, 2


def to_device(rt, a):
    ptr = C.c_void_p()
    assert rt.hipMalloc(C.byref(ptr), a.nbytes) == 0
    assert rt.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0
    return ptr


@pytest.mark.gpu
@pytest.mark.parametrize("shape,batches", [((80, 9, 9, 32, 4), (48, 5)),     # 9x9 lattice: 40 fields per chunk
                                           ((20, 3, 3, 8, 3), (16, 3))])
def test_noalloc_matches_forward_gpu(hip_lib, shape, batches):
    lt, ly, lx, hidden, layers = shape
    rt = hip_runtime()
    net = ConvNet(lt, ly, lx, 2, hidden, layers, max_batch=max(batches), seed=5, lib=hip_lib)
    d = desc_of(net)
    per_field = 2 * hidden * lt * ly * lx
    print(f"noalloc {shape}: {SCRATCH_FLOATS // per_field} fields per chunk, batches {batches}", flush=True)
    params = to_device(rt, np.ascontiguousarray(net.get_params(), dtype=np.float32))
    rng = np.random.default_rng(17)
    bufs = [params]
    try:
        for batch in batches:
            x = rng.normal(size=(batch, 2, lt, ly, lx)).astype(np.float32)
            want = net.forward(x)
            xd = to_device(rt, x)
            yd = to_device(rt, np.full_like(x, np.nan))
            bufs += [xd, yd]
            assert rt.hipDeviceSynchronize() == 0
            rc = hip_lib.bafq_net_forward_noalloc(C.byref(d), params, batch, xd, yd)
            assert rt.hipDeviceSynchronize() == 0
            assert rc == 0
            got = np.empty_like(x)
            assert rt.hipMemcpy(got.ctypes.data, yd, got.nbytes, D2H) == 0
            np.testing.assert_array_equal(got, want)
    finally:
        for b in bufs:
            rt.hipFree(b)
        net.close()
