"""Memory kinds at the boundary: managed, pinned host, pageable host.

The reference prefetches managed memory to the current device before its
passes (src/transpose.cu:100-104) and exits on any other pointer.  The
library accepts every pointer the device reaches (device memory of the
current device, managed memory, pinned host memory), prefetches managed
memory, and refuses pageable host memory before anything is enqueued.
Results are bit-exact against the oracle whatever the memory kind.
"""
import ctypes
import os
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

HIP_MEM_ATTACH_GLOBAL = 1


def hip_runtime():
    """The HIP runtime already mapped into this process (the one torch and the
    library share), opened again to allocate managed memory."""
    torch.cuda.init()
    for line in open("/proc/self/maps"):
        path = line.split()[-1]
        if os.path.basename(path).startswith("libamdhip64.so"):
            hip = ctypes.CDLL(path)
            hip.hipMallocManaged.restype = ctypes.c_int
            hip.hipMallocManaged.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t,
                                             ctypes.c_uint]
            hip.hipFree.restype = ctypes.c_int
            hip.hipFree.argtypes = [ctypes.c_void_p]
            return hip
    pytest.skip("HIP runtime not mapped into the process")


class Managed:
    """hipMallocManaged buffer of n elements with a host numpy view."""

    def __init__(self, n, esz):
        self.hip = hip_runtime()
        self.ptr = ctypes.c_void_p()
        rc = self.hip.hipMallocManaged(ctypes.byref(self.ptr), max(n * esz, 4096),
                                       HIP_MEM_ATTACH_GLOBAL)
        if rc != 0 or not self.ptr.value:
            pytest.skip(f"hipMallocManaged unavailable (hipError {rc})")
        ct = ctypes.c_uint32 if esz == 4 else ctypes.c_uint64
        self.view = np.ctypeslib.as_array((ct * n).from_address(self.ptr.value))

    def free(self):
        if self.ptr.value:
            assert self.hip.hipFree(self.ptr) == 0
            self.ptr = ctypes.c_void_p()


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_memory_kinds(ipt):
    d = torch.zeros(1024, dtype=torch.int32, device="cuda")
    assert ipt.memory_kind(d) == (ipt.MEM_DEVICE, torch.cuda.current_device())
    h = torch.zeros(1024, dtype=torch.int32).pin_memory()
    assert ipt.memory_kind(h)[0] == ipt.MEM_HOST
    a = np.zeros(1024, dtype=np.uint32)
    assert ipt.memory_kind(a.ctypes.data)[0] == ipt.MEM_PAGEABLE
    m = Managed(1024, 4)
    try:
        assert ipt.memory_kind(m.ptr.value)[0] == ipt.MEM_MANAGED
    finally:
        m.free()


# (d1, d2, d3): GCS, square swap, block-cycle, decomposition, slice-in-LDS.
HOST_SHAPES = [(512, 384, 8), (1024, 1024, 1), (4096, 7, 1), (9000, 300, 1), (100, 300, 2)]


@pytest.mark.parametrize("esz", [4, 8])
def test_managed_memory_matches_oracle(ipt, oracle, esz):
    """Transposes of managed memory through the C ABI (prefetch first, as the
    reference), every dispatch route, and a batch with a managed instance."""
    L = ipt.lib()
    for k, (d1, d2, d3) in enumerate(HOST_SHAPES):
        n = d1 * d2 * d3
        print(f"[case] managed d1={d1} d2={d2} d3={d3} esz={esz}", flush=True)
        m = Managed(n, esz)
        try:
            a = oracle.random_bits(n, esz, 4100 + k)
            m.view[:] = a
            st = L.ipt_transpose_213(m.ptr, esz, d1, d2, d3, stream_ptr())
            assert st == 0, (st, L.ipt_last_error())
            torch.cuda.synchronize()
            assert np.array_equal(m.view, oracle.seq_tt_213(a, d1, d2, d3)), (d1, d2, d3, esz)
        finally:
            m.free()
    # a batch mixing a device instance and a managed one
    d1, d2, d3 = 300, 200, 3
    n = d1 * d2 * d3
    print(f"[case] managed batch esz={esz}", flush=True)
    m = Managed(n, esz)
    try:
        a = oracle.random_bits(n, esz, 4200)
        b = oracle.random_bits(n, esz, 4201)
        m.view[:] = a
        tb = torch.from_numpy(b.view(np.int32 if esz == 4 else np.int64)).cuda()
        descs = (ipt.MatrixDesc * 2)()
        descs[0] = ipt.MatrixDesc(m.ptr.value, esz, ipt.VARIANT_AUTO, d1, d2, d3)
        descs[1] = ipt.MatrixDesc(tb.data_ptr(), esz, ipt.VARIANT_AUTO, d1, d2, d3)
        h = ctypes.c_void_p()
        assert L.ipt_batch_create(descs, 2, None, ctypes.byref(h)) == 0, L.ipt_last_error()
        try:
            assert L.ipt_batch_run(h, stream_ptr()) == 0
            torch.cuda.synchronize()
            assert np.array_equal(m.view, oracle.seq_tt_213(a, d1, d2, d3))
            got_b = tb.cpu().numpy().view(np.uint32 if esz == 4 else np.uint64)
            assert np.array_equal(got_b, oracle.seq_tt_213(b, d1, d2, d3))
        finally:
            assert L.ipt_batch_destroy(h) == 0
    finally:
        m.free()
    ipt.device_check()


@pytest.mark.parametrize("esz", [4, 8])
def test_pinned_host_memory_matches_oracle(ipt, oracle, esz):
    """Pinned host tensors are transposed in place by the GPU over the host
    link; a rank-4 permutation too."""
    dt = torch.int32 if esz == 4 else torch.int64
    for k, (d1, d2, d3) in enumerate(HOST_SHAPES):
        n = d1 * d2 * d3
        print(f"[case] pinned d1={d1} d2={d2} d3={d3} esz={esz}", flush=True)
        a = oracle.random_bits(n, esz, 4300 + k)
        t = torch.from_numpy(a.view(np.int32 if esz == 4 else np.int64).copy()).pin_memory()
        assert t.dtype == dt
        ipt.transpose(t, d1, d2, d3)
        torch.cuda.synchronize()
        got = t.numpy().view(np.uint32 if esz == 4 else np.uint64)
        assert np.array_equal(got, oracle.seq_tt_213(a, d1, d2, d3)), (d1, d2, d3, esz)
    dims, perm = [6, 5, 4, 3], [2, 0, 3, 1]
    print(f"[case] pinned permute dims={dims} perm={perm} esz={esz}", flush=True)
    a = oracle.random_bits(360, esz, 4400)
    t = torch.from_numpy(a.view(np.int32 if esz == 4 else np.int64).copy()).pin_memory()
    ipt.permute(t, dims, perm)
    torch.cuda.synchronize()
    got = t.numpy().view(np.uint32 if esz == 4 else np.uint64)
    assert np.array_equal(got, oracle.seq_tt(a, dims, perm))
    ipt.device_check()


def test_pageable_host_memory_is_refused(ipt):
    """A malloc'd host buffer is refused by every entry point before anything
    is enqueued (a kernel would fault on it), and left untouched."""
    L = ipt.lib()
    a = np.arange(64 * 48, dtype=np.uint32)
    kind, _ = ipt.memory_kind(a.ctypes.data)
    if kind != ipt.MEM_PAGEABLE:
        pytest.skip(f"this system reports malloc'd memory as kind {kind}")
    p = ctypes.c_void_p(a.ctypes.data)
    print("[case] pageable refused: transpose_213", flush=True)
    assert L.ipt_transpose_213(p, 4, 64, 48, 1, stream_ptr()) == 1
    assert b"pageable" in L.ipt_last_error()
    print("[case] pageable refused: variants, permute, batch", flush=True)
    for v in (ipt.VARIANT_GCS, ipt.VARIANT_C2R, ipt.VARIANT_BLOCK_CYCLE):
        assert L.ipt_transpose_variant(p, 4, 64, 48, 1, v, None, stream_ptr()) == 1
    dims = (ctypes.c_int64 * 3)(4, 16, 48)
    perm = (ctypes.c_int * 3)(2, 0, 1)
    assert L.ipt_permute(p, 4, 3, dims, perm, stream_ptr()) == 1
    descs = (ipt.MatrixDesc * 1)()
    descs[0] = ipt.MatrixDesc(a.ctypes.data, 4, ipt.VARIANT_AUTO, 64, 48, 1)
    h = ctypes.c_void_p()
    assert L.ipt_batch_create(descs, 1, None, ctypes.byref(h)) == 1
    assert not h.value
    torch.cuda.synchronize()
    assert np.array_equal(a, np.arange(64 * 48, dtype=np.uint32))
    ipt.device_check()


def test_host_resident_rates(ipt, oracle):
    """Informational: the [512][384][64] f32 transpose on device, managed and
    pinned host memory (the last two cross the host link)."""
    d1, d2, d3, esz = 512, 384, 64, 4
    n = d1 * d2 * d3
    a = oracle.random_bits(n, esz, 4500)
    L = ipt.lib()

    def rate(ptr):
        for _ in range(2):
            assert L.ipt_transpose_213(ptr, esz, d1, d2, d3, stream_ptr()) == 0
        torch.cuda.synchronize()
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            assert L.ipt_transpose_213(ptr, esz, d1, d2, d3, stream_ptr()) == 0
        torch.cuda.synchronize()
        return 2.0 * n * esz * reps / (time.perf_counter() - t0) / 1e9

    dev = torch.from_numpy(a.view(np.int32)).cuda()
    r_dev = rate(ctypes.c_void_p(dev.data_ptr()))
    pin = torch.from_numpy(a.view(np.int32).copy()).pin_memory()
    r_pin = rate(ctypes.c_void_p(pin.data_ptr()))
    m = Managed(n, esz)
    try:
        m.view[:] = a
        r_man = rate(m.ptr)
    finally:
        m.free()
    print(f"[rate] {d1}x{d2}x{d3} f32 GB/s: device {r_dev:.0f}, managed {r_man:.0f}, "
          f"pinned host {r_pin:.0f}", flush=True)
