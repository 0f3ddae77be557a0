"""ibim_apply_noalloc: the product of ibim_apply on device vectors with its
workspace in the library's fixed device array (no allocation).

Checked through the C ABI against ibim_apply on the same operator and vectors,
bit for bit (the same kernels and summation order), on two calls in a row of
which the second has a smaller operator (the array still holds the first
call's planes), for stored pair storage, the FP64 and mixed matrix-free pair
operators, and the dense-plane layout of the combined-field tube's
matrix-free sweep; sizes beyond the array are refused before any launch.
The vectors' device memory is the test's own (HIP runtime through ctypes).
"""
import ctypes as C
import os

import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu


def sphere_problem(gpu, n, kind=0):
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(n))
    pb.geometry(5)
    h = 4.0 / n
    pb.tube(gpu.params(3.0, 3 * h, 2 * h, delta_in=0.3 * h if kind else 0.0, dirichlet=1), kind)
    return pb


def hip_runtime():
    """The HIP runtime instance libibim.so is bound to: its DT_NEEDED soname,
    opened without loading anything new (another module, torch, may ship its
    own copy of the runtime file)."""
    import ibim
    ibim.lib()
    return C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOW | os.RTLD_NOLOAD)


class DeviceVectors:
    """count complex128 values per vector in device memory of the test's own
    (HIP runtime through ctypes), filled from and read back to numpy."""

    def __init__(self, *host):
        self.hip = hip_runtime()
        self.n = len(host[0])
        self.ptrs = []
        for h in host:
            ptr = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(ptr), C.c_size_t(16 * self.n)) == 0
            a = np.ascontiguousarray(h, dtype=np.complex128)
            assert self.hip.hipMemcpy(ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(16 * self.n), 1) == 0
            self.ptrs.append(ptr)

    def read(self, i):
        out = np.zeros(self.n, dtype=np.complex128)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptrs[i], C.c_size_t(16 * self.n), 2) == 0
        return out

    def free(self):
        for ptr in self.ptrs:
            self.hip.hipFree(ptr)
        self.ptrs = []


def noalloc(pb, op, p, q):
    nan = np.full(len(p), complex(np.nan, np.nan))
    dev = DeviceVectors(p, q, nan, nan)  # outputs start as NaN: every element must be written
    try:
        pb.apply_noalloc(dev.ptrs[0].value, dev.ptrs[1].value, dev.ptrs[2].value, dev.ptrs[3].value, op)
        return dev.read(2), dev.read(3)
    finally:
        dev.free()


def vectors(count, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(count) + 1j * rng.standard_normal(count) for _ in range(2)]


def check(gpu, pb, op, seed):
    p, q = vectors(pb.count, seed)
    y, z = pb.apply(p, q, op)
    yn, zn = noalloc(pb, op, p, q)
    assert np.array_equal(yn, y) and np.array_equal(zn, z)


def test_noalloc_product_equals_ibim_apply_on_two_calls(gpu):
    big = sphere_problem(gpu, 40)
    small = sphere_problem(gpu, 16)
    assert small.count < big.count
    big.assemble(gpu.C128_PAIR)
    check(gpu, big, gpu.OP_STORED, 1)
    small.assemble(gpu.C64_PAIR)
    check(gpu, small, gpu.OP_STORED, 2)
    # matrix-free pair operators (their tables exist after the pair assembly)
    for op in (gpu.OP_FREE64, gpu.OP_FREE32):
        check(gpu, big, op, 3)
        check(gpu, small, op, 4)


def test_noalloc_dense_plane_layout(gpu):
    # the combined-field (combo) tube sweeps matrix-free into dense planes
    pb = sphere_problem(gpu, 16, kind=1)
    check(gpu, pb, gpu.OP_FREE64, 5)
    pb.assemble(gpu.C128)
    check(gpu, pb, gpu.OP_STORED, 6)


def test_noalloc_refuses_what_it_cannot_do_without_allocating(gpu):
    # beyond the array: refused before any launch
    pb = sphere_problem(gpu, 96)
    assert pb.count > 30000
    pb.assemble(gpu.C64_PAIR)
    p, q = vectors(pb.count, 7)
    with pytest.raises(gpu.IbimError, match="scratch"):
        noalloc(pb, gpu.OP_STORED, p, q)
    # matrix-free tables not built yet: they would need an allocation
    fresh = sphere_problem(gpu, 16)
    p, q = vectors(fresh.count, 8)
    with pytest.raises(gpu.IbimError, match="tables"):
        noalloc(fresh, gpu.OP_FREE64, p, q)
    # and it still works after the refusals
    fresh.assemble(gpu.C128_PAIR)
    check(gpu, fresh, gpu.OP_STORED, 9)
