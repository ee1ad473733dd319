"""dmrgx_converge_check_noalloc: CONVERGE_CHECK (sml/CONVERGE_CHECK.c:7-23) over CSR arrays
and a vector already in device memory, with no device allocation (block partials in the
library's fixed per-GPU scratch array).  Checked against the allocating CONVERGE_CHECK on
the same arrays -- same kernels and schedule, so the same bits -- on two calls in a row,
the second on a smaller superblock (the scratch keeps the first call's partials; nothing
may leak into the second), and against the numpy L1 residual."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import dmrg_c_amd as D
import oracle_py as O
from conftest import GOLDEN
from dmrg_c_amd import snapshot

pytestmark = pytest.mark.gpu


def _lib():
    L = D.lib()
    L.dmrgx_device_alloc.argtypes = [C.c_int64, C.POINTER(C.c_void_p)]
    L.dmrgx_device_free.argtypes = [C.c_void_p]
    L.dmrgx_copy_to_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.dmrgx_converge_check_noalloc.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_double, C.POINTER(C.c_double)]
    L.CONVERGE_CHECK.restype = C.c_double
    L.CONVERGE_CHECK.argtypes = [C.POINTER(D.CRS1), C.c_void_p, C.c_double, C.c_int]
    return L


class DeviceArrays:
    def __init__(self, L, *arrays):
        self.L, self.ptrs = L, []
        for a in arrays:
            p = C.c_void_p()
            D.check(L.dmrgx_device_alloc(max(a.nbytes, 8), C.byref(p)))
            D.check(L.dmrgx_copy_to_device(p, a.ctypes.data, a.nbytes))
            self.ptrs.append(p)

    def free(self):
        for p in self.ptrs:
            self.L.dmrgx_device_free(p)


def _case(name, seed):
    s = snapshot.load(f"{GOLDEN}/{name}.npz")
    row, col, val = O.expand_sym(s.n, s.ham.row, s.ham.col, s.ham.val)
    v = s.meta["gs_vec"] + 1e-6 * np.random.default_rng(seed).standard_normal(s.n)
    return s.n, row, col, val, v, float(s.meta["gs_energy"][0])


def test_noalloc_matches_converge_check_twice():
    old = os.getcwd()
    os.chdir(tempfile.mkdtemp())
    try:
        L = _lib()
        for name, seed in (("klm_L6_m24_ixy_s04", 1), ("hubbard_L8_m32_s05", 2)):   # n 2158, then 200
            n, row, col, val, v, e = _case(name, seed)
            print(f"noalloc converge check {name} n={n} nnz={row[-1]}", flush=True)
            dev = DeviceArrays(L, row, col, val, v)
            out = C.c_double()
            D.check(L.dmrgx_converge_check_noalloc(n, int(row[-1]), *dev.ptrs, e, C.byref(out)))
            dev.free()
            H = D.HostCRS(row, col, val)
            want = L.CONVERGE_CHECK(H.ptr, v.ctypes.data, e, 1)
            assert np.float64(out.value).view(np.int64) == np.float64(want).view(np.int64), (out.value, want)
            r = np.abs(e * v - O.spmv_full(row, col, val, v)).sum()
            assert abs(out.value - r) <= 1e-12 + 1e-10 * r
    finally:
        os.chdir(old)


def test_noalloc_rejects_bad_arguments():
    L = _lib()
    out = C.c_double()
    assert L.dmrgx_converge_check_noalloc(-1, 0, None, None, None, None, 0.0, C.byref(out)) == -1   # DMRGX_ERR_ARG
    assert L.dmrgx_converge_check_noalloc(10, 30, None, None, None, None, 0.0, C.byref(out)) == -1
