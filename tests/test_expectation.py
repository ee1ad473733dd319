"""On-site and intersite expectation values (DMRG_EXPECTATION_ONSITE,
DMRG_EXPECTATION_INTERSITE_Q0, DMRG_V_M_*_Q0) against the reference's own outputs."""
import ctypes as C
import glob
import os
import tempfile

import numpy as np
import pytest

import dmrg_c_amd as D
import oracle_py as O
from conftest import GOLDEN
from dmrg_c_amd import snapshot

EXPECTS = sorted(glob.glob(os.path.join(GOLDEN, "*_expect_*.npz")))
IDS = [os.path.basename(p)[:-4] for p in EXPECTS]


@pytest.mark.parametrize("path", EXPECTS, ids=IDS)
def test_oracle_expectation_matches_reference(path):
    """Oracle restatement on the reference's own inputs reproduces its outputs."""
    e = snapshot.load_expect(path)
    out = O.expectation(e)
    assert np.allclose(out, e.out, rtol=1e-12, atol=1e-13)


class _Call:
    """Reference-layout arguments for one expectation call."""

    def __init__(self, e):
        self.keep = []

        def crs(c):
            h = D.HostCRS(c.row, c.col, c.val, c.row_dim)
            h.crs.col_dim = c.col_dim
            self.keep.append(h)
            return h.ptr

        def crs_array(lst):
            arr = (C.POINTER(D.CRS1) * max(1, len(lst)))(*[crs(c) for c in lst])
            self.keep.append(arr)
            return arr

        self.M_LL, self.M_RR, self.M_CF = crs_array(e.M_LL), crs_array(e.M_RR), crs_array(e.M_CF)
        self.M_On = crs(e.M_On)
        B = D.DMRG_BASIS()
        for k in ("LL", "LR", "RR", "RL"):
            a = np.ascontiguousarray(getattr(e, k), np.int16)
            self.keep.append(a)
            setattr(B, f"{k}_LLLRRRRL", a.ctypes.data_as(C.POINTER(C.c_short)))
        B.dim_LLLRRRRL, B.dim_LL, B.dim_RR, B.dim_onsite = e.n, e.dim_LL, e.dim_RR, e.dim_onsite
        self.B = B
        st = D.DMRG_STATUS()
        st.LL_site, st.RR_site, st.dim_LLLRRRRL = e.LL_site, e.RR_site, e.n
        self.St = st


def run_expectation(e, vec):
    L = D.lib()
    L.DMRG_EXPECTATION_ONSITE.argtypes = [C.c_void_p, C.POINTER(D.CRS1), C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_int, C.POINTER(D.DMRG_BASIS), C.POINTER(D.DMRG_STATUS)]
    L.DMRG_EXPECTATION_INTERSITE_Q0.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(D.CRS1), C.c_void_p, C.c_void_p,
                                                C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                                C.POINTER(D.DMRG_BASIS), C.POINTER(D.DMRG_STATUS)]
    a = _Call(e)
    v = np.ascontiguousarray(vec, np.float64)
    t1, t2 = np.zeros(e.n), np.zeros(e.n)
    out = np.zeros(len(e.out))
    if e.kind == 0:
        L.DMRG_EXPECTATION_ONSITE(C.cast(a.M_LL, C.c_void_p), a.M_On, C.cast(a.M_RR, C.c_void_p), out.ctypes.data,
                                  v.ctypes.data, t1.ctypes.data, 1, C.byref(a.B), C.byref(a.St))
    else:
        L.DMRG_EXPECTATION_INTERSITE_Q0(C.cast(a.M_CF, C.c_void_p), C.cast(a.M_LL, C.c_void_p), a.M_On,
                                        C.cast(a.M_RR, C.c_void_p), out.ctypes.data, e.origin, v.ctypes.data,
                                        t1.ctypes.data, t2.ctypes.data, 1, C.byref(a.B), C.byref(a.St))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("path", EXPECTS, ids=IDS)
def test_gpu_expectation_same_vector(path):
    """Device expectation values on the reference's vector: within 1e-12 of its outputs."""
    e = snapshot.load_expect(path)
    out = run_expectation(e, e.vec)
    assert np.allclose(out, e.out, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("path", [p for p in EXPECTS if p.endswith("_0.npz") or p.endswith("_1.npz")],
                         ids=[i for i in IDS if i.endswith("_0") or i.endswith("_1")])
def test_gpu_expectation_end_to_end(path):
    """Ground state from this library's solver on the same superblock, then expectation
    values: within 1e-8 of the reference's (the acceptance bound for observables)."""
    e = snapshot.load_expect(path)
    name = os.path.basename(path).split("_expect_")[0]
    step = os.path.basename(path).split("_expect_s")[1].split("_")[0]
    s = snapshot.load(os.path.join(GOLDEN, f"{name}_s{step}.npz"))
    assert np.array_equal(s.LL, e.LL) and s.n == e.n
    old = os.getcwd()
    os.chdir(tempfile.mkdtemp())
    try:
        E, err, vec, _ = D.diagonalize_superblock_sym(s.ham.row, s.ham.col, s.ham.val)
    finally:
        os.chdir(old)
    out = run_expectation(e, vec)
    assert np.abs(out - e.out).max() <= 1e-8
