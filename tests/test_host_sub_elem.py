"""Sub-Hamiltonian element makers of the matrix-free path (DMRG_MAKE_ELEM_LL_LLLR, _LR_LLLR,
_LLLR_LLLR, _LRRL_LRRL, _RL_RRRL, _RR_RRRL, _RRRL_RRRL; include/dmrg.h:342-350) against the
reference's own functions compiled from its sources (oracle/_ref/libdmrgref.so): on seeded
random two-slot sub-bases (about a third of the product states absent), random sparse
operators with entries near the 1e-15 drop and repeated targets, parity arrays and both
sign orders, every row of the sub-basis is accumulated term after term into the same
DMRG_A_BASIS by both, and the accumulators (entry count, slot indices, values, inverse map)
must be identical bit for bit.  Host code only (no GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

import dmrg_c_amd as D
from conftest import REPO

REF_LIB = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")

SLOT = {"LL": 0, "LR": 1, "RR": 2, "RL": 3}
# name -> (sub-basis slots, operator slots, sign: parity slot and the two type strings)
MAKERS = {
    "DMRG_MAKE_ELEM_LL_LLLR": (("LL", "LR"), ("LL",), None),
    "DMRG_MAKE_ELEM_LR_LLLR": (("LL", "LR"), ("LR",), None),
    "DMRG_MAKE_ELEM_LLLR_LLLR": (("LL", "LR"), ("LL", "LR"), ("LL", "LL_LR", "LR_LL")),
    "DMRG_MAKE_ELEM_LRRL_LRRL": (("LR", "RL"), ("LR", "RL"), ("LR", "LR_RL", "RL_LR")),
    "DMRG_MAKE_ELEM_RL_RRRL": (("RR", "RL"), ("RL",), None),
    "DMRG_MAKE_ELEM_RR_RRRL": (("RR", "RL"), ("RR",), None),
    "DMRG_MAKE_ELEM_RRRL_RRRL": (("RR", "RL"), ("RR", "RL"), ("RR", "RR_RL", "RL_RR")),
}


def _rand_op(rng, dim):
    rows, cols, vals = [0], [], []
    for _ in range(dim):
        k = int(rng.integers(0, 4))
        c = rng.integers(0, dim, size=k)          # repeated targets allowed
        v = rng.choice([1.0, -0.5, 2.0 ** -52, 3e-16, rng.normal()], size=k)
        cols.extend(c)
        vals.extend(v)
        rows.append(len(cols))
    return (np.array(rows, np.int64), np.array(cols, np.int32), np.array(vals, np.float64))


class Acc:
    """A DMRG_A_BASIS over numpy storage."""

    def __init__(self, cap, n_inv):
        self.f = [np.zeros(cap, np.int16) for _ in range(4)]
        self.val = np.zeros(cap)
        self.inv = np.full(n_inv, -1, np.int32)
        sp = lambda a: a.ctypes.data_as(C.POINTER(C.c_short))
        self.s = D.DMRG_A_BASIS(sp(self.f[0]), sp(self.f[1]), sp(self.f[3]), sp(self.f[2]),
                                self.inv.ctypes.data_as(C.POINTER(C.c_int)),
                                self.val.ctypes.data_as(C.POINTER(C.c_double)), 0)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(MAKERS))
def test_sub_element_makers_equal_reference(name, seed):
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libdmrgref.so not built (make -C oracle ref)")
    rng = np.random.default_rng(seed)
    (sa, sb), op_slots, sign = MAKERS[name]
    dims = {"LL": 9, "LR": 4, "RR": 7, "RL": 4}
    da, db = dims[sa], dims[sb]
    present = rng.random((da, db)) < 0.65
    inv2 = np.where(present, np.cumsum(present).reshape(da, db) - 1, -1).astype(np.int32)
    n_sub = int(present.sum())
    row_ptrs = (C.POINTER(C.c_int) * da)(*[inv2[i].ctypes.data_as(C.POINTER(C.c_int)) for i in range(da)])
    ops = [_rand_op(rng, dims[s]) for s in op_slots]
    crs = [D.HostCRS(*o) for o in ops]
    ele = np.ascontiguousarray(rng.integers(0, 5, size=dims[sign[0]] if sign else 1), np.int32)
    libs = {"ref": C.CDLL(REF_LIB), "lib": D.lib()}
    head = [C.POINTER(D.DMRG_BASIS_ONSITE), C.POINTER(D.DMRG_A_BASIS), C.c_void_p, C.POINTER(D.CRS1)]
    for L in libs.values():
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = (head + [C.POINTER(D.CRS1), C.c_double, C.c_void_p, C.POINTER(C.c_int), C.c_char_p, C.c_char_p]
                       if sign else head + [C.c_double, C.POINTER(C.c_int)])
    accs = {k: Acc(4096, n_sub) for k in libs}
    en = {k: C.c_int(0) for k in libs}
    states = np.argwhere(present)
    for qa, qb in states:
        q = {sa: int(qa), sb: int(qb)}
        B = D.DMRG_BASIS_ONSITE(q.get("LL", 0), q.get("LR", 0), q.get("RR", 0), q.get("RL", 0), 0, dims["RR"], 4)
        for k in libs:
            en[k].value = 0
        for t in range(3):   # several terms into the same row accumulator, as a model's term list does
            coeff = float(rng.choice([0.5, -1.25, 1e-16, rng.normal()]))
            typ = rng.choice([sign[1], sign[2]]).encode() if sign else b""
            flag = rng.choice([b"Yes", b"No"]) if sign else b""
            for k, L in libs.items():
                fn = getattr(L, name)
                if sign:
                    fn(C.byref(B), C.byref(accs[k].s), C.cast(row_ptrs, C.c_void_p), crs[0].ptr, crs[1].ptr, coeff,
                       ele.ctypes.data, C.byref(en[k]), typ, flag)
                else:
                    fn(C.byref(B), C.byref(accs[k].s), C.cast(row_ptrs, C.c_void_p), crs[0].ptr, coeff, C.byref(en[k]))
        a, b = accs["lib"], accs["ref"]
        m = en["ref"].value
        assert en["lib"].value == m
        for s in (sa, sb):
            assert np.array_equal(a.f[SLOT[s]][:m], b.f[SLOT[s]][:m])
        assert np.array_equal(a.val[:m].view(np.int64), b.val[:m].view(np.int64))
        assert np.array_equal(a.inv, b.inv)
        for acc in accs.values():   # reset for the next row
            acc.inv[:] = -1
