"""Test helper: reference-layout structs (DMRG_BASIS, CRS1 operators) built from a
snapshot, so the C ABI can be driven exactly as a dmrg_C program drives it."""
from __future__ import annotations

import ctypes as C

import numpy as np

import dmrg_c_amd as D


class BasisBundle:
    """DMRG_BASIS (+ OBC sub-bases when the snapshot has them) over numpy storage."""

    def __init__(self, snap):
        self.keep = []
        B = D.DMRG_BASIS()
        d, dl, dr, n = snap.dim_onsite, snap.dim_LL, snap.dim_RR, snap.n

        def s16(a):
            a = np.ascontiguousarray(a, np.int16)
            self.keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_short))

        def i32(a):
            a = np.ascontiguousarray(a, np.int32)
            self.keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int))

        def rows2d(flat, nrow, ncol):
            flat = np.ascontiguousarray(flat, np.int32).reshape(nrow, ncol)
            self.keep.append(flat)
            arr = (C.POINTER(C.c_int) * nrow)(*[flat[i].ctypes.data_as(C.POINTER(C.c_int)) for i in range(nrow)])
            self.keep.append(arr)
            return C.cast(arr, C.POINTER(C.POINTER(C.c_int)))

        B.LL_LLLRRRRL, B.LR_LLLRRRRL = s16(snap.LL), s16(snap.LR)
        B.RR_LLLRRRRL, B.RL_LLLRRRRL = s16(snap.RR), s16(snap.RL)
        self.inv = snap.inv_table()
        B.Inv_LLLRRRRL = i32(self.inv)
        B.dim_LLLRRRRL, B.dim_LL, B.dim_RR, B.dim_onsite = n, dl, dr, d
        if snap.obc is not None:
            o = snap.obc
            B.LL_LLLR, B.LR_LLLR = s16(o["LL_LLLR"]), s16(o["LR_LLLR"])
            B.LR_LRRL, B.RL_LRRL = s16(o["LR_LRRL"]), s16(o["RL_LRRL"])
            B.RR_RRRL, B.RL_RRRL = s16(o["RR_RRRL"]), s16(o["RL_RRRL"])
            B.dim_LLLR, B.dim_LRRL, B.dim_RRRL = len(o["LL_LLLR"]), len(o["LR_LRRL"]), len(o["RR_RRRL"])
            inv = np.full(dl * d, -1, np.int32)
            inv[o["LL_LLLR"].astype(np.int64) * d + o["LR_LLLR"]] = np.arange(B.dim_LLLR)
            B.Inv_LLLR = rows2d(inv, dl, d)
            inv = np.full(d * d, -1, np.int32)
            inv[o["LR_LRRL"].astype(np.int64) * d + o["RL_LRRL"]] = np.arange(B.dim_LRRL)
            B.Inv_LRRL = rows2d(inv, d, d)
            inv = np.full(dr * d, -1, np.int32)
            inv[o["RR_RRRL"].astype(np.int64) * d + o["RL_RRRL"]] = np.arange(B.dim_RRRL)
            B.Inv_RRRL = rows2d(inv, dr, d)
            self.M = {k: D.HostCRS(o[k].row, o[k].col, o[k].val, o[k].row_dim) for k in ("M_LLLR", "M_LRRL", "M_RRRL")}
            # empty sign matrix and zero parities, as xxz DIAGONALIZE_SUPERBLOCK.c:18-19
            self.M["M_LRRL_Sign"] = D.HostCRS(np.zeros(B.dim_LLLR + 1, np.int64), np.zeros(1, np.int32),
                                              np.zeros(1), B.dim_LLLR)
            self.ele_rr = np.zeros(max(B.dim_RRRL, dr), np.int32)
        self.B = B

    @property
    def ptr(self):
        return C.pointer(self.B)


def host_ops(snap):
    """CRS1 objects for the snapshot's operators (kept alive by the returned list)."""
    out = []
    for o in snap.ops:
        h = D.HostCRS(o.row, o.col, o.val, o.row_dim)
        h.crs.col_dim = o.col_dim
        out.append(h)
    return out


def call_terms(snap, ops, eles, basis_onsite, a_basis, inv_ptr, elem_num):
    """Run the snapshot's term list through the library's DMRG_MAKE_ELEM_* symbols, the way
    a model's MAKE_ELEMENT_HAM_LLLRRRRL does."""
    L = D.lib()
    types = {4: (b"LL_LR", b"LR_LL"), 5: (b"LL_RL", b"RL_LL"), 6: (b"LL_RR", b"RR_LL"), 7: (b"LR_RL", b"RL_LR"),
             8: (b"LR_RR", b"RR_LR"), 9: (b"RR_RL", b"RL_RR")}
    names = ["LL", "LR", "RR", "RL", "LLLR", "LLRL", "LLRR", "LRRL", "LRRR", "RRRL", "ZERO"]
    for t in snap.terms:
        k = int(t["kind"])
        fn = getattr(L, f"DMRG_MAKE_ELEM_{names[k]}_LLLRRRRL")
        if k == 10:
            fn(basis_onsite, a_basis, inv_ptr, elem_num)
            continue
        c = float(t["coeff"])
        if k < 4:
            fn(basis_onsite, a_basis, inv_ptr, ops[int(t["op1"])].ptr, c, elem_num)
            continue
        mode = int(t["sign_mode"])
        ty = types[k][mode - 1] if mode else None
        flag = b"Yes" if mode else b"No"
        e = [eles[int(t[f"ele{i}"])] if int(t[f"ele{i}"]) >= 0 else None for i in range(3)]
        ep = [x.ctypes.data_as(C.POINTER(C.c_int)) if x is not None else None for x in e]
        a, b = ops[int(t["op1"])].ptr, ops[int(t["op2"])].ptr
        if k in (4, 8, 9):
            fn(basis_onsite, a_basis, inv_ptr, a, b, c, ep[0], elem_num, ty, flag)
        elif k in (6, 7):
            fn(basis_onsite, a_basis, inv_ptr, a, b, c, ep[0], ep[1], elem_num, ty, flag)
        else:
            fn(basis_onsite, a_basis, inv_ptr, a, b, c, ep[0], ep[1], ep[2], elem_num, ty, flag)
