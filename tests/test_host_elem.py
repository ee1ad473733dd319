"""Host element makers (DMRG_MAKE_ELEM_*_LLLRRRRL outside a recording) reproduce the
reference's lower-triangle CRS when driven row by row like GET_HAM_LLLRRRRL.c:69-94;
the recorder captures the model's term list.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import dmrg_c_amd as D
from conftest import GOLDEN
from dmrg_c_amd import snapshot
from sbhelp import call_terms, host_ops

CASES = ["xxz_L8_m16_s03", "xxz_L12_m24_jz05_s13", "hubbard_L8_m32_s04", "klm_L6_m24_ixy_s02"]


@pytest.mark.parametrize("name", CASES)
def test_host_element_makers_rebuild_reference_crs(name):
    s = snapshot.load(f"{GOLDEN}/{name}.npz")
    n = s.n
    ops = host_ops(s)
    eles = [np.ascontiguousarray(e, np.int32) for e in s.eles]
    inv = s.inv_table()
    inv_ptr = inv.ctypes.data_as(C.POINTER(C.c_int))
    cap = 4096
    ll, lr, rr, rl = (np.zeros(cap, np.int16) for _ in range(4))
    acc_inv = np.full(n, -1, np.int32)
    vals = np.zeros(cap)
    A = D.DMRG_A_BASIS(ll.ctypes.data_as(C.POINTER(C.c_short)), lr.ctypes.data_as(C.POINTER(C.c_short)),
                       rl.ctypes.data_as(C.POINTER(C.c_short)), rr.ctypes.data_as(C.POINTER(C.c_short)),
                       acc_inv.ctypes.data_as(C.POINTER(C.c_int)), vals.ctypes.data_as(C.POINTER(C.c_double)), 0)
    d, dr = s.dim_onsite, s.dim_RR
    rows, cols, vs = [0], [], []
    for r in range(n):
        B = D.DMRG_BASIS_ONSITE(int(s.LL[r]), int(s.LR[r]), int(s.RR[r]), int(s.RL[r]), r, dr, d)
        en = C.c_int(0)
        call_terms(s, ops, eles, C.byref(B), C.byref(A), inv_ptr, C.byref(en))
        m = en.value
        idx = ((ll[:m].astype(np.int64) * d + lr[:m]) * dr + rr[:m]) * d + rl[:m]
        invs = inv[idx]
        keep = ((invs >= 0) & (invs <= r) & (np.abs(vals[:m]) > 1e-15)) | (invs == r)
        order = np.argsort(invs[keep], kind="stable")
        cols.extend(invs[keep][order])
        vs.extend(vals[:m][keep][order])
        rows.append(len(cols))
        acc_inv[invs] = -1  # the reset loop of MAKE_ELEMENT_HAM_LLLRRRRL.c:82-90
    assert np.array_equal(np.array(rows), s.ham.row)
    assert np.array_equal(np.array(cols, np.int32), s.ham.col)
    assert np.array_equal(np.array(vs).view(np.int64), s.ham.val.view(np.int64))


def test_recorder_captures_term_list():
    s = snapshot.load(f"{GOLDEN}/hubbard_L8_m32_s04.npz")
    L = D.lib()
    ops = host_ops(s)
    eles = [np.ascontiguousarray(e, np.int32) for e in s.eles]
    assert L.dmrgx_record_begin() == 0
    assert L.dmrgx_record_begin() != 0  # nested recording is refused
    B = D.DMRG_BASIS_ONSITE(0, 0, 0, 0, 0, s.dim_RR, s.dim_onsite)
    call_terms(s, ops, eles, C.byref(B), None, None, None)
    T = L.dmrgx_record_end()
    assert L.dmrgx_terms_count(T) == len(s.terms)
    L.dmrgx_terms_free(T)


def test_superblock_rejects_out_of_range_inputs():
    """Validation happens on the host before any device work (no GPU needed)."""
    s = snapshot.load(f"{GOLDEN}/xxz_L8_m16_s03.npz")
    bad = snapshot.load(f"{GOLDEN}/xxz_L8_m16_s03.npz")
    bad.LL = bad.LL.copy()
    bad.LL[0] = s.dim_LL  # outside its block
    with pytest.raises(D.DmrgxError) as e:
        D.Superblock(bad)
    assert e.value.code == -1
