"""Operator renormalisation product CRS_CRS_CCS_PRODUCT (sml/CRS_CRS_CCS_PRODUCT.c:6-124):
Out = M1 (M2 M3) with entries of magnitude <= 1e-15 dropped after each pass, as
DMRG_TRANS_MAT_ONE/TWO/C_C use it for T^dagger O T.

Pins of the oracle (or_crs_crs_ccs_product): the reference's own routine compiled from its
sources (oracle/_ref/libdmrgref.so, skipped when not built), bit for bit, and a dense numpy
triple product.  The library's drop-in is compared bit for bit with the oracle on the GPU.
Cases: random sparse matrices (rows in random column order, a repeated column, entries
small enough to be dropped), and T^dagger O T with the transformation matrix of a
reference run's truncation."""
import ctypes as C
import subprocess
import sys
import os

import numpy as np
import pytest

import oracle_py
from conftest import REPO

REF_LIB = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")


def rand_crs(rng, n, m, density, shuffle=False, dup=False, tiny=False):
    rows, cols, vals = [0], [], []
    for _ in range(n):
        c = np.nonzero(rng.random(m) < density)[0]
        if shuffle:
            rng.shuffle(c)
        v = rng.uniform(-1, 1, c.size)
        if tiny and c.size:
            v[rng.random(c.size) < 0.2] *= 1e-16
        if dup and c.size:
            c = np.r_[c, c[0]]
            v = np.r_[v, rng.uniform(-1, 1)]
        cols += list(c)
        vals += list(v)
        rows.append(len(cols))
    return np.array(rows, np.int64), np.array(cols, np.int32), np.array(vals, np.float64)


def crs_dense(r, c, v, n, m, last_wins=False):
    A = np.zeros((n, m))
    for i in range(n):
        for e in range(r[i], r[i + 1]):
            if last_wins:
                A[i, c[e]] = v[e]
            else:
                A[i, c[e]] += v[e]
    return A


def oracle(M1, M2, M3, n1, n2, n3, k):
    L = oracle_py.lib()
    fn = L.or_crs_crs_ccs_product
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    args = [n1, *[a.ctypes.data for a in M1], n2, *[a.ctypes.data for a in M2], n3, k,
            *[a.ctypes.data for a in M3]]
    row = np.empty(n1 + 1, np.int64)
    nnz = fn(*args, row.ctypes.data, None, None)
    col, val = np.empty(max(nnz, 1), np.int32), np.empty(max(nnz, 1))
    fn(*args, row.ctypes.data, col.ctypes.data, val.ctypes.data)
    return row, col[:nnz], val[:nnz]


def call(fn, M1, M2, M3, n1, n2, n3, k, out_cap=None, work_cap=None, work_out=None):
    """The reference's signature: CRS_CRS_CCS_PRODUCT(M1, M2, M3, Out, Work, Work_Vec).
    out_cap / work_cap: Out->max_val / Work->max_val (default: dense size); work_out, a
    list, receives Work's (col pointers, rows, values) as the function left them."""
    import dmrg_c_amd as D

    def crs(r, c, v, n, m):
        return D.CRS1(len(v), n + 1, n, m, c.ctypes.data_as(C.POINTER(C.c_int)),
                      r.ctypes.data_as(C.POINTER(C.c_long)), v.ctypes.data_as(C.POINTER(C.c_double)))
    m1, m2 = crs(*M1, n1, n2), crs(*M2, n2, n3)
    m3 = D.CCS1(len(M3[2]), k + 1, n3, k, M3[0].ctypes.data_as(C.POINTER(C.c_long)),
                M3[1].ctypes.data_as(C.POINTER(C.c_int)), M3[2].ctypes.data_as(C.POINTER(C.c_double)))
    cap = max(n1 * k, 1)
    orow, ocol, oval = np.zeros(n1 + 1, np.int64), np.zeros(cap, np.int32), np.zeros(cap)
    out = crs(orow, ocol, oval, n1, k)
    out.max_val = cap if out_cap is None else out_cap
    wcap = max(n2 * k, 1)
    wcol, wrow, wval = np.zeros(k + 1, np.int64), np.zeros(wcap, np.int32), np.zeros(wcap)
    work = D.CCS1(wcap if work_cap is None else work_cap, k + 1, 0, 0, wcol.ctypes.data_as(C.POINTER(C.c_long)),
                  wrow.ctypes.data_as(C.POINTER(C.c_int)), wval.ctypes.data_as(C.POINTER(C.c_double)))
    wvec = np.zeros(max(n2, n3, 1))
    fn.argtypes = [C.POINTER(D.CRS1), C.POINTER(D.CRS1), C.POINTER(D.CCS1), C.POINTER(D.CRS1), C.POINTER(D.CCS1),
                   C.c_void_p]
    fn(C.byref(m1), C.byref(m2), C.byref(m3), C.byref(out), C.byref(work), wvec.ctypes.data)
    if work_out is not None:
        work_out[:] = [wcol, wrow, wval]
    assert out.row_dim == n1 and out.col_dim == k
    nnz = int(orow[n1])
    return orow, ocol[:nnz], oval[:nnz]


def random_case(seed, n1, n2, n3, k, density=0.2):
    rng = np.random.default_rng(seed)
    M1 = rand_crs(rng, n1, n2, density, shuffle=True, dup=True)
    M2 = rand_crs(rng, n2, n3, density, shuffle=True, tiny=True)
    c3 = rand_crs(rng, k, n3, density)   # CCS of M3 = CRS of its transpose
    return M1, M2, c3, (n1, n2, n3, k)


def truncation_case():
    """T^dagger O T: T from the oracle truncation of a reference run's target sector."""
    from test_density import SUBS
    from test_truncate import oracle_info
    S = SUBS[0]
    info, _, _ = oracle_info(S, max(1, S.dim_LLLR // 3))
    T = info["T"]
    m, k = T.shape
    rng = np.random.default_rng(11)
    O = rand_crs(rng, m, m, 0.05)
    Td = [np.array(np.r_[0, np.cumsum((T.T != 0).sum(1))], np.int64),
          np.concatenate([np.nonzero(row)[0] for row in T.T]).astype(np.int32),
          np.concatenate([row[row != 0] for row in T.T])]
    M3 = [np.array(np.r_[0, np.cumsum((T != 0).sum(0))], np.int64),
          np.concatenate([np.nonzero(col)[0] for col in T.T]).astype(np.int32),
          np.concatenate([col[col != 0] for col in T.T])]
    return Td, O, M3, (k, m, m, k)


CASES = {f"random{s}": random_case(s, *dims) for s, dims in enumerate([(7, 9, 11, 5), (40, 60, 50, 30),
                                                                        (1, 1, 1, 1), (64, 300, 200, 17)])}
CASES["truncation"] = truncation_case()


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_dense(name):
    M1, M2, M3, (n1, n2, n3, k) = CASES[name]
    r, c, v = oracle(M1, M2, M3, n1, n2, n3, k)
    want = crs_dense(*M1, n1, n2, last_wins=True) @ crs_dense(*M2, n2, n3) @ crs_dense(*M3, k, n3, True).T
    got = crs_dense(r, c, v, n1, k)
    np.testing.assert_allclose(got, np.where(np.abs(want) > 1e-15, want, 0.0), atol=1e-13)


def test_dropped_threshold_matters():
    """The pass-1 drop is part of the result: an intermediate sum below 1e-15 that M1 would
    scale above it is gone from the output (an oracle without the drop fails here)."""
    rng = np.random.default_rng(3)
    M1 = (np.array([0, 1], np.int64), np.array([0], np.int32), np.array([1e5]))
    M2 = (np.array([0, 1], np.int64), np.array([0], np.int32), np.array([1e-16]))
    M3 = (np.array([0, 1], np.int64), np.array([0], np.int32), np.array([rng.uniform(1, 2)]))
    r, c, v = oracle(M1, M2, M3, 1, 1, 1, 1)
    exact = 1e5 * 1e-16 * M3[2][0]
    assert exact > 1e-15 and r[1] == 0   # the reference drops W = 1e-16 * m3 before M1 scales it


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_reference(name):
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libdmrgref.so not built (make -C oracle ref)")
    R = C.CDLL(REF_LIB)
    M1, M2, M3, dims = CASES[name]
    want = call(R.CRS_CRS_CCS_PRODUCT, M1, M2, M3, *dims)
    got = oracle(M1, M2, M3, *dims)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_device_matches_oracle(name):
    import dmrg_c_amd as D
    M1, M2, M3, dims = CASES[name]
    print(f"crs_crs_ccs {name} dims {dims}", flush=True)
    got = call(D.lib().CRS_CRS_CCS_PRODUCT, M1, M2, M3, *dims)
    want = oracle(M1, M2, M3, *dims)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


# ------------------------------------------------------------------ capacity checks
# The reference checks the kept count against Work->max_val / Out->max_val before every
# position it evaluates (sml/CRS_CRS_CCS_PRODUCT.c:56,101), so it exits as soon as the
# entries kept before the last evaluated position reach the capacity -- also when the last
# position is dropped and the result would have fit exactly.  Each call runs in a child
# process (the drop-in exits); the capacities at the boundary are taken from the
# reference's own Work and Out for a case whose last positions are dropped.
CAP_CASE = "random0"


def _cap_boundaries():
    R = C.CDLL(REF_LIB)
    M1, M2, M3, (n1, n2, n3, k) = CASES[CAP_CASE]
    w = []
    orow, ocol, _ = call(R.CRS_CRS_CCS_PRODUCT, M1, M2, M3, n1, n2, n3, k, work_out=w)
    wcol, wrow, _ = w
    w_nnz, o_nnz = int(wcol[k]), int(orow[n1])
    w_last = w_nnz > wcol[k - 1] and wrow[w_nnz - 1] == n2 - 1     # W(n2-1, k-1) kept
    o_last = o_nnz > orow[n1 - 1] and ocol[o_nnz - 1] == k - 1     # Out(n1-1, k-1) kept
    return w_nnz - int(w_last), o_nnz - int(o_last), w_nnz, o_nnz


def _run_child(which, out_cap, work_cap):
    code = (f"import sys; sys.path[:0] = [{os.path.dirname(os.path.abspath(__file__))!r}, {REPO!r}]\n"
            "import ctypes as C, test_transmat as T\n"
            f"M1, M2, M3, dims = T.CASES[{CAP_CASE!r}]\n"
            + ("fn = C.CDLL(T.REF_LIB).CRS_CRS_CCS_PRODUCT\n" if which == "ref"
               else "import dmrg_c_amd as D; fn = D.lib().CRS_CRS_CCS_PRODUCT\n")
            + f"T.call(fn, M1, M2, M3, *dims, out_cap={out_cap}, work_cap={work_cap})\n")
    env = dict(os.environ, DMRGX_SML_OUT="0")
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, timeout=120).returncode


def _cap_cases():
    w_before, o_before, w_nnz, o_nnz = _cap_boundaries()
    big = 10 ** 6
    # (Out->max_val, Work->max_val, exits?)
    return [(o_before, big, True), (o_before + 1, big, False), (big, w_before, True), (big, w_before + 1, False),
            (o_nnz, big, o_nnz == o_before), (big, w_nnz, w_nnz == w_before)]


def test_reference_capacity_rule():
    """Pins the rule on the reference itself (and that the case exercises a dropped last
    position for both passes, where the rule differs from 'capacity >= kept entries')."""
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libdmrgref.so not built (make -C oracle ref)")
    w_before, o_before, w_nnz, o_nnz = _cap_boundaries()
    assert w_nnz == w_before and o_nnz == o_before, "case should drop the last position of W and Out"
    for out_cap, work_cap, exits in _cap_cases():
        assert (_run_child("ref", out_cap, work_cap) != 0) == exits, (out_cap, work_cap, exits)


@pytest.mark.gpu
def test_device_capacity_rule():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libdmrgref.so not built (make -C oracle ref)")
    for out_cap, work_cap, exits in _cap_cases():
        print(f"crs_crs_ccs capacity out={out_cap} work={work_cap}", flush=True)
        assert (_run_child("dev", out_cap, work_cap) != 0) == exits, (out_cap, work_cap, exits)
