"""Single-slot products between sectors: DMRG_V_M_{LL,LR,RR,RL}_Q1..Q4
(dmrg/DMRG_V_M_*_Q1..Q4.c), the products the expectation routines of the DMRG programs
apply to the ground state (main/dmrg/hubbard_vf/src/expectations/EXPECTATION_INTERSITE_Q2.c,
SC_CORRELATIONS.c).

The whole basis is built from the reference's own block quantum numbers and target-sector
basis (tests/golden/qn_*.npz, dumped from the compiled reference): sectors keyed by
(electrons, S^z) in enumeration order -- the order GET_WHOLE_BASIS_Q2_SUPERBLOCK
(main/dmrg/hubbard_vf/src/allocations/GET_WHOLE_BASIS_Q2_SUPERBLOCK.c:69-118) gives with one
thread -- the target sector in the reference's order.  Operators are random sparse
matrices on the slot's block that shift its quantum numbers, so that every state they
reach lies in the input sector.

Pins of the oracle: a dense Kronecker-product operator on the whole product space restricted
to the two sectors, and the reference's own DMRG_V_M_*_Q1 / _Q2 compiled from its sources
(oracle/_ref/libdmrgref.so, skipped when not built).  The library is then compared bit for
bit with the oracle on the GPU."""
import ctypes as C
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_py
from conftest import GOLDEN, REPO

QN_FILES = sorted(glob.glob(os.path.join(GOLDEN, "qn_*.npz")))
# the production-size steps (m=400, m=800) run on the GPU only (the m=800 one in test_klm_m800.py)
SMALL = [f for f in QN_FILES if "m400" not in f and "m800" not in f]
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")
SLOTS = ("LL", "LR", "RR", "RL")


def _load(path):
    z = np.load(path)
    n_qn, dim_LL, dim_RR, d, t_sz, t_ele = (int(v) for v in z["header"])
    q = dict(n_qn=n_qn, dim_LL=dim_LL, dim_RR=dim_RR, d=d, target=(t_ele if n_qn == 2 else 0, t_sz))
    for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR", "sz_site", "ele_site"):
        q[k] = z[k].astype(np.int64)
    if n_qn == 1:
        for k in ("ele_LL", "ele_RR", "ele_site"):
            q[k] = np.zeros_like(q[k])
    return q, tuple(z[k] for k in ("LL", "LR", "RR", "RL"))


def whole_basis(q, target_basis):
    """(inv, sectors): inv[flat state] = index within its sector; sectors[(ele, sz)] =
    (flat indices, LL, LR, RR, RL) in sector order."""
    dims = (q["dim_LL"], q["d"], q["dim_RR"], q["d"])
    LL, LR, RR, RL = (a.ravel() for a in np.meshgrid(*[np.arange(n) for n in dims], indexing="ij"))
    ele = q["ele_LL"][LL] + q["ele_site"][LR] + q["ele_RR"][RR] + q["ele_site"][RL]
    sz = q["sz_LL"][LL] + q["sz_site"][LR] + q["sz_RR"][RR] + q["sz_site"][RL]
    keys, gid = np.unique(np.stack([ele, sz], 1), axis=0, return_inverse=True)
    gid = gid.ravel()
    inv = np.full(LL.size, -1, np.int32)
    sectors = {}
    for g, (e, s) in enumerate(keys):
        flat = np.nonzero(gid == g)[0]
        if (int(e), int(s)) == q["target"]:
            tb = [np.asarray(a, np.int64) for a in target_basis]
            flat = np.ravel_multi_index(tuple(tb), dims)
        inv[flat] = np.arange(flat.size, dtype=np.int32)
        sectors[(int(e), int(s))] = (flat, LL[flat].astype(np.int16), LR[flat].astype(np.int16),
                                     RR[flat].astype(np.int16), RL[flat].astype(np.int16))
    return inv, sectors


def slot_qn(q, slot):
    return {0: (q["ele_LL"], q["sz_LL"]), 1: (q["ele_site"], q["sz_site"]), 2: (q["ele_RR"], q["sz_RR"]),
            3: (q["ele_site"], q["sz_site"])}[slot]


def shift_operator(q, slot, rng, shifts):
    """Random CRS (row, col, val) on the slot's block with nonzeros only where the row state's
    quantum numbers are the column state's plus (dele, dsz); the first shift in `shifts`
    whose output sector exists and whose operator is not empty."""
    e, s = slot_qn(q, slot)
    for de, ds in shifts:
        out_key = (q["target"][0] + de, q["target"][1] + ds)
        allowed = (e[:, None] == e[None, :] + de) & (s[:, None] == s[None, :] + ds)
        mask = allowed & (rng.random(allowed.shape) < 0.6)
        if mask.any():
            dense = np.where(mask, rng.uniform(-1, 1, mask.shape), 0.0)
            m = sp.csr_matrix(dense)
            m.sort_indices()
            return (m.indptr.astype(np.int64), m.indices.astype(np.int32), m.data.copy()), dense, (de, ds), out_key
    return None


def dense_product(q, slot, dense_op, sectors, out_key, x, sign):
    dims = (q["dim_LL"], q["d"], q["dim_RR"], q["d"])
    mats = [sp.identity(n, format="csr") for n in dims]
    mats[slot] = sp.csr_matrix(dense_op)
    O = mats[0]
    for m in mats[1:]:
        O = sp.kron(O, m, format="csr")
    out_flat = sectors[out_key][0]
    in_flat = sectors[q["target"]][0]
    y = O[out_flat][:, in_flat] @ x
    if sign and slot > 0:
        LL, LR, RR, _ = np.unravel_index(out_flat, dims)
        e = q["ele_LL"][LL] + (q["ele_site"][LR] if slot >= 2 else 0) + (q["ele_RR"][RR] if slot >= 3 else 0)
        y = np.where(e % 2 == 0, 1.0, -1.0) * y
    return y


def oracle_v_m(q, slot, op, inv, sec, x, sign):
    L = oracle_py.lib()
    fn = L.or_v_m_sector
    fn.restype = None
    fn.argtypes = [C.c_int, C.c_int32, C.c_int32, C.c_int32] + [C.c_void_p] * 11 + [C.c_int32, C.c_void_p,
                                                                                      C.c_void_p]
    _, LL, LR, RR, RL = sec
    eles = [np.ascontiguousarray(q[k], np.int32) for k in ("ele_LL", "ele_site", "ele_RR")]
    y = np.empty(LL.size)
    x = np.ascontiguousarray(x)
    fn(slot, LL.size, q["dim_RR"], q["d"], LL.ctypes.data, LR.ctypes.data, RR.ctypes.data, RL.ctypes.data,
       inv.ctypes.data, op[0].ctypes.data, op[1].ctypes.data, op[2].ctypes.data, *[e.ctypes.data for e in eles],
       1 if sign else 0, x.ctypes.data, y.ctypes.data)
    return y


def cases(files, seed=0):
    """(name, q, inv, sectors, slot, op, dense, out_key, x) for every slot of every file."""
    out = []
    for path in files:
        q, tb = _load(path)
        inv, sectors = whole_basis(q, tb)
        rng = np.random.default_rng(seed + len(out))
        n_in = len(tb[0])
        x = rng.uniform(-1, 1, n_in)
        for slot in range(4):
            shifts = [(0, 2), (0, -2), (1, 1), (1, -1), (-1, 1), (0, 1), (0, -1)]
            got = shift_operator(q, slot, rng, [s for s in shifts if (q["target"][0] + s[0],
                                                                      q["target"][1] + s[1]) in sectors])
            if got is None:
                continue
            op, dense, _, out_key = got
            out.append((os.path.basename(path)[3:-4], q, inv, sectors, slot, op, dense, out_key, x))
    return out


def _one_per_run(files):
    """The first step of each reference run (xxz, spin-1 xxz, hubbard, klm)."""
    seen, out = set(), []
    for f in files:
        run = os.path.basename(f).rsplit("_s", 1)[0]
        if run not in seen:
            seen.add(run)
            out.append(f)
    return out


SMALL_CASES = cases(_one_per_run(SMALL))


@pytest.mark.parametrize("sign", [False, True])
@pytest.mark.parametrize("case", SMALL_CASES, ids=[f"{c[0]}-{SLOTS[c[4]]}" for c in SMALL_CASES])
def test_oracle_matches_dense_product(case, sign):
    name, q, inv, sectors, slot, op, dense, out_key, x = case
    want = dense_product(q, slot, dense, sectors, out_key, x, sign)
    got = oracle_v_m(q, slot, op, inv, sectors[out_key], x, sign)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * max(1.0, np.abs(want).max()))


def test_sign_rule_planted_mistake_fails():
    """A sign that leaves the LR electrons out of the RR product's parity must differ from
    the dense product somewhere on a fixture with electrons."""
    hub = [c for c in SMALL_CASES if c[1]["n_qn"] == 2 and c[4] == 2]
    assert hub
    name, q, inv, sectors, slot, op, dense, out_key, x = hub[0]
    want = dense_product(q, slot, dense, sectors, out_key, x, True)
    q_bad = dict(q, ele_site=np.zeros_like(q["ele_site"]))   # LR electrons left out of the sign
    got = oracle_v_m(q_bad, slot, op, inv, sectors[out_key], x, True)
    assert not np.allclose(got, want)


def _ref_lib():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libdmrgref.so not built (make -C oracle ref)")
    return C.CDLL(REF_LIB)


class _CRS1(C.Structure):
    _fields_ = [("max_val", C.c_long), ("max_row", C.c_int), ("row_dim", C.c_int), ("col_dim", C.c_int),
                ("Col", C.POINTER(C.c_int)), ("Row", C.POINTER(C.c_long)), ("Val", C.POINTER(C.c_double))]


def _ref_call(q, slot, op, inv, sec, x, sign, qn_depth):
    """The reference's DMRG_V_M_<slot>_Q<qn_depth> on a DMRG_WHOLE_BASIS_Q<qn_depth>
    (include/dmrg.h:140-170) whose output sector sits at index 1 of every level."""
    R = _ref_lib()
    _, LL, LR, RR, RL = sec
    n = LL.size
    keep = []

    def nest(arr, depth):
        """depth levels of pointer arrays of length 2, the data at [1][1]..[1]."""
        if depth == 0:
            return arr.ctypes.data
        inner = nest(arr, depth - 1)
        p = (C.c_void_p * 2)(None, inner)
        keep.append(p)
        return C.addressof(p)

    arrs = [np.ascontiguousarray(a, np.int16) for a in (LL, LR, RL, RR)]
    dim_leaf = np.array([0, n], np.int32)
    keep += arrs + [dim_leaf]
    Wt = type("W", (C.Structure,), {"_fields_": [("LL", C.c_void_p), ("LR", C.c_void_p), ("RL", C.c_void_p),
                                                 ("RR", C.c_void_p), ("Inv", C.c_void_p), ("Dim", C.c_void_p),
                                                 ("max_dim", C.c_int), ("dim_RR", C.c_int),
                                                 ("dim_onsite", C.c_int)]})
    dim_ptr = dim_leaf.ctypes.data
    for _ in range(qn_depth - 1):
        p = (C.c_void_p * 2)(None, dim_ptr)
        keep.append(p)
        dim_ptr = C.addressof(p)
    W = Wt(*[nest(a, qn_depth) for a in arrs], inv.ctypes.data, dim_ptr, n, q["dim_RR"], q["d"])
    row, col, val = op
    M = _CRS1(len(val), len(row) - 1, len(row) - 1, len(row) - 1, col.ctypes.data_as(C.POINTER(C.c_int)),
              row.ctypes.data_as(C.POINTER(C.c_long)), val.ctypes.data_as(C.POINTER(C.c_double)))
    eles = [np.ascontiguousarray(q[k], np.int32) for k in ("ele_LL", "ele_site", "ele_RR")]
    x = np.ascontiguousarray(x)
    y = np.zeros(n)
    fn = getattr(R, f"DMRG_V_M_{SLOTS[slot]}_Q{qn_depth}")
    qn = [C.c_int(1)] * qn_depth
    flag = b"Yes" if sign else b"No"
    if slot == 0:
        fn(C.byref(M), *qn, C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data), C.c_int(1), C.byref(W))
    else:
        e = [C.c_void_p(a.ctypes.data) for a in eles[:slot]]
        fn(C.byref(M), *qn, C.c_void_p(x.ctypes.data), *e, C.c_char_p(flag), C.c_void_p(y.ctypes.data),
           C.c_int(1), C.byref(W))
    return y


@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("sign", [False, True])
@pytest.mark.parametrize("case", SMALL_CASES, ids=[f"{c[0]}-{SLOTS[c[4]]}" for c in SMALL_CASES])
def test_oracle_matches_reference(case, sign, depth):
    name, q, inv, sectors, slot, op, dense, out_key, x = case
    want = _ref_call(q, slot, op, inv, sectors[out_key], x, sign, depth)
    got = oracle_v_m(q, slot, op, inv, sectors[out_key], x, sign)
    np.testing.assert_array_equal(got, want)


def test_argument_checks_without_device():
    import dmrg_c_amd as D
    L = D.lib()
    h = C.c_void_p()
    L.dmrgx_whole_basis_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.POINTER(C.c_void_p)]
    inv = np.zeros(16, np.int32)
    assert L.dmrgx_whole_basis_create(0, 2, 2, inv.ctypes.data, C.byref(h)) != 0
    assert L.dmrgx_whole_basis_create(2, 2, 2, None, C.byref(h)) != 0
    assert L.dmrgx_whole_basis_create(40000, 2, 2, inv.ctypes.data, C.byref(h)) != 0
    bad = np.full(16, -2, np.int32)
    assert L.dmrgx_whole_basis_create(2, 2, 2, bad.ctypes.data, C.byref(h)) != 0
    L.dmrgx_v_m_sector.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32] + [C.c_void_p] * 4 + \
                                  [C.c_int32, C.c_void_p]
    assert L.dmrgx_v_m_sector(None, 0, None, 0, 0, None, None, None, None, 0, None) != 0


GPU_CASES = SMALL_CASES + cases([f for f in QN_FILES if "m400" in f], seed=7)


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=[f"{c[0]}-{SLOTS[c[4]]}" for c in GPU_CASES])
def test_device_matches_oracle(case):
    import dmrg_c_amd as D
    name, q, inv, sectors, slot, op, dense, out_key, x = case
    print(f"v_m_sector {name} slot {SLOTS[slot]} out {out_key} n_in {x.size}", flush=True)
    W = D.WholeBasis(q["dim_LL"], q["dim_RR"], q["d"], inv)
    try:
        ids = {k: W.add_sector(*s[1:]) for k, s in sectors.items()}
        eles = dict(ele_LL=q["ele_LL"], ele_LR=q["ele_site"], ele_RR=q["ele_RR"])
        for sign in (False, True):
            want = oracle_v_m(q, slot, op, inv, sectors[out_key], x, sign)
            got = W.v_m(slot, op, ids[out_key], x, sign=sign, **eles)
            np.testing.assert_array_equal(got, want)
        # the same product again after another sector's product: no state kept between calls
        other = next(k for k in sectors if k != out_key)
        W.v_m(slot, op, ids[other], np.zeros(max(sectors[k][0].size for k in sectors)))
        np.testing.assert_array_equal(W.v_m(slot, op, ids[out_key], x, **eles),
                                      oracle_v_m(q, slot, op, inv, sectors[out_key], x, False))
    finally:
        W.free()


@pytest.mark.gpu
def test_device_rejects_out_of_range_inverse():
    """An operator that reaches a state whose index lies past the input vector is an error,
    detected on the device (the reference would read past V)."""
    import dmrg_c_amd as D
    q, tb = _load(SMALL[0])
    inv, sectors = whole_basis(q, tb)
    big = max(sectors, key=lambda k: sectors[k][0].size)
    W = D.WholeBasis(q["dim_LL"], q["dim_RR"], q["d"], inv)
    try:
        sid = W.add_sector(*sectors[big][1:])
        d = q["d"]
        ident = (np.arange(d + 1, dtype=np.int64), np.arange(d, dtype=np.int32), np.ones(d))
        with pytest.raises(D.DmrgxError, match="outside the input vector"):
            W.v_m(1, ident, sid, np.ones(1))
        ok = W.v_m(1, ident, sid, np.arange(sectors[big][0].size, dtype=np.float64))
        np.testing.assert_array_equal(ok, np.arange(sectors[big][0].size, dtype=np.float64))
    finally:
        W.free()


def _sx_like(q, slot, rng, dsz_abs=2, conserve=False):
    """Random CRS on the slot's block: S^z-conserving (conserve) or changing S^z by +-2."""
    e, s = slot_qn(q, slot)
    diff = s[:, None] - s[None, :]
    mask = (diff == 0) if conserve else (np.abs(diff) == dsz_abs)
    mask &= (e[:, None] == e[None, :]) & (rng.random(mask.shape) < 0.6)
    m = sp.csr_matrix(np.where(mask, rng.uniform(-1, 1, mask.shape), 0.0))
    m.sort_indices()
    return m.indptr.astype(np.int64), m.indices.astype(np.int32), m.data.copy()


@pytest.mark.gpu
def test_device_intersite_correlation_matches_composition():
    """EXPECTATION_INTERSITE_Q2 (main/dmrg/xxz_vf/src/expectations/EXPECTATION_INTERSITE_Q2.c)
    on a three-sector whole basis (target and target -+ 2, other states -1 in the inverse table,
    as GET_WHOLE_BASIS_Q2_SUPERBLOCK builds it) against the same steps composed from the
    oracle's sector products and numpy inner products."""
    import dmrg_c_amd as D
    path = next(f for f in SMALL if "xxz_L12" in f)
    q, tb = _load(path)
    inv_all, sectors = whole_basis(q, tb)
    t = q["target"]
    keys = [t, (t[0], t[1] - 2), (t[0], t[1] + 2)]
    inv = np.full_like(inv_all, -1)
    for k in keys:
        inv[sectors[k][0]] = inv_all[sectors[k][0]]
    rng = np.random.default_rng(21)
    LL_site, RR_site = 3, 2
    x = rng.uniform(-1, 1, sectors[t][0].size)
    W = D.WholeBasis(q["dim_LL"], q["dim_RR"], q["d"], inv)
    try:
        ids = [W.add_sector(*sectors[k][1:]) for k in keys]
        for origin in (0, 2, LL_site + 1, LL_site + 2):
            M_CF = [_sx_like(q, 0, rng, conserve=True) for _ in range(max(LL_site - origin + 1, 0))]
            M_LL = {origin: _sx_like(q, 0, rng)}
            M_On = _sx_like(q, 1, rng)
            M_RR = [_sx_like(q, 2, rng) for _ in range(RR_site + 1)]
            print(f"intersite {os.path.basename(path)} origin {origin}", flush=True)
            got = W.intersite(*ids, M_CF, M_LL, M_On, M_RR, origin, LL_site, RR_site, x)
            if origin > LL_site + 1:
                assert not got.any()
                continue
            ov = lambda slot, op, k, v: oracle_v_m(q, slot, op, inv, sectors[k], v, False)
            first = (1, M_On) if origin == LL_site + 1 else (0, M_LL[origin])
            t1 = [ov(*first, k, x) for k in keys[1:]]
            want = np.zeros(RR_site + LL_site + 4 - origin)
            for site in range(origin, LL_site + 1):
                want[site - origin] = x @ ov(0, M_CF[site - origin], t, x)
            for qq, slot in enumerate((1, 3)):
                want[LL_site + 1 + qq - origin] = sum(t1[i] @ ov(slot, M_On, keys[1 + i], x) for i in range(2))
            for site in range(RR_site, -1, -1):
                want[RR_site + LL_site + 3 - site - origin] = sum(t1[i] @ ov(2, M_RR[site], keys[1 + i], x)
                                                                  for i in range(2))
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    finally:
        W.free()
