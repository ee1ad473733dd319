"""Reduced density matrices of the left half (DMRG_GET_DENSITY_MATRIX,
dmrg/DMRG_GET_DENSITY_MATRIX.c:5-107), the first step of the truncation
(DMRG_GET_SYSTEM_INFO_Q1..Q3).

The sub-basis comes from the reference's own target-sector bases (tests/golden/qn_*.npz):
LLLR states by first appearance with their partner counts (as the programs'
GET_BASIS_SUPERBLOCK fills LL_LLLR / LR_LLLR / Sum_LLLR), blocks = runs of LLLR states with
equal (electrons, S^z).  Pins of the oracle: the reference's own DMRG_GET_DENSITY_MATRIX
compiled from its sources (oracle/_ref/libdmrgref.so, skipped when not built) and numpy's
Psi Psi^T with the trace equal to the block's share of |Vec|^2; a planted dropped term
must fail.  The library is compared bit for bit with the oracle on the GPU."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

import oracle_py
from conftest import GOLDEN, REPO

QN_FILES = sorted(glob.glob(os.path.join(GOLDEN, "qn_*.npz")))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")


def _one_per_run(files):
    seen, out = set(), []
    for f in files:
        run = os.path.basename(f).rsplit("_s", 1)[0]
        if run not in seen:
            seen.add(run)
            out.append(f)
    return out


class Sub:
    """Target-sector basis of a reference run with its LLLR sub-basis and blocks."""

    def __init__(self, path, seed=0):
        z = np.load(path)
        n_qn, self.dim_LL, self.dim_RR, self.d, _, _ = (int(v) for v in z["header"])
        self.LL, self.LR, self.RR, self.RL = (np.ascontiguousarray(z[k], np.int16) for k in ("LL", "LR", "RR", "RL"))
        self.n = self.LL.size
        dims = (self.dim_LL, self.d, self.dim_RR, self.d)
        flat = np.ravel_multi_index(tuple(a.astype(np.int64) for a in (self.LL, self.LR, self.RR, self.RL)), dims)
        self.inv = np.full(int(np.prod(dims)), -1, np.int32)
        self.inv[flat] = np.arange(self.n, dtype=np.int32)
        key = self.LL.astype(np.int64) * self.d + self.LR
        first = np.r_[True, key[1:] != key[:-1]]
        self.LL_LLLR = np.ascontiguousarray(self.LL[first])
        self.LR_LLLR = np.ascontiguousarray(self.LR[first])
        self.Sum_LLLR = np.diff(np.r_[np.nonzero(first)[0], self.n]).astype(np.int32)
        self.dim_LLLR = self.LL_LLLR.size
        ele = z["ele_LL"][self.LL_LLLR] + z["ele_site"][self.LR_LLLR] if n_qn == 2 else np.zeros(self.dim_LLLR)
        sz = z["sz_LL"][self.LL_LLLR] + z["sz_site"][self.LR_LLLR]
        self.n_qn = n_qn
        self.q_ele, self.q_sz = np.ascontiguousarray(ele, np.int32), np.ascontiguousarray(sz, np.int32)
        q = np.stack([ele, sz], 1)
        starts = np.r_[0, np.nonzero(np.any(q[1:] != q[:-1], axis=1))[0] + 1, self.dim_LLLR]
        self.dim_block = np.diff(starts).astype(np.int32)
        self.block_start = starts[:-1]
        rng = np.random.default_rng(seed)
        v = rng.standard_normal(self.n)
        self.vec = v / np.linalg.norm(v)
        self.name = os.path.basename(path)[3:-4]

    def basis_struct(self, D):
        B = D.DMRG_BASIS()
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        B.LL_LLLRRRRL, B.LR_LLLRRRRL = P(self.LL, C.c_short), P(self.LR, C.c_short)
        B.RL_LLLRRRRL, B.RR_LLLRRRRL = P(self.RL, C.c_short), P(self.RR, C.c_short)
        B.Inv_LLLRRRRL = P(self.inv, C.c_int)
        B.dim_LLLRRRRL, B.dim_LL, B.dim_RR, B.dim_onsite = self.n, self.dim_LL, self.dim_RR, self.d
        B.LL_LLLR, B.LR_LLLR = P(self.LL_LLLR, C.c_short), P(self.LR_LLLR, C.c_short)
        B.dim_LLLR = self.dim_LLLR
        B.Sum_LLLR = P(self.Sum_LLLR, C.c_int)
        return B

    def psi(self, b):
        s0, db = self.block_start[b], self.dim_block[b]
        base = np.r_[0, np.cumsum(self.Sum_LLLR)]
        k = self.Sum_LLLR[s0]
        return np.stack([self.vec[base[s0 + r]:base[s0 + r] + k] for r in range(db)])


def oracle_block(S, b, vec=None):
    L = oracle_py.lib()
    fn = L.or_density_matrix
    fn.restype = C.c_int
    fn.argtypes = [C.c_int32] * 4 + [C.c_void_p] * 7 + [C.c_int32, C.c_int32, C.c_void_p]
    db = int(S.dim_block[b])
    out = np.empty((db, db))
    v = np.ascontiguousarray(S.vec if vec is None else vec)
    rc = fn(S.n, S.dim_RR, S.d, S.dim_LLLR, S.LL_LLLR.ctypes.data, S.LR_LLLR.ctypes.data, S.Sum_LLLR.ctypes.data,
            S.RR.ctypes.data, S.RL.ctypes.data, S.inv.ctypes.data, v.ctypes.data, int(S.block_start[b]), db,
            out.ctypes.data)
    assert rc == 0
    return out


SUBS = [Sub(f, i) for i, f in enumerate(_one_per_run([f for f in QN_FILES if "m400" not in f and "m800" not in f]))]


@pytest.mark.parametrize("S", SUBS, ids=[s.name for s in SUBS])
def test_oracle_matches_numpy_and_trace(S):
    tot = 0.0
    for b in range(len(S.dim_block)):
        got = oracle_block(S, b)
        P = S.psi(b)
        np.testing.assert_allclose(got, P @ P.T, rtol=1e-13, atol=1e-16)
        tot += np.trace(got)
    assert abs(tot - 1.0) < 1e-13   # |Vec| = 1, every state in exactly one block


def test_dropped_term_fails():
    """An oracle that loses the last partner of every block must fail the numpy check."""
    S = SUBS[0]
    b = int(np.argmax(S.Sum_LLLR[S.block_start] * (S.dim_block > 1)))
    P = S.psi(b)
    short = S.vec.copy()
    base = np.r_[0, np.cumsum(S.Sum_LLLR)]
    for r in range(S.dim_block[b]):
        short[base[S.block_start[b] + r + 1] - 1] = 0.0
    assert not np.allclose(oracle_block(S, b, short), P @ P.T, rtol=1e-13, atol=1e-16)


def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libdmrgref.so not built (make -C oracle ref)")
    R = C.CDLL(REF_LIB)
    import dmrg_c_amd as D
    R.DMRG_GET_DENSITY_MATRIX.restype = C.POINTER(C.POINTER(C.c_double))
    R.DMRG_GET_DENSITY_MATRIX.argtypes = [C.c_int, C.c_void_p, C.POINTER(D.DMRG_BASIS), C.c_void_p, C.c_int]
    R.FREE_ARRAY_DOUBLE2.argtypes = [C.POINTER(C.POINTER(C.c_double)), C.c_long]
    return R, D


@pytest.mark.parametrize("S", SUBS, ids=[s.name for s in SUBS])
def test_oracle_matches_reference(S):
    R, D = _ref()
    B = S.basis_struct(D)
    for b in range(len(S.dim_block)):
        db = int(S.dim_block[b])
        M = R.DMRG_GET_DENSITY_MATRIX(b, S.dim_block.ctypes.data, C.byref(B), S.vec.ctypes.data, 1)
        want = np.array([[M[r][c] for c in range(db)] for r in range(db)])
        R.FREE_ARRAY_DOUBLE2(M, db)
        np.testing.assert_array_equal(oracle_block(S, b), want)


def test_size_query_and_argument_checks_without_device():
    import dmrg_c_amd as D
    L = D.lib()
    L.dmrgx_density_matrices.argtypes = [C.POINTER(D.DMRG_BASIS), C.c_int32, C.c_void_p, C.c_void_p,
                                         C.POINTER(C.c_int64), C.c_void_p]
    S = SUBS[0]
    B = S.basis_struct(D)
    n = C.c_int64()
    assert L.dmrgx_density_matrices(C.byref(B), len(S.dim_block), S.dim_block.ctypes.data, None, C.byref(n),
                                    None) == 0
    assert n.value == int(np.sum(S.dim_block.astype(np.int64) ** 2))
    bad = np.array([S.dim_LLLR + 1], np.int32)
    assert L.dmrgx_density_matrices(C.byref(B), 1, bad.ctypes.data, None, C.byref(n), None) != 0
    assert L.dmrgx_density_matrices(C.byref(B), 0, S.dim_block.ctypes.data, None, C.byref(n), None) != 0


GPU_SUBS = SUBS + [Sub(f, 9) for f in QN_FILES if "m400" in f]


@pytest.mark.gpu
@pytest.mark.parametrize("S", GPU_SUBS, ids=[s.name for s in GPU_SUBS])
def test_device_matches_oracle(S):
    import dmrg_c_amd as D
    print(f"density {S.name} n={S.n} blocks={len(S.dim_block)} max block={S.dim_block.max()}", flush=True)
    B = S.basis_struct(D)
    mats = D.density_matrices(B, S.dim_block, S.vec)
    for b, M in enumerate(mats):
        np.testing.assert_array_equal(M, oracle_block(S, b))


@pytest.mark.gpu
def test_drop_in_single_block_and_errors():
    import dmrg_c_amd as D
    L = D.lib()
    L.DMRG_GET_DENSITY_MATRIX.restype = C.POINTER(C.POINTER(C.c_double))
    L.DMRG_GET_DENSITY_MATRIX.argtypes = [C.c_int, C.c_void_p, C.POINTER(D.DMRG_BASIS), C.c_void_p, C.c_int]
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    S = SUBS[2]
    B = S.basis_struct(D)
    b = int(np.argmax(S.dim_block))
    db = int(S.dim_block[b])
    M = L.DMRG_GET_DENSITY_MATRIX(b, S.dim_block.ctypes.data, C.byref(B), S.vec.ctypes.data, 1)
    got = np.array([[M[r][c] for c in range(db)] for r in range(db)])
    for r in range(db):
        libc.free(C.cast(M[r], C.c_void_p))
    libc.free(C.cast(M, C.c_void_p))
    np.testing.assert_array_equal(got, oracle_block(S, b))
    # partners that differ within a block are the reference's error, here a status code
    merged = S.dim_block.copy()
    if len(merged) > 1:
        merged = np.r_[merged[0] + merged[1], merged[2:]].astype(np.int32)
        with pytest.raises(D.DmrgxError):
            D.density_matrices(B, merged, S.vec)
