"""Truncation: DMRG_GET_SYSTEM_INFO_Q1/Q2 (dmrg/DMRG_GET_SYSTEM_INFO_Q2.c:7-236) -- density
matrices of the LLLR blocks, their eigendecompositions and the kept states with the
transformation matrix.

Inputs: the reference's own target-sector bases (tests/golden/qn_*.npz) with the LLLR
sub-basis by first appearance (tests/test_density.py) and a seeded unit vector; the LLLR
quantum numbers are those the programs pass (hubbard/klm: Tot_Ele_LLLR, Tot_Sz_LLLR; xxz:
Tot_Sz_LLLR; main/dmrg/<model>_vf/src/main/RENORMALIZE.c).

Pins of the oracle's selection (or_system_select, fed numpy's eigh of the oracle's density
matrices): the reference's own DMRG_GET_SYSTEM_INFO_Q1/Q2 compiled from its sources with
LAPACK (oracle/_ref/libdmrgref.so, skipped when not built).  The library is then compared
with the oracle on the GPU.  Eigenvectors are compared through what does not depend on
their sign (or on a rotation within a degenerate set): per quantum-number block the
projector onto the kept columns, each column's Rayleigh quotient against its eigenvalue,
and orthonormality."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_py
from conftest import REPO
from test_density import QN_FILES, SUBS, Sub, oracle_block

REF_LIB = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")


def qns(S):
    return (S.q_ele, S.q_sz) if S.n_qn == 2 else (S.q_sz, None)


def oracle_info(S, max_dim):
    """(dict like dmrg_c_amd.system_info_dict, block matrices)."""
    q1, q2 = qns(S)
    keys = np.stack([q1, q2 if q2 is not None else np.zeros_like(q1)], 1)
    uniq, counts = np.unique(keys, axis=0, return_counts=True)   # ascending (q1, q2), as the reference
    dim_block = counts.astype(np.int32)
    assert np.array_equal(dim_block, S.dim_block)                # blocks are the runs of equal numbers
    rho = [oracle_block(S, b) for b in range(len(dim_block))]
    eig = [np.linalg.eigh(m) for m in rho]
    w = np.concatenate([e[0] for e in eig])
    L = oracle_py.lib()
    fn = L.or_system_select
    fn.restype = C.c_int32
    fn.argtypes = [C.c_int32] + [C.c_void_p] * 4 + [C.c_int32] + [C.c_void_p] * 4
    qb1 = np.ascontiguousarray(uniq[:, 0], np.int32)
    qb2 = np.ascontiguousarray(uniq[:, 1], np.int32)
    m = int(dim_block.sum())
    val, sums = np.empty(m), np.empty(2)
    kb, ki = np.empty(m, np.int32), np.empty(m, np.int32)
    k = fn(len(dim_block), dim_block.ctypes.data, qb1.ctypes.data, qb2.ctypes.data, w.ctypes.data, max_dim,
           val.ctypes.data, sums.ctypes.data, kb.ctypes.data, ki.ctypes.data)
    assert k > 0
    base = np.r_[0, np.cumsum(dim_block)]
    T = np.zeros((m, k))
    for c in range(k):
        T[base[kb[c]]:base[kb[c] + 1], c] = eig[kb[c]][1][:, ki[c]]
    return ({"dim_renorm": k, "val_dist": val, "sum_val_dm": sums[0], "tr_error": sums[1], "q1": qb1[kb[:k]],
             "q2": qb2[kb[:k]] if S.n_qn == 2 else None, "T": T}, rho, base)


def assert_same_info(got, want, rho, base, tol_vec=1e-9):
    assert got["dim_renorm"] == want["dim_renorm"]
    np.testing.assert_allclose(got["val_dist"], want["val_dist"], rtol=0, atol=1e-14)
    assert abs(got["sum_val_dm"] - want["sum_val_dm"]) < 1e-13
    assert abs(got["tr_error"] - want["tr_error"]) < 1e-13
    np.testing.assert_array_equal(got["q1"], want["q1"])
    if want["q2"] is not None:
        np.testing.assert_array_equal(got["q2"], want["q2"])
    T, Tw = got["T"], want["T"]
    k = T.shape[1]
    np.testing.assert_allclose(T.T @ T, np.eye(k), atol=1e-12)
    for b in range(len(base) - 1):
        sl = slice(base[b], base[b + 1])
        np.testing.assert_allclose(T[sl] @ T[sl].T, Tw[sl] @ Tw[sl].T, atol=tol_vec)
    kept_vals = np.sort(want["val_dist"][:k])[::-1]
    ray = np.array([T[base[b]:base[b + 1], c] @ rho[b] @ T[base[b]:base[b + 1], c]
                    for c in range(k) for b in [int(np.searchsorted(base, np.nonzero(T[:, c])[0][0], "right") - 1)]])
    np.testing.assert_allclose(np.sort(ray)[::-1], kept_vals, atol=1e-13)


def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libdmrgref.so not built (make -C oracle ref)")
    import dmrg_c_amd as D
    R = C.CDLL(REF_LIB)
    P = C.POINTER(D.DMRG_SYSTEM_INFO)
    R.DMRG_GET_SYSTEM_INFO_Q1.restype = P
    R.DMRG_GET_SYSTEM_INFO_Q1.argtypes = [C.c_void_p, C.POINTER(D.DMRG_BASIS), C.c_void_p, C.c_int, C.c_int,
                                          C.POINTER(D.DMRG_TIME)]
    R.DMRG_GET_SYSTEM_INFO_Q2.restype = P
    R.DMRG_GET_SYSTEM_INFO_Q2.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(D.DMRG_BASIS), C.c_void_p, C.c_int,
                                          C.c_int, C.POINTER(D.DMRG_TIME)]
    R.DMRG_FREE_SYSTEM_INFO.argtypes = [P]
    return R, D


CASES = [(S, md) for S in SUBS for md in (max(1, S.dim_LLLR // 3), S.dim_LLLR + 5)]


@pytest.mark.parametrize("case", CASES, ids=[f"{S.name}-max{md}" for S, md in CASES])
def test_oracle_matches_reference(case):
    S, max_dim = case
    R, D = _ref()
    B = S.basis_struct(D)
    t = D.DMRG_TIME()
    q1, q2 = qns(S)
    if S.n_qn == 2:
        p = R.DMRG_GET_SYSTEM_INFO_Q2(q1.ctypes.data, q2.ctypes.data, C.byref(B), S.vec.ctypes.data, max_dim, 1,
                                      C.byref(t))
    else:
        p = R.DMRG_GET_SYSTEM_INFO_Q1(q1.ctypes.data, C.byref(B), S.vec.ctypes.data, max_dim, 1, C.byref(t))
    try:
        want = D.system_info_dict(p.contents, S.n_qn)
    finally:
        R.DMRG_FREE_SYSTEM_INFO(p)
    assert want["T_dagger_is_transpose"]
    got, rho, base = oracle_info(S, max_dim)
    assert_same_info(got, want, rho, base)


def test_selection_planted_mistake_fails():
    """An oracle cut that ignores the degeneracy rule (cut exactly at max_dim) must change
    dim_renorm on a vector with two equal blocks."""
    L = oracle_py.lib()
    fn = L.or_system_select
    fn.restype = C.c_int32
    fn.argtypes = [C.c_int32] + [C.c_void_p] * 4 + [C.c_int32] + [C.c_void_p] * 4
    dim_block = np.array([2, 2], np.int32)
    q = np.array([0, 1], np.int32)
    w = np.array([0.1, 0.4, 0.1, 0.4])   # block eigenvalues, ascending, equal across the blocks
    val, sums, kb, ki = np.empty(4), np.empty(2), np.empty(4, np.int32), np.empty(4, np.int32)
    k = fn(2, dim_block.ctypes.data, q.ctypes.data, q.ctypes.data, w.ctypes.data, 1, val.ctypes.data,
           sums.ctypes.data, kb.ctypes.data, ki.ctypes.data)
    assert k != 1   # the cut may not split the degenerate pair 0.4, 0.4
    assert k == -1  # and no smaller cut exists: the reference stops with an error
    k3 = fn(2, dim_block.ctypes.data, q.ctypes.data, q.ctypes.data, w.ctypes.data, 3, val.ctypes.data,
            sums.ctypes.data, kb.ctypes.data, ki.ctypes.data)
    assert k3 == 2 and abs(sums[1] - 0.2) < 1e-15
    np.testing.assert_array_equal(kb[:2], [0, 1])   # ties keep block order


GPU_CASES = CASES + [(S, md) for S in [Sub(f, 9) for f in QN_FILES if "m400" in f] for md in (400, 1000)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=[f"{S.name}-max{md}" for S, md in GPU_CASES])
def test_device_matches_oracle(case):
    import dmrg_c_amd as D
    S, max_dim = case
    print(f"system info {S.name} dim_LLLR={S.dim_LLLR} max_dim={max_dim}", flush=True)
    B = S.basis_struct(D)
    q1, q2 = qns(S)
    got = D.system_info(B, S.n_qn, q1, q2, S.vec, max_dim)
    assert got["T_dagger_is_transpose"]
    assert got["shapes"] == (S.dim_LLLR, got["dim_renorm"], got["dim_renorm"], S.dim_LLLR)
    want, rho, base = oracle_info(S, max_dim)
    assert_same_info(got, want, rho, base)


BIG_CASES = [(S, md) for S in [Sub(f, 9) for f in QN_FILES if "m400" in f] for md in (400, 1000)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", BIG_CASES, ids=[f"{S.name}-max{md}" for S, md in BIG_CASES])
def test_device_library_solver_matches_oracle(case):
    """The same selection through rocSOLVER dsyevd for every block (mode 1), and the own
    solver's result (mode 0) against it: equal cut, eigenvalues within 1e-14, equal kept
    subspaces per block."""
    import dmrg_c_amd as D
    S, max_dim = case
    print(f"system info, rocSOLVER and own solver, {S.name} max_dim={max_dim}", flush=True)
    B = S.basis_struct(D)
    q1, q2 = qns(S)
    D.set_block_eigensolver(1)
    try:
        lib_info = D.system_info(B, S.n_qn, q1, q2, S.vec, max_dim)
    finally:
        D.set_block_eigensolver(0)
    own = D.system_info(B, S.n_qn, q1, q2, S.vec, max_dim)
    want, rho, base = oracle_info(S, max_dim)
    assert_same_info(lib_info, want, rho, base)
    assert_same_info(own, lib_info, rho, base)


M800 = [f for f in QN_FILES if "m800" in f]


@pytest.mark.gpu
@pytest.mark.parametrize("path", M800, ids=[os.path.basename(f)[3:-4] for f in M800])
def test_device_mixed_solvers_m800(path):
    """The KLM m=800 step: 61 blocks, 13 of them above 192 states (rocSOLVER) and 48 at or
    below (the own solver) in one call, against rocSOLVER for every block: equal cut and
    quantum numbers, eigenvalues within 1e-14, equal kept subspaces per block, orthonormal
    columns."""
    import dmrg_c_amd as D
    S = Sub(path, 4)
    assert (S.dim_block > 192).any() and (S.dim_block <= 192).any()
    print(f"system info {S.name} dim_LLLR={S.dim_LLLR} blocks={len(S.dim_block)} max_dim=800", flush=True)
    B = S.basis_struct(D)
    q1, q2 = qns(S)
    D.set_block_eigensolver(1)
    try:
        want = D.system_info(B, S.n_qn, q1, q2, S.vec, 800)
    finally:
        D.set_block_eigensolver(0)
    got = D.system_info(B, S.n_qn, q1, q2, S.vec, 800)
    assert got["dim_renorm"] == want["dim_renorm"]
    np.testing.assert_allclose(got["val_dist"], want["val_dist"], rtol=0, atol=1e-14)
    assert abs(got["tr_error"] - want["tr_error"]) < 1e-13
    np.testing.assert_array_equal(got["q1"], want["q1"])
    np.testing.assert_array_equal(got["q2"], want["q2"])
    T, Tw = got["T"], want["T"]
    k = T.shape[1]
    np.testing.assert_allclose(T.T @ T, np.eye(k), atol=1e-12)
    base = np.r_[0, np.cumsum(S.dim_block)]
    for b in range(len(base) - 1):
        sl = slice(base[b], base[b + 1])
        np.testing.assert_allclose(T[sl] @ T[sl].T, Tw[sl] @ Tw[sl].T, atol=1e-9)


def test_block_eigensolver_mode_checked():
    import dmrg_c_amd as D
    L = D.lib()
    L.dmrgx_set_block_eigensolver.argtypes = [C.c_int]
    assert L.dmrgx_set_block_eigensolver(2) == -1   # DMRGX_ERR_ARG
    assert L.dmrgx_set_block_eigensolver(-1) == -1
    assert L.dmrgx_set_block_eigensolver(0) == 0


@pytest.mark.gpu
def test_device_order_error():
    import dmrg_c_amd as D
    S = SUBS[0]
    B = S.basis_struct(D)
    q1, q2 = qns(S)
    with pytest.raises(D.DmrgxError, match="out of order"):
        D.system_info(B, S.n_qn, q1[::-1].copy(), None if q2 is None else q2[::-1].copy(), S.vec, 10)


@pytest.mark.gpu
def test_device_blocks_not_contiguous_error():
    """(0,0)... (1,1)... (0,0)...: every neighbouring pair is in order, but the (0,0) states
    form two runs, so the run count (3) differs from the distinct pairs (2); the reference
    exits at DMRG_GET_DENSITY_MATRIX_INFO_Q2.c:69-74, the library returns an argument error."""
    import dmrg_c_amd as D
    S = next(s for s in SUBS if s.n_qn == 2)
    B = S.basis_struct(D)
    m = S.dim_LLLR
    assert m >= 3
    q1 = np.zeros(m, np.int32)
    q2 = np.zeros(m, np.int32)
    q1[m // 3:2 * m // 3] = 1
    q2[m // 3:2 * m // 3] = 1
    with pytest.raises(D.DmrgxError, match="not contiguous"):
        D.system_info(B, 2, q1, q2, S.vec, 10)
