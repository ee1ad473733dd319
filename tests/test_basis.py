"""Superblock basis construction (GET_BASIS_SUPERBLOCK of each DMRG program,
main/dmrg/<model>_vf/src/allocations/GET_BASIS_SUPERBLOCK.c, with
dmrg/DMRG_QUICK_SORT_BASIS_Q1/Q2.c): the basis arrays are compared bit for bit with the
reference's own (tests/golden/qn_*.npz, dumped from the compiled reference by
oracle/harness/snap.c), for the oracle on the CPU and for the library on the GPU."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN

QN_FILES = sorted(glob.glob(os.path.join(GOLDEN, "qn_*.npz")))


def _qn(path):
    z = np.load(path)
    n_qn, dim_LL, dim_RR, d, t_sz, t_ele = (int(v) for v in z["header"])
    q = dict(n_qn=n_qn, target_sz=t_sz, target_ele=t_ele)
    for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR", "sz_site", "ele_site"):
        q[k] = z[k]
    ref = tuple(z[k] for k in ("LL", "LR", "RR", "RL"))
    return q, ref


def _random_qn(rng, n_qn, dim_LL, dim_RR, d):
    sz_site = rng.integers(-2, 3, d)
    ele_site = rng.integers(0, 3, d)
    return dict(n_qn=n_qn, sz_LL=rng.integers(-6, 7, dim_LL), ele_LL=rng.integers(0, 9, dim_LL),
                sz_RR=rng.integers(-6, 7, dim_RR), ele_RR=rng.integers(0, 9, dim_RR), sz_site=sz_site,
                ele_site=ele_site, target_sz=int(rng.integers(-2, 3)), target_ele=int(rng.integers(4, 12)))


def _lexsort_basis(q):
    """The reference's definition restated with numpy: all states of the target sector, ordered
    by (electrons of LL+LR, S^z of LL+LR, LL, LR, RR, RL)."""
    LL, LR, RR, RL = np.meshgrid(np.arange(len(q["sz_LL"])), np.arange(len(q["sz_site"])),
                                 np.arange(len(q["sz_RR"])), np.arange(len(q["sz_site"])), indexing="ij")
    LL, LR, RR, RL = (a.ravel() for a in (LL, LR, RR, RL))
    szL = q["sz_LL"][LL] + q["sz_site"][LR]
    sz = szL + q["sz_RR"][RR] + q["sz_site"][RL]
    keep = sz == q["target_sz"]
    eleL = np.zeros_like(szL)
    if q["n_qn"] == 2:
        eleL = q["ele_LL"][LL] + q["ele_site"][LR]
        keep &= (eleL + q["ele_RR"][RR] + q["ele_site"][RL]) == q["target_ele"]
    idx = np.flatnonzero(keep)
    order = np.lexsort((RL[idx], RR[idx], LR[idx], LL[idx], szL[idx], eleL[idx]))
    idx = idx[order]
    return tuple(a[idx].astype(np.int16) for a in (LL, LR, RR, RL))


@pytest.mark.parametrize("path", QN_FILES, ids=[os.path.basename(p)[3:-4] for p in QN_FILES])
def test_oracle_basis_matches_reference(path):
    import oracle_py as O
    q, ref = _qn(path)
    got = O.superblock_basis(q)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_basis_matches_lexsort(seed):
    import oracle_py as O
    rng = np.random.default_rng(seed)
    q = _random_qn(rng, 1 + seed % 2, int(rng.integers(1, 30)), int(rng.integers(1, 30)), int(rng.integers(2, 5)))
    got, want = O.superblock_basis(q), _lexsort_basis(q)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_reference_fixtures_satisfy_the_sector():
    """Every state of a reference basis is in the target sector, with no repeats."""
    for path in QN_FILES:
        q, (LL, LR, RR, RL) = _qn(path)
        sz = q["sz_LL"][LL] + q["sz_site"][LR] + q["sz_RR"][RR] + q["sz_site"][RL]
        assert np.all(sz == q["target_sz"])
        key = ((LL.astype(np.int64) * 64 + LR) * 40000 + RR) * 64 + RL
        assert len(np.unique(key)) == len(key)


# (basis fixture, superblock snapshot of the same run and step)
PAIRS = [("qn_hubbard_L8_m32_s03", "hubbard_L8_m32_s03"), ("qn_hubbard_L8_m32_s05", "hubbard_L8_m32_s05"),
         ("qn_hubbard_L8_m32_s07", "hubbard_L8_m32_s07"), ("qn_klm_L6_m32_s03", "klm_L6_m32_s03"),
         ("qn_xxz_L12_m24_s04", "xxz_L12_m24_jz05_s04"), ("qn_xxz_L12_m24_s13", "xxz_L12_m24_jz05_s13")]


@pytest.mark.parametrize("qn_name,snap_name", PAIRS, ids=[p[1] for p in PAIRS])
def test_basis_fixtures_agree_with_snapshots(qn_name, snap_name):
    """The two fixture sets dump the same reference run: their bases must be identical."""
    from dmrg_c_amd import snapshot
    _, ref = _qn(os.path.join(GOLDEN, qn_name + ".npz"))
    s = snapshot.load(os.path.join(GOLDEN, snap_name + ".npz"))
    for a, b in zip(ref, (s.LL, s.LR, s.RR, s.RL)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("path", QN_FILES, ids=[os.path.basename(p)[3:-4] for p in QN_FILES])
def test_basis_size_query(path):
    """The sector dimension (host bookkeeping, no device work) equals the reference's."""
    import ctypes as C
    import dmrg_c_amd as D
    q, ref = _qn(path)
    b, keep = D._basis_qn(q)
    n = C.c_int64()
    D.check(D.lib().dmrgx_superblock_basis(C.byref(b), C.byref(n), None, None, None, None))
    assert n.value == len(ref[0])


def test_basis_argument_checks():
    import ctypes as C
    import dmrg_c_amd as D
    rng = np.random.default_rng(0)
    q = _random_qn(rng, 2, 5, 5, 4)
    for bad in (dict(n_qn=3), dict(sz_LL=np.zeros(0, np.int32))):
        b, keep = D._basis_qn(dict(q, **bad))
        n = C.c_int64()
        assert D.lib().dmrgx_superblock_basis(C.byref(b), C.byref(n), None, None, None, None) != 0


@pytest.mark.gpu
@pytest.mark.parametrize("path", QN_FILES, ids=[os.path.basename(p)[3:-4] for p in QN_FILES])
def test_gpu_basis_matches_reference(path):
    import dmrg_c_amd as D
    q, ref = _qn(path)
    got = D.superblock_basis(q)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n_qn,dims", [(0, 1, (3, 7, 2)), (1, 2, (17, 9, 4)), (2, 2, (64, 64, 8)),
                                            (3, 1, (200, 150, 3)), (4, 2, (396, 396, 4))])
def test_gpu_basis_matches_oracle(seed, n_qn, dims):
    """Random block quantum numbers, up to the Hubbard m=400 block sizes (n ~ 1e5)."""
    import dmrg_c_amd as D
    import oracle_py as O
    q = _random_qn(np.random.default_rng(seed), n_qn, *dims)
    want = O.superblock_basis(q)
    if len(want[0]) == 0:
        pytest.skip("empty sector")
    got = D.superblock_basis(q)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("qn_name,snap_name", PAIRS, ids=[p[1] for p in PAIRS])
def test_gpu_superblock_from_quantum_numbers(qn_name, snap_name):
    """Superblock whose basis is built on the device from the quantum numbers: its assembled
    CRS is the reference's (Row/Col/Val bit-exact), as with the snapshot's basis arrays."""
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    q, _ = _qn(os.path.join(GOLDEN, qn_name + ".npz"))
    s = snapshot.load(os.path.join(GOLDEN, snap_name + ".npz"))
    sb = D.Superblock(s, qn=q)
    try:
        assert sb.n == s.n
        sb.assemble(full=False)
        row, col, val = sb.download_lower()
    finally:
        sb.free()
    assert np.array_equal(row, s.ham.row)
    assert np.array_equal(col, s.ham.col)
    assert np.array_equal(val, s.ham.val)
