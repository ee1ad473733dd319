"""Kronecker form of the superblock Hamiltonian (dmrgx_superblock_kron): per sector of the
superblock basis, the LL/LR/LLLR terms as a dense matrix on the rows of the sector's block of
the vector and the RR/RL/RRRL terms as one on its columns, the coupling terms row by row.

The operator is the one GET_HAM_LLLRRRRL assembles (main/dmrg/<model>_vf/src/allocations/
GET_HAM_LLLRRRRL.c:11-114), so its product must equal the product with the reference's own
CRS (expanded to full) to rounding, on every golden superblock of the three models, and its
ground state must be the reference's within the FP64 bounds."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_ids, golden_snapshots

SNAPS = golden_snapshots()
BIG = [os.path.join(GOLDEN, f) for f in ("hubbard_L64_m400_s31.npz", "xxz_L64_m300_s31.npz")]


def _ref_full(s):
    import scipy.sparse as sp
    L = sp.csr_matrix((s.ham.val, s.ham.col, s.ham.row), shape=(s.n, s.n))
    return (L + L.T - sp.diags(L.diagonal())).tocsr()


@pytest.mark.gpu
@pytest.mark.parametrize("variant", range(16), ids=["vec64", "mfma64", "mfma128", "mfma64x128", "mfma128x64", "mfma64k32",
                                                     "mfma64o4", "mfma128o2", "vec64-xcd", "mfma64-xcd", "mfma128-xcd",
                                                     "mfma64x128-xcd", "mfma128x64-xcd", "mfma64k32-xcd",
                                                     "mfma64o4-xcd", "mfma128o2-xcd"])
@pytest.mark.parametrize("path", SNAPS + BIG, ids=golden_ids(SNAPS + BIG))
def test_gpu_kron_product_equals_reference_crs(path, variant):
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(path)
    sb = D.Superblock(s)
    K = sb.kron()
    K.set_kernel(variant)
    try:
        x = np.random.default_rng(11).standard_normal(s.n)
        y = K.matvec(x)
        if s.ham is not None:
            y_ref = _ref_full(s) @ x
        else:   # no dumped CRS: the device assembly (bit-exact with the reference's)
            A = sb.assemble()
            y_ref = A.matvec(x)
            A.free()
        scale = np.abs(y_ref).max()
        assert np.abs(y - y_ref).max() <= 1e-12 * scale, np.abs(y - y_ref).max() / scale
    finally:
        K.free()
        sb.free()


@pytest.mark.gpu
@pytest.mark.parametrize("path", SNAPS + BIG[:1], ids=golden_ids(SNAPS + BIG[:1]))
def test_gpu_kron_ground_state(path):
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(path)
    sb = D.Superblock(s)
    K = sb.kron()
    try:
        E, vec, st = K.ground_state()
    finally:
        K.free()
        sb.free()
    e_ref = float(s.meta["gs_energy"][0])
    assert abs(E - e_ref) <= 1e-10 * max(1.0, abs(e_ref)), (E, e_ref)
    assert st.gs_error < 1e-9
    if "gs_vec" in s.meta:
        ref = s.meta["gs_vec"]
        assert np.abs(vec - np.sign(vec @ ref) * ref).max() <= 1e-8


@pytest.mark.gpu
def test_gpu_kron_autotune_and_bad_variant():
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(BIG[0])
    sb = D.Superblock(s)
    K = sb.kron()
    try:
        v, us = K.autotune()
        assert 0 <= v < K.VARIANTS and us > 0
        for bad in (-1, K.VARIANTS):
            with pytest.raises(D.DmrgxError):
                K.set_kernel(bad)
        E, _, _ = K.ground_state()
        e_ref = float(s.meta["gs_energy"][0])
        assert abs(E - e_ref) <= 1e-10 * abs(e_ref)
    finally:
        K.free()
        sb.free()


@pytest.mark.gpu
def test_gpu_kron_rejects_fp32():
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(SNAPS[-1])
    sb = D.Superblock(s)
    K = sb.kron()
    try:
        with pytest.raises(D.DmrgxError):
            K.ground_state(precision=1)
    finally:
        K.free()
        sb.free()


@pytest.mark.gpu
@pytest.mark.parametrize("path", SNAPS[-3:] + BIG[:1], ids=golden_ids(SNAPS[-3:] + BIG[:1]))
def test_gpu_residual_stop(path):
    """lan_type 2 ("Residual", no reference counterpart): Lanczos until the Ritz pair's residual
    bound is below ii_acc / (2 sqrt(n)); the reference's L1 residual test then decides.  The
    energy within the FP64 bound of the reference's, on the Kronecker form and the CRS."""
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(path)
    e_ref = float(s.meta["gs_energy"][0])
    sb = D.Superblock(s)
    K = sb.kron()
    A = sb.assemble()
    try:
        for solver in (K, A):
            E, vec, st = solver.ground_state(lan_type=2)
            assert abs(E - e_ref) <= 1e-10 * max(1.0, abs(e_ref)), (E, e_ref)
            assert st.gs_error < 1e-9
            _, _, st0 = solver.ground_state()
            assert st.spmv_calls <= st0.spmv_calls + 2
        with pytest.raises(D.DmrgxError):
            A.ground_state(lan_type=3)
    finally:
        K.free()
        A.free()
        sb.free()
