"""Mixed precision: FP32 CRS values in the Lanczos SpMV, FP64 recurrence scalars, FP64
inverse-iteration/CG refinement against the FP64 matrix, FP64 Rayleigh quotient (precision 1);
or FP32 Lanczos to a loose Ritz residual, then FP64 Lanczos from that Ritz vector to the
residual bound, the FP64 residual test and FP64 inverse iteration if it fails (precision 2).

Acceptance bound for this form: |dE/E| <= 1e-9 against the reference's FP64 energy, and the
refined vector passes the reference's own inverse-iteration test (L1 residual below
ii_acc = 1e-9, measured with the FP64 Rayleigh quotient of the Ritz vector as the shift)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_ids, golden_snapshots

SNAPS = [p for p in golden_snapshots() if "_s0" in p or "_s1" in p]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("path", SNAPS, ids=golden_ids(SNAPS))
def test_gpu_mixed_precision_energy(path, precision):
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(path)
    if s.n < 2:
        pytest.skip("1x1 superblock")
    A = D.DeviceMatrix(s.ham.row, s.ham.col, s.ham.val, sym=True)
    try:
        E, vec, st = A.ground_state(precision=precision)
    finally:
        A.free()
    e_ref = float(s.meta["gs_energy"][0])
    assert abs(E - e_ref) <= 1e-9 * abs(e_ref), (E, e_ref)
    assert st.gs_error < 1e-9, st.gs_error
    assert 0 < st.lanczos_fp32_steps <= st.lanczos_steps
    if "gs_vec" in s.meta:
        ref = s.meta["gs_vec"]
        assert np.abs(vec - np.sign(vec @ ref) * ref).max() <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 2])
def test_gpu_mixed_precision_benchmark_superblock(precision):
    """The n = 144542 Hubbard L=64 m=400 superblock: mixed precision within 1e-9 of the
    reference's FP64 energy."""
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(os.path.join(GOLDEN, "hubbard_L64_m400_s31.npz"))
    sb = D.Superblock(s)
    A = sb.assemble()
    sb.free()
    try:
        E, vec, st = A.ground_state(precision=precision)
    finally:
        A.free()
    e_ref = float(s.meta["gs_energy"][0])
    assert abs(E - e_ref) <= 1e-9 * abs(e_ref), (E, e_ref)
    assert st.gs_error < 1e-9, st.gs_error
    if "gs_vec" in s.meta:
        ref = s.meta["gs_vec"]
        assert np.abs(vec - np.sign(vec @ ref) * ref).max() <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 2])
def test_gpu_mixed_precision_ed(precision):
    """ED (XXZ L=16, n = 12870): mixed precision within 1e-9 of the FP64 solve."""
    from dmrg_c_amd import ed
    P = ed.EdProblem.from_model(ed.Model("xxz", L=16, sz=0, J_z=0.8, h_z=0.1))
    A = P.assemble()
    P.free()
    try:
        e64, _, _ = A.ground_state(precision=0)
        e32, _, st = A.ground_state(precision=precision)
    finally:
        A.free()
    assert abs(e32 - e64) <= 1e-9 * abs(e64)
    assert st.gs_error < 1e-9, st.gs_error
