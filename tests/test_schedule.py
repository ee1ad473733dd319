"""SpMV schedules (dmrgx_matrix_set_schedule / _get_schedule / _autotune).

The load policy changes no arithmetic, so the solve is bit-identical with and without
non-temporal loads; lanes per row and the workgroup cap change the summation order of each
row, so energies agree with the reference's to the FP64 bound under every schedule."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _matrix():
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    s = snapshot.load(os.path.join(GOLDEN, "hubbard_L64_m400_s31.npz"))
    sb = D.Superblock(s)
    A = sb.assemble()
    sb.free()
    return A, float(s.meta["gs_energy"][0])


@pytest.mark.gpu
def test_gpu_schedule_argument_checks():
    import dmrg_c_amd as D
    A, _ = _matrix()
    try:
        for bad in ((0, 0, 0), (3, 0, 0), (128, 0, 0), (8, -1, 0), (8, 0, 2)):
            with pytest.raises(D.DmrgxError):
                A.set_schedule(*bad)
        A.set_schedule(16, 2048, 1)
        assert A.schedule == (16, 2048, 1)
    finally:
        A.free()


@pytest.mark.gpu
def test_gpu_load_policy_bit_identical():
    A, e_ref = _matrix()
    try:
        x = np.random.default_rng(3).standard_normal(A.n)
        out = {}
        for nt in (0, 1):
            A.set_schedule(8, 4096, nt)
            out[nt] = (A.matvec(x), A.ground_state())
        assert np.array_equal(out[0][0], out[1][0])
        assert out[0][1][0] == out[1][1][0]
        assert np.array_equal(out[0][1][1], out[1][1][1])
    finally:
        A.free()


@pytest.mark.gpu
@pytest.mark.parametrize("sched", [(1, 0, 0), (4, 1024, 1), (16, 0, 0), (32, 8192, 1), (64, 2048, 0)])
def test_gpu_schedules_agree(sched):
    A, e_ref = _matrix()
    try:
        x = np.random.default_rng(5).standard_normal(A.n)
        A.set_schedule(8, 4096, 0)
        y0 = A.matvec(x)
        A.set_schedule(*sched)
        y1 = A.matvec(x)
        assert np.abs(y1 - y0).max() <= 1e-12 * np.abs(y0).max()
        E, _, _ = A.ground_state()
        assert abs(E - e_ref) <= 1e-10 * abs(e_ref)
    finally:
        A.free()


@pytest.mark.gpu
def test_gpu_autotune_sets_schedule():
    A, e_ref = _matrix()
    try:
        sched, us = A.autotune()
        assert A.schedule == sched and us > 0
        g, cap, nt = sched
        assert g & (g - 1) == 0 and 1 <= g <= 64 and cap >= 0 and nt in (0, 1)
        E, _, _ = A.ground_state()
        assert abs(E - e_ref) <= 1e-10 * abs(e_ref)
    finally:
        A.free()
