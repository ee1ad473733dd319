"""SpMV schedules (dmrgx_matrix_set_schedule / _get_schedule / _autotune).

The load bits of the variant (non-temporal loads, loads in flight per lane) change no
arithmetic, so the solve is bit-identical under them; lanes per row change the summation
order of each row, and the workgroup cap and the row-range bit the order of the fused dot
products, so energies agree with the reference's to the FP64 bound under every schedule."""
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
        for bad in ((0, 0, 0), (3, 0, 0), (128, 0, 0), (8, -1, 0), (8, 0, 8), (8, 0, -1)):
            with pytest.raises(D.DmrgxError):
                A.set_schedule(*bad)
        A.set_schedule(16, 2048, 7)
        assert A.schedule == (16, 2048, 7)
    finally:
        A.free()


def _solves(A, sched, x):
    A.set_schedule(*sched)
    return A.matvec(x), A.ground_state(), A.ground_state(precision=1)


def _same(p, q):
    assert np.array_equal(p[0], q[0])
    for k in (1, 2):
        assert p[k][0] == q[k][0]
        assert np.array_equal(p[k][1], q[k][1])


@pytest.mark.gpu
def test_gpu_load_variants_bit_identical():
    A, e_ref = _matrix()
    try:
        x = np.random.default_rng(3).standard_normal(A.n)
        # capped grid: the load bits give the same bits within each row mapping
        for base in (0, 4):
            ref = _solves(A, (8, 4096, base), x)
            for v in (1, 2, 3):
                _same(ref, _solves(A, (8, 4096, base | v), x))
            assert abs(ref[1][0] - e_ref) <= 1e-10 * abs(e_ref)
        # one workgroup per row group: the row mapping changes nothing either
        ref = _solves(A, (16, 0, 0), x)
        for v in range(1, 8):
            _same(ref, _solves(A, (16, 0, v), x))
    finally:
        A.free()


@pytest.mark.gpu
@pytest.mark.parametrize("sched", [(1, 0, 0), (4, 1024, 5), (16, 0, 2), (32, 8192, 7), (64, 2048, 4), (8, 16384, 6)])
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
        g, cap, v = sched
        assert g & (g - 1) == 0 and 1 <= g <= 64 and cap >= 0 and v in range(8)
        E, _, _ = A.ground_state()
        assert abs(E - e_ref) <= 1e-10 * abs(e_ref)
    finally:
        A.free()


@pytest.mark.gpu
def test_gpu_time_schedule_leaves_schedule():
    A, _ = _matrix()
    try:
        A.set_schedule(8, 2048, 1)
        us = A.time_schedule(32, 0, 6, 5)
        assert us > 0 and A.schedule == (8, 2048, 1)
    finally:
        A.free()
