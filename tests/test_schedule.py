"""SpMV schedules (dmrgx_matrix_set_schedule / _get_schedule / _autotune).

Entries in flight per lane and the storage format (CSR, value dictionary, 16-bit column
differences) change no arithmetic, so the solve is bit-identical under them; lanes per row
and paired-entry loads change the summation order of each row, and the workgroup cap and
the row-range bit the order of the fused dot products, so energies agree with the
reference's to the FP64 bound under every schedule.  The sliced form (format 3, 1, 2 or 4
lanes per row) sums each row in an order of its own: its products agree with the CSR ones
to rounding, and its solves are the same bits single and batched."""
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
        for bad in ((0, 0, 0), (3, 0, 0), (128, 0, 0), (8, -1, 0), (8, 0, 24), (8, 0, 32), (2, 0, 40), (8, 0, -1),
                    (8, 0, 64), (2, 0, 64 | 33), (2, 0, 64 | 40), (2, 0, 128)):
            with pytest.raises(D.DmrgxError):
                A.set_schedule(*bad)
        A.set_schedule(16, 2048, 23)
        assert A.schedule == (16, 2048, 23)
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
            for v in (2, 8, 9, 10, 11, 16, 17, 18, 19):   # bit 0 is ignored outside CSR
                _same(ref, _solves(A, (8, 4096, base | v), x))
            pair = _solves(A, (8, 4096, base | 1), x)        # paired loads: own order
            _same(pair, _solves(A, (8, 4096, base | 3), x))
            assert abs(pair[1][0] - e_ref) <= 1e-10 * abs(e_ref)
            assert abs(ref[1][0] - e_ref) <= 1e-10 * abs(e_ref)
        # one workgroup per row group: the row mapping changes nothing either
        ref = _solves(A, (16, 0, 0), x)
        for v in range(1, 24):
            if v & 1 and v < 8:
                continue
            _same(ref, _solves(A, (16, 0, v), x))
    finally:
        A.free()


@pytest.mark.gpu
def test_gpu_sliced_loads_ahead_bit_identical():
    """Bit 6 (matrix loads one pass group ahead of the x gathers) changes no arithmetic:
    the sliced form with 16-bit columns gives the same product and solves with and
    without it, plain or non-temporal, with any lanes per row and row mapping."""
    A, e_ref = _matrix()
    try:
        x = np.random.default_rng(4).standard_normal(A.n)
        for g in (1, 2, 4):
            for v in (25, 27, 29, 31):
                for cap in (4096, 0):
                    ref = _solves(A, (g, cap, v), x)
                    _same(ref, _solves(A, (g, cap, v | 64), x))
            assert abs(ref[1][0] - e_ref) <= 1e-10 * abs(e_ref)
    finally:
        A.free()


@pytest.mark.gpu
@pytest.mark.parametrize("sched", [(1, 0, 0), (4, 1024, 5), (16, 0, 10), (32, 8192, 15), (64, 2048, 12), (8, 16384, 6), (2, 0, 18), (8, 2048, 23),
                                   (1, 0, 24), (2, 4096, 24), (4, 8192, 28), (2, 16384, 31)])
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
        assert g & (g - 1) == 0 and 1 <= g <= 64 and cap >= 0 and (v & ~64) in range(40)
        assert ((v >> 3) & 7) < 3 or g <= 4
        assert not v & 64 or ((v >> 3) & 7) == 3   # loads ahead: sliced form only
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


@pytest.mark.gpu
def test_gpu_value_dictionary():
    """The benchmark superblock's 26.8 M entries hold ~78 k distinct values: the dictionary
    has exactly the distinct values of the matrix, and its products are the plain products
    bit for bit."""
    import ctypes as C
    import dmrg_c_amd as D
    A, _ = _matrix()
    try:
        L = D.lib()
        L.dmrgx_matrix_dict_values.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        nv = C.c_int64()
        D.check(L.dmrgx_matrix_dict_values(A.handle, C.byref(nv)))
        _, _, val = A.download()
        assert nv.value == np.unique(val.view(np.uint64)).size
        x = np.random.default_rng(11).standard_normal(A.n)
        A.set_schedule(16, 8192, 2)
        y0 = A.matvec(x)
        A.set_schedule(16, 8192, 10)
        assert np.array_equal(A.matvec(x), y0)
    finally:
        A.free()


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [8, 16], ids=["dictionary", "delta"])
def test_gpu_formats_ed_and_engine(fmt):
    """ED matrices through the value dictionary (a handful of distinct values) and the delta
    form, single and batched, bit-identical to the plain form."""
    from dmrg_c_amd import ed, engine
    mats = []
    for m in (ed.Model("xxz", L=16, sz=0, J_z=0.7, h_z=0.2), ed.Model("hubbard", L=8, N=8, sz=0, U=3.0)):
        P = ed.EdProblem.from_model(m)
        mats.append(P.assemble())
        P.free()
    try:
        plain = []
        for A in mats:
            g, cap, _ = A.schedule
            A.set_schedule(g, cap, 0)
            plain.append(A.ground_state())
            A.set_schedule(g, cap, fmt)
        single = [A.ground_state() for A in mats]
        batch = engine.ground_state_batch(mats)
        for p, d, b in zip(plain, single, batch):
            assert p[0] == d[0] == b[0]
            assert np.array_equal(p[1], d[1]) and np.array_equal(p[1], b[1])
        m32 = [A.ground_state(precision=1) for A in mats]
        b32 = engine.ground_state_batch(mats, engine.default_params(precision=1))
        for d, b in zip(m32, b32):
            assert d[0] == b[0] and np.array_equal(d[1], b[1])
    finally:
        for A in mats:
            A.free()


@pytest.mark.gpu
@pytest.mark.parametrize("lanes,cols16", [(1, False), (2, False), (4, False), (1, True), (2, True), (4, True),
                                          (1, "idx"), (2, "idx"), (4, "idx")])
def test_gpu_sliced_ed_and_engine(lanes, cols16):
    """ED matrices and the m=400 superblock in the sliced form (32- or 16-bit columns, or
    16-bit columns with 16-bit value indices -- the superblock, with 77,946 distinct values,
    reads the 16-bit-column form there): energies within 1e-10 of the CSR form's
    (FP32-refined within 1e-9), single and batched solves the same bits; with 16-bit columns
    and value indices also the same bits as with 32-bit columns."""
    from dmrg_c_amd import ed, engine
    mats = []
    for m in (ed.Model("xxz", L=16, sz=0, J_z=0.7, h_z=0.2), ed.Model("hubbard", L=8, N=8, sz=0, U=3.0)):
        P = ed.EdProblem.from_model(m)
        mats.append(P.assemble())
        P.free()
    A, e_ref = _matrix()
    mats.append(A)
    try:
        plain, wide = [], []
        for M in mats:
            M.set_schedule(8, 4096, 2)
            plain.append(M.ground_state())
            M.set_schedule(lanes, 4096, 24 | (4 if lanes == 2 else 0))
            if cols16:
                wide.append(M.ground_state())
                v = 32 if cols16 == "idx" else 25
                M.set_schedule(lanes, 4096, v | (4 if lanes == 2 else 0) | (2 if lanes == 4 else 0))
        print(f"sliced lanes={lanes} 16-bit columns={cols16}", flush=True)
        single = [M.ground_state() for M in mats]
        for w, d in zip(wide, single):
            assert d[0] == w[0] and np.array_equal(d[1], w[1])
        batch = engine.ground_state_batch(mats)
        for p, d, b in zip(plain, single, batch):
            assert abs(d[0] - p[0]) <= 1e-10 * abs(p[0])
            assert d[0] == b[0] and np.array_equal(d[1], b[1])
        assert abs(single[-1][0] - e_ref) <= 1e-10 * abs(e_ref)
        m32 = [M.ground_state(precision=1) for M in mats]
        b32 = engine.ground_state_batch(mats, engine.default_params(precision=1))
        for p, d, b in zip(plain, m32, b32):
            assert abs(d[0] - p[0]) <= 1e-9 * abs(p[0])
            assert d[0] == b[0] and np.array_equal(d[1], b[1])
    finally:
        for M in mats:
            M.free()
