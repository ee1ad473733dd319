"""Edge cases of the device paths: ragged CSR rows (empty, one entry, thousands) under every
lanes-per-row count and storage format (the sliced form included); one-state sectors; non-convergence reported as a
status code by the single and batched solves; malformed input rejected before any launch."""
import numpy as np
import pytest


def _ragged(n=5000, seed=0, distinct=None):
    """Random full CSR: row lengths 0 (10 %), 1, a few hundred and a few thousand."""
    rng = np.random.default_rng(seed)
    lens = rng.choice([0, 1, 3, 17, 64, 250, 2000], size=n, p=[0.1, 0.1, 0.2, 0.3, 0.2, 0.09, 0.01])
    row = np.zeros(n + 1, np.int64)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in lens]).astype(np.int32)
    if distinct:
        val = rng.choice(rng.standard_normal(distinct), size=row[-1])
    else:
        val = rng.standard_normal(row[-1])
    return row, col, val


@pytest.mark.gpu
@pytest.mark.parametrize("n", [70000, 300000])
def test_gpu_sliced_16bit_columns_with_wide_gaps(n):
    """The sliced form's 16-bit columns on rows whose column gaps exceed 65534 (kept whole in
    the exception list), next to runs of small gaps and rows of every length: products equal
    to the 32-bit sliced form's bit for bit and to scipy's within 1e-12, for 1, 2 and 4 lanes
    per row, plain and non-temporal loads; exceptions counted on the host must match."""
    import scipy.sparse as sp
    import dmrg_c_amd as D
    rng = np.random.default_rng(n)
    lens = rng.choice([0, 1, 2, 9, 40, 130], size=n, p=[0.05, 0.1, 0.15, 0.4, 0.25, 0.05])
    cols = []
    for r, k in enumerate(lens):
        near = rng.integers(max(0, r - 3000), min(n, r + 3000), size=k // 2)   # small gaps
        far = rng.integers(0, n, size=k - k // 2)                              # gaps up to n
        cols.append(np.unique(np.concatenate([near, far])))
    row = np.zeros(n + 1, np.int64)
    row[1:] = np.cumsum([len(c) for c in cols])
    col = np.concatenate(cols).astype(np.int32)
    val = rng.standard_normal(row[-1])
    A = D.DeviceMatrix(row, col, val, sym=False)
    try:
        x = rng.standard_normal(n)
        ref = sp.csr_matrix((val, col, row), shape=(n, n)) @ x
        for g in (1, 2, 4):
            # exceptions expected: per slot column of each 64/g-row slice, the differences of
            # consecutive slots (padding repeats the row's last column) outside 0..65534
            C_ = 64 // g
            rl = np.diff(row)
            width = np.repeat([((rl[i:i + C_] + g - 1) // g).max() for i in range(0, n, C_)], C_)[:n]
            exp = 0
            for r in range(n):
                c = cols[r].astype(np.int64)
                for l in range(min(g, len(c))):
                    sub = c[l::g]
                    exp += int(np.count_nonzero(np.diff(sub) > 65534))
                    if len(sub) < width[r] and c[-1] - sub[-1] > 65534:   # padding: the row's last column
                        exp += 1
            print(f"sliced 16-bit n={n} lanes={g} exceptions expected {exp}", flush=True)
            assert A.sell16_exceptions(g) == exp
            A.set_schedule(g, 0, 24)
            y32 = A.matvec(x)
            assert np.abs(y32 - ref).max() <= 1e-12 * (np.abs(ref).max() + 1)
            for v in (25, 27, 29, 31):
                A.set_schedule(g, 2048 if v & 4 else 0, v)
                assert np.array_equal(A.matvec(x), y32), (g, v)
        assert exp > 0
    finally:
        A.free()


@pytest.mark.gpu
@pytest.mark.parametrize("distinct", [None, 40], ids=["random_values", "few_values"])
def test_gpu_ragged_rows_every_schedule(distinct):
    import scipy.sparse as sp
    import dmrg_c_amd as D
    row, col, val = _ragged(distinct=distinct)
    n = len(row) - 1
    A = D.DeviceMatrix(row, col, val, sym=False)
    try:
        x = np.random.default_rng(1).standard_normal(n)
        ref = sp.csr_matrix((val, col, row), shape=(n, n)) @ x
        for g in (1, 2, 4, 8, 16, 32, 64):
            base, pair = None, None
            for v in (0, 1, 2, 3, 5, 8, 10, 16, 18, 23):
                A.set_schedule(g, 2048 if v & 4 else 0, v)
                y = A.matvec(x)
                assert np.abs(y - ref).max() <= 1e-12 * (np.abs(ref).max() + 1)
                if v & 4:
                    continue
                if v in (1, 3):                              # paired CSR loads: their own order
                    pair = y if pair is None else pair
                    assert np.array_equal(y, pair), (g, v)
                else:
                    base = y if base is None else base
                    assert np.array_equal(y, base), (g, v)   # formats and in-flight depth: same bits
        # sliced form: 1, 2, 4 lanes per row; its other load bits and the row mapping leave
        # the products' bits alone
        for g in (1, 2, 4):
            sl = None
            for v in (24, 25, 26, 27, 28, 31, 32, 34, 36, 39):   # 32..39: value indices where the table is small
                A.set_schedule(g, 2048 if v & 4 else 0, v)
                y = A.matvec(x)
                assert np.abs(y - ref).max() <= 1e-12 * (np.abs(ref).max() + 1), (g, v)
                sl = y if sl is None else sl
                assert np.array_equal(y, sl), (g, v)
            assert np.all(A.matvec(np.ones(n))[np.diff(row) == 0] == 0.0)
        A.set_schedule(8, 0, 0)
        assert np.all(A.matvec(np.ones(n))[np.diff(row) == 0] == 0.0)   # empty rows give 0
    finally:
        A.free()


@pytest.mark.gpu
def test_gpu_one_state_sector():
    """XXZ with every spin up: a 1 x 1 matrix, whose eigenpair is the element itself, alone
    and through the engine."""
    from dmrg_c_amd import ed, engine
    P = ed.EdProblem.from_model(ed.Model("xxz", L=6, sz=6, J_z=0.9, h_z=0.3))
    A = P.assemble()
    P.free()
    try:
        assert A.n == 1
        _, _, val = A.download()
        E, vec, st = A.ground_state()
        assert E == val[0] and vec[0] == 1.0
        (Eb, vb, _), = engine.ground_state_batch([A])
        assert Eb == E and vb[0] == 1.0
    finally:
        A.free()


@pytest.mark.gpu
def test_gpu_no_convergence_is_a_status():
    """A Lanczos limit of 5 steps on an n = 12870 chain: dmrgx_ground_state returns
    DMRGX_ERR_NOCONV (no exit), and the engine reports the same status for that request
    while the others in the batch still finish."""
    import dmrg_c_amd as D
    from dmrg_c_amd import ed, engine
    mats = []
    for m in (ed.Model("xxz", L=16, sz=0, J_z=0.7), ed.Model("xxz", L=12, sz=0, J_z=1.1)):
        P = ed.EdProblem.from_model(m)
        mats.append(P.assemble())
        P.free()
    try:
        with pytest.raises(D.DmrgxError) as ei:
            mats[0].ground_state(diag_max_step=5)
        assert ei.value.code == -3
        E = engine.Engine()
        E.submit(mats[0], engine.default_params(diag_max_step=5), tag=0)
        E.submit(mats[1], engine.default_params(), tag=1)
        res = {r.tag: r for r in E.run_until_done()}
        E.free()
        assert res[0].status == -3
        assert res[1].status == 0 and res[1].energy == mats[1].ground_state()[0]
    finally:
        for A in mats:
            A.free()


@pytest.mark.gpu
def test_gpu_malformed_input_rejected():
    """A lower-triangle CRS whose row lacks its diagonal, and decreasing row pointers, are
    refused with status codes (the reference exits in DIAG_ADD_CRS1)."""
    import dmrg_c_amd as D
    row = np.array([0, 1, 2, 3], np.int64)
    col = np.array([0, 0, 1], np.int32)        # row 1 holds (1, 0) only: no diagonal
    val = np.array([1.0, 0.5, 2.0])
    with pytest.raises(D.DmrgxError):
        A = D.DeviceMatrix(row, col, val, sym=True)
        A.ground_state()
    bad_row = np.array([0, 2, 1, 3], np.int64)
    with pytest.raises(D.DmrgxError):
        D.DeviceMatrix(bad_row, np.array([0, 1, 2], np.int32), np.ones(3), sym=False)
