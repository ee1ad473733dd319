"""GPU parity of libdmrg_c_amd.so against the oracle and the reference's golden snapshots.
Everything goes through the C ABI.  Tolerances: CRS structure and values bit-exact;
SpMV 1e-13 relative; energy 1e-10 relative; eigenvector 1e-8 (max-norm, up to sign)."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import dmrg_c_amd as D
import oracle_py as O
from conftest import GOLDEN, golden_ids, golden_snapshots
from dmrg_c_amd import snapshot
from sbhelp import BasisBundle, call_terms, host_ops

pytestmark = pytest.mark.gpu

SNAPS = golden_snapshots()


@pytest.fixture(autouse=True, scope="module")
def _quiet_files():
    old = os.getcwd()
    d = tempfile.mkdtemp()
    os.chdir(d)  # SML_out/ step files land here
    yield
    os.chdir(old)


def _gap_ok(s):
    """Ground state non-degenerate (vector comparison meaningful)."""
    if s.n > 1500:
        return True
    fr, fc, fv = O.expand_sym(s.n, s.ham.row, s.ham.col, s.ham.val)
    Dm = np.zeros((s.n, s.n))
    for i in range(s.n):
        Dm[i, fc[fr[i]:fr[i + 1]]] = fv[fr[i]:fr[i + 1]]
    w = np.linalg.eigvalsh(Dm)
    return w[1] - w[0] > 1e-6


@pytest.mark.parametrize("path", SNAPS, ids=golden_ids(SNAPS))
def test_device_assembly_bit_exact(path):
    s = snapshot.load(path)
    sb = D.Superblock(s)
    A = sb.assemble(full=True)
    row, col, val = sb.download_lower()
    assert np.array_equal(row, s.ham.row)
    assert np.array_equal(col, s.ham.col)
    assert np.array_equal(val.view(np.int64), s.ham.val.view(np.int64))
    fr, fc, fv = A.download()
    er, ec, ev = O.expand_sym(s.n, s.ham.row, s.ham.col, s.ham.val)
    assert np.array_equal(fr, er) and np.array_equal(fc, ec) and np.array_equal(fv, ev)


@pytest.mark.parametrize("path", SNAPS[::3], ids=golden_ids(SNAPS[::3]))
def test_spmv_matches_oracle(path):
    s = snapshot.load(path)
    A = D.DeviceMatrix(s.ham.row, s.ham.col, s.ham.val, sym=True)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(s.n)
    y = A.matvec(x)
    yo = O.spmv_sym(s.ham.row, s.ham.col, s.ham.val, x)
    assert np.abs(y - yo).max() <= 1e-13 * max(1.0, np.abs(yo).max())
    # reference symbol on a host CRS1
    H = D.HostCRS(s.ham.row, s.ham.col, s.ham.val)
    out = np.empty(s.n)
    D.lib().MATRIX_VECTOR_PRODUCT_SYM(H.ptr, x.ctypes.data, out.ctypes.data, None, 1)
    assert np.array_equal(out, y)


@pytest.mark.parametrize("path", SNAPS, ids=golden_ids(SNAPS))
def test_diagonalize_superblock_sym(path):
    """DMRG_DIAGONALIZE_SUPERBLOCK_SYM on the reference's CRS: E within 1e-10 of the
    reference's, GS vector equal up to sign, residual below ii_acc."""
    s = snapshot.load(path)
    e_ref = float(s.meta["gs_energy"][0])
    E, err, vec, _ = D.diagonalize_superblock_sym(s.ham.row, s.ham.col, s.ham.val)
    assert abs(E - e_ref) <= 1e-10 * abs(e_ref), (E, e_ref)
    assert err < 1e-9
    assert abs(np.linalg.norm(vec) - 1) < 1e-12
    if _gap_ok(s):
        v_ref = s.meta["gs_vec"]
        sg = np.sign(np.dot(vec, v_ref))
        assert np.abs(vec - sg * v_ref).max() <= 1e-8


@pytest.mark.parametrize("path", [p for p in SNAPS if snapshot.load(p).n >= 1000],
                         ids=golden_ids([p for p in SNAPS if snapshot.load(p).n >= 1000]))
def test_lanczos_slow_same_start_as_oracle(path):
    """LANCZOS_SLOW_SYM with Guess "Yes": same start vector as the oracle -> same energy
    and the same convergence step within a few steps."""
    s = snapshot.load(path)
    start = O.start_vector(s.n, 11)
    k_o, e_o, v_o = O.lanczos_slow(O.CrsOperator(s.ham.row, s.ham.col, s.ham.val), start, acc=1e-12)
    H = D.HostCRS(s.ham.row, s.ham.col, s.ham.val)
    vec = start.copy()
    ev = C.c_double()
    box = D.BOX_LAN(H.ptr, vec.ctypes.data_as(C.POINTER(C.c_double)), C.pointer(ev), 1e-12, 1, 800, 800, 1,
                    b"Yes", b"Normal")
    D.lib().LANCZOS_SLOW_SYM(C.byref(box))
    assert abs(ev.value - e_o) <= 1e-10 * abs(e_o)
    steps = int(open("SML_out/LANCZOS_SLOW_SYM_Step.txt").read().split()[-1])
    assert abs(steps - k_o) <= 3


@pytest.mark.parametrize("block,lan_type", [(20, b"Normal"), (37, b"Normal"), (800, b"Tight")],
                         ids=["restart20", "restart37", "tight"])
@pytest.mark.parametrize("path", [p for p in SNAPS if snapshot.load(p).n >= 1000][:3],
                         ids=golden_ids([p for p in SNAPS if snapshot.load(p).n >= 1000][:3]))
def test_lanczos_sym_restarts_and_tight(path, block, lan_type):
    """LANCZOS_SYM restarting from the Ritz vector every `block` steps (sml/LANCZOS_SYM.c:248-255)
    and the "Tight" criterion reach the reference's energy; the chunks the driver queues ahead
    never cross a restart."""
    s = snapshot.load(path)
    H = D.HostCRS(s.ham.row, s.ham.col, s.ham.val)
    vec = O.start_vector(s.n, 5)
    ev = C.c_double()
    box = D.BOX_LAN(H.ptr, vec.ctypes.data_as(C.POINTER(C.c_double)), C.pointer(ev), 1e-12, 1, 800, block, 1,
                    b"Yes", lan_type)
    D.lib().LANCZOS_SYM(C.byref(box))
    e_ref = float(s.meta["gs_energy"][0])
    assert abs(ev.value - e_ref) <= 1e-10 * abs(e_ref)
    H2 = O.spmv_sym(s.ham.row, s.ham.col, s.ham.val, vec)
    assert np.abs(H2 - ev.value * vec).max() <= 1e-4 * max(1.0, abs(e_ref))


def test_start_vector_matches_oracle():
    n = 1000
    assert np.array_equal(D.start_vector(n, 5), O.start_vector(n, 5))


@pytest.mark.parametrize("path", [p for p in SNAPS if "xxz" in p and snapshot.load(p).n > 8],
                         ids=golden_ids([p for p in SNAPS if "xxz" in p and snapshot.load(p).n > 8]))
def test_obc_matvec_and_solvers(path):
    """DMRG_MATRIX_VECTOR_PRODUCT_OBC equals the assembled operator; the OBC Lanczos and
    inverse iteration reach the reference's energy."""
    s = snapshot.load(path)
    bb = BasisBundle(s)
    L = D.lib()
    rng = np.random.default_rng(2)
    x = rng.standard_normal(s.n)
    y = np.empty(s.n)
    M = bb.M
    ele = bb.ele_rr.ctypes.data_as(C.POINTER(C.c_int))
    L.DMRG_MATRIX_VECTOR_PRODUCT_OBC(M["M_LLLR"].ptr, M["M_LRRL"].ptr, M["M_LRRL_Sign"].ptr, M["M_RRRL"].ptr,
                                     x.ctypes.data, y.ctypes.data, ele, 1, bb.ptr)
    y_crs = O.spmv_sym(s.ham.row, s.ham.col, s.ham.val, x)
    assert np.abs(y - y_crs).max() <= 1e-12 * max(1.0, np.abs(y_crs).max())
    if s.n < 64:
        return
    e_ref = float(s.meta["gs_energy"][0])
    vec = np.zeros(s.n)
    ev = C.c_double()
    box = D.DMRG_BOX_LAN(vec.ctypes.data_as(C.POINTER(C.c_double)), C.pointer(ev), 1e-14, 1, 800, 1, b"No",
                         b"Normal", M["M_LLLR"].ptr, M["M_LRRL"].ptr, M["M_LRRL_Sign"].ptr, M["M_RRRL"].ptr, ele,
                         bb.ptr)
    L.DMRG_LANCZOS_SLOW_OBC(C.byref(box))
    assert abs(ev.value - e_ref) <= 1e-10 * abs(e_ref)
    ii = D.DMRG_BOX_II(1e-9, 1e-11, 3, 1e-6, 1000, vec.ctypes.data_as(C.POINTER(C.c_double)), C.pointer(ev), 0.0, 1,
                       M["M_LLLR"].ptr, M["M_LRRL"].ptr, M["M_LRRL_Sign"].ptr, M["M_RRRL"].ptr, ele, bb.ptr)
    L.DMRG_INVERSE_ITERATION_OBC(C.byref(ii))
    assert ii.error < 1e-9
    tmp = np.empty(s.n)
    err = L.DMRG_CONVERGE_CHECK_OBC(ev.value, M["M_LLLR"].ptr, M["M_LRRL"].ptr, M["M_LRRL_Sign"].ptr,
                                    M["M_RRRL"].ptr, vec.ctypes.data, tmp.ctypes.data, ele, 1, bb.ptr)
    assert abs(err - ii.error) <= 1e-12 + 1e-6 * ii.error


@pytest.mark.parametrize("name", ["hubbard_L8_m32_s07", "klm_L6_m32_s04", "xxz_L12_m24_jz05_s13"])
def test_dropin_record_assemble_solve(name):
    """The drop-in sequence: record the model's term list through the DMRG_MAKE_ELEM_*
    symbols, dmrgx_get_ham_lllrrrrl (GET_HAM_LLLRRRRL body), DMRG_DIAGONALIZE_SUPERBLOCK_SYM."""
    s = snapshot.load(f"{GOLDEN}/{name}.npz")
    L = D.lib()
    ops = host_ops(s)
    eles = [np.ascontiguousarray(e, np.int32) for e in s.eles]
    bb = BasisBundle(s)
    assert L.dmrgx_record_begin() == 0
    B1 = D.DMRG_BASIS_ONSITE(0, 0, 0, 0, 0, s.dim_RR, s.dim_onsite)
    call_terms(s, ops, eles, C.byref(B1), None, None, None)
    T = L.dmrgx_record_end()
    Hp = L.dmrgx_get_ham_lllrrrrl(bb.ptr, T)
    H = Hp.contents
    nnz = H.Row[s.n]
    row = np.ctypeslib.as_array(H.Row, shape=(s.n + 1,)).copy()
    col = np.ctypeslib.as_array(H.Col, shape=(nnz,)).copy()
    val = np.ctypeslib.as_array(H.Val, shape=(nnz,)).copy()
    assert np.array_equal(row, s.ham.row) and np.array_equal(col, s.ham.col)
    assert np.array_equal(val.view(np.int64), s.ham.val.view(np.int64))
    P = D.default_dmrg_parameter()
    st = D.DMRG_STATUS()
    gs = np.zeros(s.n)
    st.GS_Vec = gs.ctypes.data_as(C.POINTER(C.c_double))
    L.DMRG_DIAGONALIZE_SUPERBLOCK_SYM(Hp, C.byref(D.DMRG_TIME()), C.byref(P), C.byref(st), 1)
    e_ref = float(s.meta["gs_energy"][0])
    assert abs(st.gs_val - e_ref) <= 1e-10 * abs(e_ref)
    L.dmrgx_terms_free(T)


def test_cg_and_converge_check_symbols():
    s = snapshot.load(f"{GOLDEN}/hubbard_L8_m32_s03.npz")
    L = D.lib()
    H = D.HostCRS(s.ham.row, s.ham.col, s.ham.val)
    v = s.meta["gs_vec"].copy()
    e = float(s.meta["gs_energy"][0])
    r_gpu = L.CONVERGE_CHECK_SYM(H.ptr, v.ctypes.data, e, 1)
    y = O.spmv_sym(s.ham.row, s.ham.col, s.ham.val, v)
    r_cpu = np.abs(e * v - y).sum()
    assert abs(r_gpu - r_cpu) <= 1e-12 + 1e-8 * r_cpu
    # CG on H + 5 I (positive definite): solution satisfies the system
    fr, fc, fv = O.expand_sym(s.n, s.ham.row, s.ham.col, s.ham.val)
    shifted = s.ham.val.copy()
    diag_pos = s.ham.row[1:] - 1
    shifted[diag_pos] += 20.0
    Hs = D.HostCRS(s.ham.row, s.ham.col, shifted)
    b = np.random.default_rng(0).standard_normal(s.n)
    x = np.ones(s.n)
    box = D.BOX_CG(Hs.ptr, 1e-20, b.ctypes.data_as(C.POINTER(C.c_double)), x.ctypes.data_as(C.POINTER(C.c_double)),
                   2000, 1)
    L.CONJUGATE_GRADIENT_SYM(C.byref(box))
    res = b - O.spmv_sym(s.ham.row, s.ham.col, shifted, x)
    assert np.linalg.norm(res) < 1e-9


@pytest.mark.parametrize("name", ["hubbard_L64_m400_s31", "xxz_L64_m300_s31"])
def test_big_superblock_assembly_and_solve(name):
    """Hubbard L=64, m=400 (n=144542) and XXZ L=64, m=300 (n=23524) superblocks of the
    reference's runs (last infinite-system steps; the served stream's superblocks): device
    assembly identical to the oracle's (which equals the reference's bit for bit on the
    golden set) and E within 1e-10 of the reference run's energy."""
    s = snapshot.load(os.path.join(GOLDEN, name + ".npz"))
    print(f"big superblock {name} n={s.n}", flush=True)
    sb = D.Superblock(s)
    A = sb.assemble(full=True)
    row, col, val = sb.download_lower()
    orow, ocol, oval = O.assemble(s)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    assert np.array_equal(val.view(np.int64), oval.view(np.int64))
    E, vec, st = A.ground_state()
    e_ref = float(s.meta["gs_energy"][0])
    assert abs(E - e_ref) <= 1e-10 * abs(e_ref), (E, e_ref)
    assert st.gs_error < 1e-9
