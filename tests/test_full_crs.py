"""The full-CRS reference symbols on the expanded golden superblocks, through the C ABI.

The superblock Hamiltonians of the reference's runs are stored as the lower triangle
GET_HAM_LLLRRRRL returns; expanded to the symmetric full CRS (the oracle's or_expand_sym,
pinned against dense numpy products) they are the inputs of the reference's full-CRS path:

* MATRIX_VECTOR_PRODUCT (sml/MATRIX_VECTOR_PRODUCT.c:6-21): equal to the oracle's full
  product within 1e-13, and bit for bit to MATRIX_VECTOR_PRODUCT_SYM on the lower triangle
  (both run the same device product over the same full matrix);
* DMRG_DIAGONALIZE_SUPERBLOCK (dmrg/DMRG_DIAGONALIZE_SUPERBLOCK.c:7-64, "Lanczos" and
  "Lanczos_Slow"): E within 1e-10 of the reference run's energy, vector equal up to sign,
  gs_error below ii_acc;
* LANCZOS_SLOW (sml/LANCZOS_SLOW.c:13-286) with Guess "Yes": the oracle's start vector
  gives the oracle's energy and step count (within 3);
* LANCZOS (sml/LANCZOS.c:13-270) restarting every 20 steps: the reference's energy;
* INVERSE_ITERATION (sml/INVERSE_ITERATION.c:9-117): from the reference's vector with a
  perturbation, error below ii_acc and the energy kept;
* CONJUGATE_GRADIENT (sml/CONJUGATE_GRADIENT.c:8-120) on H + 20 I: the system solved;
* CONVERGE_CHECK (sml/CONVERGE_CHECK.c:7-23): the host L1 residual.
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import dmrg_c_amd as D
import oracle_py as O
from conftest import GOLDEN
from dmrg_c_amd import snapshot

pytestmark = pytest.mark.gpu

NAMES = ["xxz_L12_m24_jz05_s09", "hubbard_L8_m32_s05", "klm_L6_m32_s03", "klm_L6_m24_ixy_s04"]


@pytest.fixture(autouse=True, scope="module")
def _quiet_files():
    old = os.getcwd()
    os.chdir(tempfile.mkdtemp())  # SML_out/ step files land here
    yield
    os.chdir(old)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Full:
    def __init__(self, name):
        s = snapshot.load(f"{GOLDEN}/{name}.npz")
        self.s = s
        self.n = s.n
        self.row, self.col, self.val = O.expand_sym(s.n, s.ham.row, s.ham.col, s.ham.val)
        self.H = D.HostCRS(self.row, self.col, self.val)
        self.e_ref = float(s.meta["gs_energy"][0])
        self.v_ref = s.meta["gs_vec"]

    def dense_gap(self):
        if self.n > 1500:
            return 1.0
        Dm = np.zeros((self.n, self.n))
        for i in range(self.n):
            Dm[i, self.col[self.row[i]:self.row[i + 1]]] = self.val[self.row[i]:self.row[i + 1]]
        w = np.linalg.eigvalsh(Dm)
        return w[1] - w[0]


@pytest.fixture(scope="module", params=NAMES)
def full(request):
    return Full(request.param)


def test_matrix_vector_product(full):
    L = D.lib()
    x = np.random.default_rng(3).standard_normal(full.n)
    y = np.empty(full.n)
    print(f"MATRIX_VECTOR_PRODUCT n={full.n}", flush=True)
    L.MATRIX_VECTOR_PRODUCT(full.H.ptr, x.ctypes.data, y.ctypes.data, 1)
    yo = O.spmv_full(full.row, full.col, full.val, x)
    assert np.abs(y - yo).max() <= 1e-13 * max(1.0, np.abs(yo).max())
    Hs = D.HostCRS(full.s.ham.row, full.s.ham.col, full.s.ham.val)
    ys = np.empty(full.n)
    L.MATRIX_VECTOR_PRODUCT_SYM(Hs.ptr, x.ctypes.data, ys.ctypes.data, None, 1)
    assert np.array_equal(y.view(np.int64), ys.view(np.int64))


@pytest.mark.parametrize("method", ["Lanczos", "Lanczos_Slow"])
def test_diagonalize_superblock_full(full, method):
    L = D.lib()
    P = D.default_dmrg_parameter(method=method)
    st = D.DMRG_STATUS()
    vec = np.zeros(full.n)
    st.GS_Vec = _dp(vec)
    print(f"DMRG_DIAGONALIZE_SUPERBLOCK {method} n={full.n}", flush=True)
    L.DMRG_DIAGONALIZE_SUPERBLOCK(full.H.ptr, C.byref(D.DMRG_TIME()), C.byref(P), C.byref(st), 1)
    assert abs(st.gs_val - full.e_ref) <= 1e-10 * abs(full.e_ref), (st.gs_val, full.e_ref)
    assert st.gs_error < 1e-9
    assert abs(np.linalg.norm(vec) - 1) < 1e-12
    if full.dense_gap() > 1e-6:
        sg = np.sign(np.dot(vec, full.v_ref))
        assert np.abs(vec - sg * full.v_ref).max() <= 1e-8


def test_lanczos_slow_same_start_as_oracle(full):
    if full.n < 1000:
        pytest.skip("n < 1000 takes the dense path in the reference")
    start = O.start_vector(full.n, 17)
    k_o, e_o, _ = O.lanczos_slow(O.CrsOperator(full.row, full.col, full.val, sym=False), start, acc=1e-12)
    vec = start.copy()
    ev = C.c_double()
    box = D.BOX_LAN(full.H.ptr, _dp(vec), C.pointer(ev), 1e-12, 1, 800, 800, 1, b"Yes", b"Normal")
    D.lib().LANCZOS_SLOW(C.byref(box))
    assert abs(ev.value - e_o) <= 1e-10 * abs(e_o)
    steps = int(open("SML_out/LANCZOS_SLOW_Step.txt").read().split()[-1])
    assert abs(steps - k_o) <= 3


def test_lanczos_restarts(full):
    if full.n < 1000:
        pytest.skip("n < 1000 takes the dense path in the reference")
    vec = O.start_vector(full.n, 23)
    ev = C.c_double()
    box = D.BOX_LAN(full.H.ptr, _dp(vec), C.pointer(ev), 1e-12, 1, 800, 20, 1, b"Yes", b"Normal")
    D.lib().LANCZOS(C.byref(box))
    assert abs(ev.value - full.e_ref) <= 1e-10 * abs(full.e_ref)
    r = O.spmv_full(full.row, full.col, full.val, vec) - ev.value * vec
    assert np.abs(r).max() <= 1e-4 * max(1.0, abs(full.e_ref))


def test_inverse_iteration_and_converge_check(full):
    L = D.lib()
    rng = np.random.default_rng(5)
    vec = full.v_ref + 1e-4 * rng.standard_normal(full.n)
    vec /= np.linalg.norm(vec)
    ev = C.c_double(full.e_ref)
    box = D.BOX_II(full.H.ptr, 1e-9, 1e-11, 3, 1e-7, 2000, 0.0, _dp(vec), C.pointer(ev), 1, b"CG")
    L.INVERSE_ITERATION(C.byref(box))
    assert box.error < 1e-9, box.error
    assert abs(ev.value - full.e_ref) <= 1e-10 * abs(full.e_ref)
    L.CONVERGE_CHECK.restype = C.c_double
    L.CONVERGE_CHECK.argtypes = [C.POINTER(D.CRS1), C.c_void_p, C.c_double, C.c_int]
    r_gpu = L.CONVERGE_CHECK(full.H.ptr, vec.ctypes.data, ev.value, 1)
    r_cpu = np.abs(ev.value * vec - O.spmv_full(full.row, full.col, full.val, vec)).sum()
    assert abs(r_gpu - r_cpu) <= 1e-12 + 1e-8 * r_cpu
    assert abs(r_gpu - box.error) <= 1e-12 + 1e-6 * box.error


def test_conjugate_gradient(full):
    shifted = full.val.copy()
    for i in range(full.n):
        seg = slice(full.row[i], full.row[i + 1])
        shifted[seg][full.col[seg] == i] += 20.0
    Hs = D.HostCRS(full.row, full.col, shifted)
    b = np.random.default_rng(0).standard_normal(full.n)
    x = np.ones(full.n)
    box = D.BOX_CG(Hs.ptr, 1e-20, _dp(b), _dp(x), 2000, 1)
    D.lib().CONJUGATE_GRADIENT(C.byref(box))
    res = b - O.spmv_full(full.row, full.col, shifted, x)
    assert np.linalg.norm(res) < 1e-9
