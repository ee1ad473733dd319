"""Exact diagonalization through libdmrg_c_amd.so.

ctypes mirrors of the reference's ED structs (include/exact.h) for the EXACT_* entry points,
and ``EdProblem``, a device-resident ED problem (basis, on-site operators, term list) built
from a model description or from arrays, assembled into a device CSR on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import CRS1, DeviceMatrix, HostCRS, check, lib as _lib

_D = C.POINTER(C.c_double)
_L = C.POINTER(C.c_long)
_I = C.POINTER(C.c_int)


class EXACT_PARAMETER(C.Structure):
    """include/exact.h:14-37"""
    _fields_ = [("diag_num", C.c_int), ("est_max_row_elem_num", C.c_int), ("Diag_Method", C.c_char * 1000),
                ("diag_acc", C.c_double), ("diag_min_step", C.c_int), ("diag_max_step", C.c_int),
                ("cg_acc", C.c_double), ("cg_max_step", C.c_int), ("inv_iter_acc", C.c_double),
                ("inv_iter_diag_add", C.c_double), ("inv_iter_max_step", C.c_int), ("param_iter", C.c_int),
                ("param_iter_now", C.c_int)]


class EXACT_TIME(C.Structure):
    """include/exact.h:39-51"""
    _fields_ = [(k, C.c_double) for k in ("diag", "inv_iter", "make_ham", "find_dim", "make_basis", "exp_values",
                                          "cf", "sc", "total")]


class EXACT_HAM_INFO(C.Structure):
    """include/exact.h:54-64"""
    _fields_ = [("dim", C.c_int), ("diag_num", C.c_int), ("Vector", C.POINTER(_D)), ("Value", _D), ("Error", _D),
                ("mem_ham", C.c_double), ("Ham", C.POINTER(CRS1))]


class EXACT_BASIS_INFO(C.Structure):
    """include/exact.h:66-74"""
    _fields_ = [("dim", C.c_int), ("Basis", _L), ("tot_sz", C.c_int), ("tot_ele", C.c_int),
                ("tot_parity", C.c_int)]


class EXACT_A_BASIS(C.Structure):
    """include/exact.h:76-87"""
    _fields_ = [("Basis", _L), ("elem_num", C.c_long), ("Check", _L), ("Val", _D), ("max_row", C.c_int),
                ("dim_onsite", C.c_int), ("dim_whole", C.c_long)]


class EXACT_WHOLE_BASIS_Q1(C.Structure):
    """include/exact.h:89-95"""
    _fields_ = [("Dim", _I), ("Basis", C.POINTER(_L)), ("max_dim", C.c_int)]


class EXACT_WHOLE_BASIS_Q2(C.Structure):
    """include/exact.h:97-103"""
    _fields_ = [("Dim", C.POINTER(_I)), ("Basis", C.POINTER(C.POINTER(_L))), ("max_dim", C.c_int)]


class EXACT_WHOLE_BASIS_Q3(C.Structure):
    """include/exact.h:105-111"""
    _fields_ = [("Dim", C.POINTER(C.POINTER(_I))), ("Basis", C.POINTER(C.POINTER(C.POINTER(_L)))),
                ("max_dim", C.c_int)]


class ED_TERM(C.Structure):
    """dmrgx_ed_term (include/dmrg_c_amd.h)"""
    _fields_ = [("kind", C.c_int32), ("site1", C.c_int32), ("site2", C.c_int32), ("op1", C.c_int32),
                ("op2", C.c_int32), ("par_lo", C.c_int32), ("par_hi", C.c_int32), ("sign", C.c_int32),
                ("coeff", C.c_double)]


TERM_DTYPE = np.dtype([("kind", "<i4"), ("site1", "<i4"), ("site2", "<i4"), ("op1", "<i4"), ("op2", "<i4"),
                       ("par_lo", "<i4"), ("par_hi", "<i4"), ("sign", "<i4"), ("coeff", "<f8")])
assert TERM_DTYPE.itemsize == C.sizeof(ED_TERM) == 40


class ED_MODEL(C.Structure):
    """dmrgx_ed_model (include/dmrg_c_amd.h)"""
    _fields_ = [("model", C.c_int32), ("L", C.c_int32), ("spin", C.c_int32), ("N", C.c_int32), ("sz", C.c_int32),
                ("bc", C.c_int32), ("J_xy", C.c_double), ("J_z", C.c_double), ("h_z", C.c_double),
                ("D_z", C.c_double), ("t", C.c_double), ("U", C.c_double), ("V", C.c_double), ("mu", C.c_double)]


BC = {"OBC": 0, "PBC": 1, "SSD": 2}
XXZ_OPS = ("Ham_On", "Sp", "Sm", "Sz", "Zero")
HUBBARD_OPS = ("Ham_On", "CUp_D", "CUp", "CDown_D", "CDown", "NC_Up", "NC_Down", "Zero", "NC", "SzC")


@dataclass
class Model:
    """An ED model of dmrgx_ed_model_build: "xxz" (spin = 2S, sector sz = 2 Sz_tot) or
    "hubbard" (sector N electrons, sz = 2 Sz_tot)."""
    model: str = "xxz"
    L: int = 12
    spin: int = 1
    N: int = -1
    sz: int = 0
    bc: str = "OBC"
    J_xy: float = 1.0
    J_z: float = 1.0
    h_z: float = 0.0
    D_z: float = 0.0
    t: float = -1.0
    U: float = 4.0
    V: float = 0.0
    mu: float = 0.0

    def to_c(self) -> ED_MODEL:
        m = ED_MODEL()
        m.model = {"xxz": 0, "hubbard": 1}[self.model]
        m.L, m.spin, m.sz, m.bc = self.L, self.spin, self.sz, BC[self.bc]
        m.N = self.L if self.N < 0 else self.N
        for k in ("J_xy", "J_z", "h_z", "D_z", "t", "U", "V", "mu"):
            setattr(m, k, getattr(self, k))
        return m


_bound = False


def lib() -> C.CDLL:
    global _bound
    L = _lib()
    if _bound:
        return L
    vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.dmrgx_ed_create.argtypes = [i32, i32, i64, vp, i32, vp, vp, i32, vp, C.POINTER(vp)]
    L.dmrgx_ed_create_sector.argtypes = [i32, i32, vp, i32, vp, i32, i32, vp, vp, i32, vp, C.POINTER(vp)]
    L.dmrgx_ed_dim.argtypes = [vp]
    L.dmrgx_ed_dim.restype = i64
    L.dmrgx_ed_download_basis.argtypes = [vp, vp]
    L.dmrgx_ed_assemble.argtypes = [vp, C.POINTER(vp)]
    L.dmrgx_ed_apply_onsite.argtypes = [vp, vp, i32, i32, i32, vp, vp]
    L.dmrgx_ed_free.argtypes = [vp]
    L.dmrgx_ed_model_build.argtypes = [C.POINTER(ED_MODEL), C.POINTER(vp)]
    L.dmrgx_ed_model_describe.argtypes = [C.POINTER(ED_MODEL), C.POINTER(i32), C.POINTER(i32), vp, vp, vp, vp,
                                          C.POINTER(i32), vp, vp, vp, vp, vp]
    L.EXACT_DIAGONALIZE_HAMILTONIAN.argtypes = [C.POINTER(EXACT_HAM_INFO), C.POINTER(EXACT_PARAMETER),
                                                C.POINTER(EXACT_TIME), C.c_int]
    L.EXACT_V_M_Q0.argtypes = [C.POINTER(CRS1), vp, vp, C.c_int, C.c_int, C.c_int, C.POINTER(EXACT_BASIS_INFO)]
    L.EXACT_V_M_Q1.argtypes = [C.POINTER(CRS1), C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                               C.POINTER(EXACT_WHOLE_BASIS_Q1)]
    L.EXACT_V_M_Q2.argtypes = [C.POINTER(CRS1), C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, C.c_char_p, vp,
                               C.c_int, C.c_int, C.c_int, C.POINTER(EXACT_WHOLE_BASIS_Q2)]
    L.EXACT_V_M_Q3.argtypes = [C.POINTER(CRS1), C.c_int, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int, vp,
                               C.c_char_p, vp, C.c_int, C.c_int, C.c_int, C.POINTER(EXACT_WHOLE_BASIS_Q3)]
    L.EXACT_EXPECTATION_ONSITE.argtypes = [C.POINTER(CRS1), vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(EXACT_BASIS_INFO)]
    L.EXACT_EXPECTATION_INTERSITE_Q0.argtypes = [C.POINTER(CRS1), C.POINTER(CRS1), C.c_int, C.c_int, vp, vp, vp, vp,
                                                 C.c_int, C.c_int, C.POINTER(EXACT_BASIS_INFO)]
    L.EXACT_GET_A_BASIS.argtypes = [C.c_int, C.c_int]
    L.EXACT_GET_A_BASIS.restype = C.POINTER(C.POINTER(EXACT_A_BASIS))
    L.EXACT_FREE_A_BASIS.argtypes = [C.POINTER(C.POINTER(EXACT_A_BASIS)), C.c_int]
    L.EXACT_FIND_SITE_STATE.argtypes = [C.c_long, C.c_int, C.c_int]
    L.EXACT_FIND_SITE_STATE.restype = C.c_int
    L.EXACT_MAKE_ELEM_ON.argtypes = [C.c_long, C.c_int, C.c_int, C.POINTER(CRS1), C.POINTER(C.c_long), C.c_double,
                                     C.POINTER(EXACT_A_BASIS)]
    L.EXACT_MAKE_ELEM_INTER.argtypes = [C.c_long, C.c_int, C.c_int, C.c_int, C.POINTER(CRS1), C.POINTER(CRS1),
                                        C.POINTER(C.c_long), C.c_double, C.c_int, C.POINTER(EXACT_A_BASIS)]
    L.EXACT_SSD_COEFF.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_char_p]
    L.EXACT_SSD_COEFF.restype = C.c_double
    _bound = True
    return L


def _crs_array(ops: Sequence[HostCRS]):
    arr = (CRS1 * max(1, len(ops)))(*[h.crs for h in ops])
    return arr


class EdProblem:
    """Device-resident ED problem (dmrgx_ed)."""

    def __init__(self, handle: C.c_void_p):
        self.handle = handle

    @classmethod
    def from_model(cls, m: Model) -> "EdProblem":
        h = C.c_void_p()
        mc = m.to_c()
        check(lib().dmrgx_ed_model_build(C.byref(mc), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, L: int, d: int, basis: np.ndarray, ops, terms: np.ndarray,
                    n_ele: Optional[np.ndarray] = None) -> "EdProblem":
        """ops: list of (row, col, val) d x d; terms: TERM_DTYPE array."""
        hops = [HostCRS(r, c, v, d) for (r, c, v) in ops]
        arr = _crs_array(hops)
        b = np.ascontiguousarray(basis, np.int64)
        t = np.ascontiguousarray(terms, TERM_DTYPE)
        ne = None if n_ele is None else np.ascontiguousarray(n_ele, np.int32)
        h = C.c_void_p()
        check(lib().dmrgx_ed_create(L, d, len(b), b.ctypes.data, len(hops), C.cast(arr, C.c_void_p),
                                    None if ne is None else ne.ctypes.data, len(t), t.ctypes.data, C.byref(h)))
        return cls(h)

    @property
    def dim(self) -> int:
        return int(lib().dmrgx_ed_dim(self.handle))

    def basis(self) -> np.ndarray:
        b = np.empty(self.dim, np.int64)
        check(lib().dmrgx_ed_download_basis(self.handle, b.ctypes.data))
        return b

    def assemble(self) -> DeviceMatrix:
        A = C.c_void_p()
        check(lib().dmrgx_ed_assemble(self.handle, C.byref(A)))
        return DeviceMatrix._wrap(A, self.dim)

    def free(self):
        if getattr(self, "handle", None):
            lib().dmrgx_ed_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class ModelDescription:
    d: int
    ops: list            # [(row, col, val)] d x d
    terms: np.ndarray    # TERM_DTYPE
    n_ele: np.ndarray
    qn1: np.ndarray
    qn2: np.ndarray
    targets: np.ndarray  # target1, target2, qn2 used


def describe(m: Model) -> ModelDescription:
    """dmrgx_ed_model_describe: the operators, term list and sector tables of a model (host only)."""
    L = lib()
    mc = m.to_c()
    d, n_ops, n_terms = C.c_int32(), C.c_int32(), C.c_int32()
    check(L.dmrgx_ed_model_describe(C.byref(mc), C.byref(d), C.byref(n_ops), None, None, None, None,
                                    C.byref(n_terms), None, None, None, None, None))
    nnz = np.zeros(n_ops.value, np.int64)
    check(L.dmrgx_ed_model_describe(C.byref(mc), C.byref(d), C.byref(n_ops), nnz.ctypes.data, None, None, None,
                                    C.byref(n_terms), None, None, None, None, None))
    dd = d.value
    row = np.zeros(n_ops.value * (dd + 1), np.int64)
    col = np.zeros(max(1, int(nnz.sum())), np.int32)
    val = np.zeros(max(1, int(nnz.sum())))
    terms = np.zeros(n_terms.value, TERM_DTYPE)
    n_ele, qn1, qn2 = np.zeros(dd, np.int32), np.zeros(dd, np.int32), np.zeros(dd, np.int32)
    tg = np.zeros(3, np.int32)
    check(L.dmrgx_ed_model_describe(C.byref(mc), C.byref(d), C.byref(n_ops), nnz.ctypes.data, row.ctypes.data,
                                    col.ctypes.data, val.ctypes.data, C.byref(n_terms), terms.ctypes.data,
                                    n_ele.ctypes.data, qn1.ctypes.data, qn2.ctypes.data, tg.ctypes.data))
    ops, at = [], 0
    for k in range(n_ops.value):
        c = int(nnz[k])
        ops.append((row[k * (dd + 1):(k + 1) * (dd + 1)].copy(), col[at:at + c].copy(), val[at:at + c].copy()))
        at += c
    return ModelDescription(dd, ops, terms, n_ele, qn1, qn2, tg)


def default_exact_parameter(method: str = "Lanczos") -> EXACT_PARAMETER:
    """main/exact_diag/klm_vf/src/main.c:33-46"""
    P = EXACT_PARAMETER()
    P.diag_num = 1
    P.est_max_row_elem_num = 10000
    P.Diag_Method = method.encode()
    P.diag_acc = 1e-14
    P.diag_min_step = 1
    P.diag_max_step = 800
    P.cg_acc = 1e-7
    P.cg_max_step = 1000
    P.inv_iter_acc = 1e-9
    P.inv_iter_diag_add = 1e-11
    P.inv_iter_max_step = 3
    P.param_iter = 1
    return P


def exact_diagonalize(row, col, val, param: Optional[EXACT_PARAMETER] = None):
    """EXACT_DIAGONALIZE_HAMILTONIAN on a host full CRS -> (E, error, vec, time)."""
    H = HostCRS(row, col, val)
    P = param if param is not None else default_exact_parameter()
    vec = np.zeros(H.n)
    vp = (_D * 1)(vec.ctypes.data_as(_D))
    value = np.zeros(1)
    err = np.zeros(1)
    info = EXACT_HAM_INFO()
    info.Vector = C.cast(vp, C.POINTER(_D))
    info.Value = value.ctypes.data_as(_D)
    info.Error = err.ctypes.data_as(_D)
    info.Ham = H.ptr
    T = EXACT_TIME()
    lib().EXACT_DIAGONALIZE_HAMILTONIAN(C.byref(info), C.byref(P), C.byref(T), 1)
    return float(value[0]), float(err[0]), vec, T
