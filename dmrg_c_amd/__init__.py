"""Python view of libdmrg_c_amd.so, the MI355X superblock eigensolve.

The library is C ABI only (include/dmrg_c_amd.h).  This module mirrors the reference's
structs (include/SML.h, include/dmrg.h) with ctypes so that tests, the benchmark and
Python callers pass the same objects a C caller would.  There is no CPU fallback: if the
shared library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdmrg_c_amd.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

_lib: Optional[C.CDLL] = None


class LibraryMissing(RuntimeError):
    pass


class DmrgxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dmrgx error {code}: {msg}")
        self.code = code


# ------------------------------------------------------------------ structs
class CRS1(C.Structure):
    """include/SML.h:17-28"""
    _fields_ = [("max_val", C.c_long), ("max_row", C.c_int), ("row_dim", C.c_int),
                ("col_dim", C.c_int), ("Col", C.POINTER(C.c_int)), ("Row", C.POINTER(C.c_long)),
                ("Val", C.POINTER(C.c_double))]


class BOX_CG(C.Structure):
    """include/SML.h:56-65"""
    _fields_ = [("M", C.POINTER(CRS1)), ("acc", C.c_double), ("Vec", C.POINTER(C.c_double)),
                ("Out_Vec", C.POINTER(C.c_double)), ("max_step", C.c_int), ("p_threads", C.c_int)]


class BOX_II(C.Structure):
    """include/SML.h:67-86"""
    _fields_ = [("M", C.POINTER(CRS1)), ("ii_acc", C.c_double), ("ii_diag_add", C.c_double),
                ("ii_max_step", C.c_int), ("cg_acc", C.c_double), ("cg_max_step", C.c_int),
                ("error", C.c_double), ("Eig_Vec", C.POINTER(C.c_double)),
                ("Eig_Val", C.POINTER(C.c_double)), ("p_threads", C.c_int), ("Type", C.c_char * 100)]


class BOX_LAN(C.Structure):
    """include/SML.h:122-135"""
    _fields_ = [("M", C.POINTER(CRS1)), ("Eig_Vec", C.POINTER(C.c_double)),
                ("eig_val", C.POINTER(C.c_double)), ("acc", C.c_double), ("min_step", C.c_int),
                ("max_step", C.c_int), ("max_block_step", C.c_int), ("p_threads", C.c_int),
                ("Guess", C.c_char * 100), ("Type", C.c_char * 100)]


class DMRG_PARAMETER(C.Structure):
    """include/dmrg.h:6-38"""
    _fields_ = [("Initial_Guess", C.c_char * 1000), ("Enviro_Copy", C.c_char * 1000),
                ("II_Type", C.c_char * 1000), ("Lan_Con", C.c_char * 1000),
                ("max_dim_system", C.c_int), ("sweep", C.c_int), ("Diag_Method", C.c_char * 1000),
                ("diag_acc", C.c_double), ("diag_min_step", C.c_int), ("diag_max_step", C.c_int),
                ("cg_acc", C.c_double), ("cg_max_step", C.c_int), ("inv_iter_acc", C.c_double),
                ("inv_iter_diag_add", C.c_double), ("inv_iter_max_step", C.c_int),
                ("sp_LL", C.c_double), ("param_iter", C.c_int), ("param_iter_now", C.c_int),
                ("dim_LLLRRRRL_limit", C.c_int)]


class DMRG_TIME(C.Structure):
    """include/dmrg.h:40-51"""
    _fields_ = [(k, C.c_double) for k in ("diag", "inv_iter", "make_ham", "make_basis",
                                          "make_dens_mat", "trans_main", "trans_exp", "total")]


class DMRG_STATUS(C.Structure):
    """include/dmrg.h:236-267"""
    _fields_ = [("BC", C.c_char * 1000), ("Enviro_Copy", C.c_char * 100), ("LL_site", C.c_int),
                ("RR_site", C.c_int), ("max_dim_system", C.c_int), ("dim_LLLRRRRL", C.c_int),
                ("dim_LL", C.c_int), ("dim_onsite", C.c_int), ("dim_RR", C.c_int), ("spin", C.c_int),
                ("tot_sz", C.c_int), ("tot_ele", C.c_int), ("tot_ele_1", C.c_int),
                ("tot_ele_2", C.c_int), ("tot_parity", C.c_int), ("GS_Vec", C.POINTER(C.c_double)),
                ("gs_val", C.c_double), ("gs_error", C.c_double), ("percent_LL", C.c_double),
                ("sweep_now", C.c_int), ("param_iter_now", C.c_int), ("tot_iter_now", C.c_int),
                ("sweep", C.c_int), ("param_iter", C.c_int), ("tot_iter", C.c_int)]


class SolveParams(C.Structure):
    _fields_ = [("diag_acc", C.c_double), ("diag_min_step", C.c_int), ("diag_max_step", C.c_int),
                ("lan_type", C.c_int), ("guess", C.c_int), ("seed", C.c_uint64),
                ("ii_acc", C.c_double), ("ii_diag_add", C.c_double), ("ii_max_step", C.c_int),
                ("cg_acc", C.c_double), ("cg_max_step", C.c_int), ("skip_ii", C.c_int),
                ("precision", C.c_int), ("mixed_switch", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("eig_val", C.c_double), ("gs_error", C.c_double), ("lanczos_steps", C.c_int),
                ("ii_steps", C.c_int), ("cg_steps", C.c_int), ("spmv_calls", C.c_int),
                ("seconds", C.c_double), ("lanczos_fp32_steps", C.c_int)]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------------ loader
def lib() -> C.CDLL:
    """Load libdmrg_c_amd.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.dmrgx_last_error.restype = C.c_char_p
    L.dmrgx_init.argtypes = [C.c_int]
    L.dmrgx_set_seed.argtypes = [C.c_uint64]
    L.dmrgx_start_vector.argtypes = [i64, C.c_uint64, vp]
    L.dmrgx_matrix_from_arrays.argtypes = [i32, vp, vp, vp, C.c_int, C.POINTER(vp)]
    L.dmrgx_matrix_free.argtypes = [vp]
    L.dmrgx_matrix_nnz.argtypes = [vp]
    L.dmrgx_matrix_nnz.restype = i64
    L.dmrgx_matrix_dim.argtypes = [vp]
    L.dmrgx_matrix_dim.restype = i32
    L.dmrgx_matrix_download.argtypes = [vp, vp, vp, vp]
    L.dmrgx_spmv_device.argtypes = [vp, vp, vp]
    L.dmrgx_spmv_host.argtypes = [vp, vp, vp]
    L.dmrgx_default_params.argtypes = [C.POINTER(SolveParams)]
    L.dmrgx_ground_state.argtypes = [vp, C.POINTER(SolveParams), vp, C.POINTER(SolveStats)]
    L.dmrgx_matrix_set_schedule.argtypes = [vp, i32, i32, i32]
    L.dmrgx_matrix_get_schedule.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.dmrgx_matrix_autotune.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(f64)]
    L.dmrgx_matrix_time_schedule.argtypes = [vp, i32, i32, i32, i32, C.POINTER(f64)]
    L.MATRIX_VECTOR_PRODUCT_SYM.argtypes = [C.POINTER(CRS1), vp, vp, vp, C.c_int]
    L.MATRIX_VECTOR_PRODUCT.argtypes = [C.POINTER(CRS1), vp, vp, C.c_int]
    for nm in ("LANCZOS_SLOW_SYM", "LANCZOS_SYM", "LANCZOS_SLOW", "LANCZOS"):
        getattr(L, nm).argtypes = [C.POINTER(BOX_LAN)]
    for nm in ("INVERSE_ITERATION_SYM", "INVERSE_ITERATION"):
        getattr(L, nm).argtypes = [C.POINTER(BOX_II)]
    for nm in ("CONJUGATE_GRADIENT_SYM", "CONJUGATE_GRADIENT"):
        getattr(L, nm).argtypes = [C.POINTER(BOX_CG)]
    for nm in ("CONVERGE_CHECK_SYM", "CONVERGE_CHECK"):
        getattr(L, nm).argtypes = [C.POINTER(CRS1), vp, f64, C.c_int]
        getattr(L, nm).restype = f64
    for nm in ("DMRG_DIAGONALIZE_SUPERBLOCK_SYM", "DMRG_DIAGONALIZE_SUPERBLOCK"):
        getattr(L, nm).argtypes = [C.POINTER(CRS1), C.POINTER(DMRG_TIME), C.POINTER(DMRG_PARAMETER),
                                   C.POINTER(DMRG_STATUS), C.c_int]
    _lib = L
    return L


def check(code: int) -> None:
    if code != 0:
        raise DmrgxError(code, lib().dmrgx_last_error().decode())


# ------------------------------------------------------------------ helpers
class HostCRS:
    """A CRS1 over numpy arrays (kept alive by this object)."""

    def __init__(self, row, col, val, n: Optional[int] = None):
        self.row = np.ascontiguousarray(row, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        n = len(self.row) - 1 if n is None else n
        self.n = n
        self.crs = CRS1(len(self.val), n + 1, n, n,
                        self.col.ctypes.data_as(C.POINTER(C.c_int)),
                        self.row.ctypes.data_as(C.POINTER(C.c_long)), _dp(self.val))

    @property
    def ptr(self):
        return C.pointer(self.crs)


class DeviceMatrix:
    """Full CSR resident on the GPU (dmrgx_matrix)."""

    def __init__(self, row, col, val, sym: bool = True):
        L = lib()
        self._keep = HostCRS(row, col, val)
        h = C.c_void_p()
        check(L.dmrgx_matrix_from_arrays(self._keep.n, self._keep.row.ctypes.data,
                                         self._keep.col.ctypes.data, self._keep.val.ctypes.data,
                                         1 if sym else 0, C.byref(h)))
        self.handle = h
        self.n = self._keep.n
        self._keep = None

    @classmethod
    def _wrap(cls, handle, n):
        obj = cls.__new__(cls)
        obj.handle = handle
        obj.n = n
        obj._keep = None
        return obj

    @property
    def nnz(self) -> int:
        return int(lib().dmrgx_matrix_nnz(self.handle))

    def download(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        row = np.empty(self.n + 1, np.int64)
        col = np.empty(self.nnz, np.int32)
        val = np.empty(self.nnz)
        check(lib().dmrgx_matrix_download(self.handle, row.ctypes.data, col.ctypes.data, val.ctypes.data))
        return row, col, val

    def matvec(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        check(lib().dmrgx_spmv_host(self.handle, x.ctypes.data, y.ctypes.data))
        return y

    @property
    def schedule(self) -> Tuple[int, int, int]:
        """(lanes per row, workgroup cap, load variant) of this matrix's SpMV; load variant
        bit 0 = two entries per lane per load (CSR), bit 1 = eight entries in flight per lane,
        bit 2 = contiguous row ranges per workgroup and per XCD, bits 3-5 = storage format
        (0 CSR, 1 value dictionary, 2 delta columns, 3 sliced, 4 sliced with 16-bit columns
        and 16-bit value indices; for the sliced forms bit 1 = non-temporal matrix loads,
        and for format 3 bit 0 = 16-bit column differences)."""
        g, cap, nt = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().dmrgx_matrix_get_schedule(self.handle, C.byref(g), C.byref(cap), C.byref(nt)))
        return g.value, cap.value, nt.value

    def set_schedule(self, lanes: int, grid_cap: int, load_variant: int = 0) -> None:
        check(lib().dmrgx_matrix_set_schedule(self.handle, lanes, grid_cap, load_variant))

    def autotune(self) -> Tuple[Tuple[int, int, int], float]:
        """Time the schedule candidates on the device, keep the fastest; returns it and its
        time per Lanczos step + CG iteration (microseconds)."""
        g, cap, nt, us = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        check(lib().dmrgx_matrix_autotune(self.handle, C.byref(g), C.byref(cap), C.byref(nt), C.byref(us)))
        return (g.value, cap.value, nt.value), us.value

    def sell16_exceptions(self, lanes: int) -> int:
        """Columns the sliced form's 16-bit column form (lanes per row 1, 2 or 4) keeps whole."""
        e = C.c_int64()
        L = lib()
        L.dmrgx_matrix_sell16_exceptions.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
        check(L.dmrgx_matrix_sell16_exceptions(self.handle, lanes, C.byref(e)))
        return e.value

    def time_schedule(self, lanes: int, grid_cap: int, load_variant: int = 0, reps: int = 10) -> float:
        """Mean microseconds of one product + fused dot under that schedule."""
        us = C.c_double()
        check(lib().dmrgx_matrix_time_schedule(self.handle, lanes, grid_cap, load_variant, reps, C.byref(us)))
        return us.value

    def ground_state(self, start: Optional[np.ndarray] = None, **kw):
        p = SolveParams()
        lib().dmrgx_default_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        vec = np.zeros(self.n)
        if start is not None:
            p.guess = 1
            vec[:] = start
        st = SolveStats()
        check(lib().dmrgx_ground_state(self.handle, C.byref(p), vec.ctypes.data, C.byref(st)))
        return st.eig_val, vec, st

    def free(self):
        if getattr(self, "handle", None):
            lib().dmrgx_matrix_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def start_vector(n: int, seed: int) -> np.ndarray:
    out = np.empty(n)
    lib().dmrgx_start_vector(n, seed, out.ctypes.data)
    return out


def default_dmrg_parameter(method="Lanczos", lan_con="Normal", diag_acc=1e-14, cg_acc=1e-7,
                           cg_max_step=2000) -> DMRG_PARAMETER:
    """Defaults of main/dmrg/hubbard_vf/src/main/main.c:16-49."""
    P = DMRG_PARAMETER()
    P.Initial_Guess = b"No"
    P.Enviro_Copy = b"Yes"
    P.II_Type = b"CG"
    P.Lan_Con = lan_con.encode()
    P.Diag_Method = method.encode()
    P.diag_acc = diag_acc
    P.diag_min_step = 1
    P.diag_max_step = 800
    P.cg_acc = cg_acc
    P.cg_max_step = cg_max_step
    P.inv_iter_acc = 1e-9
    P.inv_iter_diag_add = 1e-11
    P.inv_iter_max_step = 3
    P.dim_LLLRRRRL_limit = 2 ** 31 - 1
    return P


def diagonalize_superblock_sym(row, col, val, param: Optional[DMRG_PARAMETER] = None):
    """DMRG_DIAGONALIZE_SUPERBLOCK_SYM on a host lower-triangle CRS -> (E, gs_error, vec, time)."""
    L = lib()
    H = HostCRS(row, col, val)
    P = param if param is not None else default_dmrg_parameter()
    st = DMRG_STATUS()
    vec = np.zeros(H.n)
    st.GS_Vec = _dp(vec)
    t = DMRG_TIME()
    L.DMRG_DIAGONALIZE_SUPERBLOCK_SYM(H.ptr, C.byref(t), C.byref(P), C.byref(st), 1)
    return st.gs_val, st.gs_error, vec, t


# ------------------------------------------------------------ superblock assembly
class DMRG_TERM(C.Structure):
    """dmrgx_term of include/dmrg_c_amd.h"""
    _fields_ = [(k, C.c_int32) for k in ("kind", "op1", "op2", "ele0", "ele1", "ele2", "sign_mode", "reserved")] + \
               [("coeff", C.c_double)]


class DMRG_BASIS_ONSITE(C.Structure):
    """include/dmrg.h:198-208"""
    _fields_ = [(k, C.c_int) for k in ("LL", "LR", "RR", "RL", "row", "dim_RR", "dim_onsite")]


class DMRG_A_BASIS(C.Structure):
    """include/dmrg.h:128-138"""
    _fields_ = [("LL_LLLRRRRL", C.POINTER(C.c_short)), ("LR_LLLRRRRL", C.POINTER(C.c_short)),
                ("RL_LLLRRRRL", C.POINTER(C.c_short)), ("RR_LLLRRRRL", C.POINTER(C.c_short)),
                ("Inv_LLLRRRRL", C.POINTER(C.c_int)), ("Val_LLLRRRRL", C.POINTER(C.c_double)),
                ("elem_num", C.c_int)]


_S = C.POINTER(C.c_short)
_I = C.POINTER(C.c_int)
_II = C.POINTER(C.POINTER(C.c_int))


class DMRG_BASIS(C.Structure):
    """include/dmrg.h:54-99"""
    _fields_ = [("LL_LLLRRRRL", _S), ("LR_LLLRRRRL", _S), ("RL_LLLRRRRL", _S), ("RR_LLLRRRRL", _S),
                ("Inv_LLLRRRRL", _I), ("dim_LLLRRRRL", C.c_int), ("dim_LL", C.c_int), ("dim_RR", C.c_int),
                ("dim_onsite", C.c_int), ("tot_sz_LLLRRRRL", C.c_int), ("tot_ele_LLLRRRRL", C.c_int),
                ("tot_ele_1_LLLRRRRL", C.c_int), ("tot_ele_2_LLLRRRRL", C.c_int),
                ("tot_parity_LLLRRRRL", C.c_int), ("LL_LLLR", _S), ("LR_LLLR", _S), ("Inv_LLLR", _II),
                ("dim_LLLR", C.c_int), ("LR_LRRL", _S), ("RL_LRRL", _S), ("Inv_LRRL", _II), ("dim_LRRL", C.c_int),
                ("RR_RRRL", _S), ("RL_RRRL", _S), ("Inv_RRRL", _II), ("dim_RRRL", C.c_int), ("Sum_LLLR", _I),
                ("Tot_Sz_LLLR", _I), ("Tot_Ele_LLLR", _I), ("Tot_Ele_1_LLLR", _I), ("Tot_Ele_2_LLLR", _I),
                ("Tot_Parity_LLLR", _I)]


class DMRG_BOX_LAN(C.Structure):
    """include/dmrg.h:269-285"""
    _fields_ = [("Eig_Vec", C.POINTER(C.c_double)), ("eig_val", C.POINTER(C.c_double)), ("acc", C.c_double),
                ("min_step", C.c_int), ("max_step", C.c_int), ("p_threads", C.c_int), ("Guess", C.c_char * 100),
                ("Type", C.c_char * 100), ("M_LLLR", C.POINTER(CRS1)), ("M_LRRL", C.POINTER(CRS1)),
                ("M_LRRL_Sign", C.POINTER(CRS1)), ("M_RRRL", C.POINTER(CRS1)), ("Ele_RR", _I),
                ("Dmrg_Basis", C.POINTER(DMRG_BASIS))]


class DMRG_BOX_CG(C.Structure):
    """include/dmrg.h:287-302"""
    _fields_ = [("acc", C.c_double), ("diag_val", C.c_double), ("Vec", C.POINTER(C.c_double)),
                ("Out_Vec", C.POINTER(C.c_double)), ("max_step", C.c_int), ("p_threads", C.c_int),
                ("M_LLLR", C.POINTER(CRS1)), ("M_LRRL", C.POINTER(CRS1)), ("M_LRRL_Sign", C.POINTER(CRS1)),
                ("M_RRRL", C.POINTER(CRS1)), ("Ele_RR", _I), ("Dmrg_Basis", C.POINTER(DMRG_BASIS))]


class DMRG_BOX_II(C.Structure):
    """include/dmrg.h:304-326"""
    _fields_ = [("ii_acc", C.c_double), ("ii_diag_add", C.c_double), ("ii_max_step", C.c_int),
                ("cg_acc", C.c_double), ("cg_max_step", C.c_int), ("Eig_Vec", C.POINTER(C.c_double)),
                ("eig_val", C.POINTER(C.c_double)), ("error", C.c_double), ("p_threads", C.c_int),
                ("M_LLLR", C.POINTER(CRS1)), ("M_LRRL", C.POINTER(CRS1)), ("M_LRRL_Sign", C.POINTER(CRS1)),
                ("M_RRRL", C.POINTER(CRS1)), ("Ele_RR", _I), ("Dmrg_Basis", C.POINTER(DMRG_BASIS))]


def _bind_superblock(L):
    vp = C.c_void_p
    L.dmrgx_superblock_create_arrays.argtypes = [C.c_int32] * 4 + [vp] * 4 + [C.c_int32, vp, C.c_int32, vp, vp,
                                                                              C.c_int32, vp, C.POINTER(vp)]
    L.dmrgx_superblock_create.argtypes = [C.POINTER(DMRG_BASIS), vp, C.POINTER(vp)]
    L.dmrgx_superblock_basis.argtypes = [vp, C.POINTER(C.c_int64), vp, vp, vp, vp]
    L.dmrgx_superblock_create_qn.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, vp, C.c_int32, vp, C.POINTER(vp)]
    L.dmrgx_superblock_n.argtypes = [vp, C.POINTER(C.c_int32)]
    L.dmrgx_superblock_assemble.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(vp)]
    L.dmrgx_superblock_download_lower.argtypes = [vp, vp, vp, vp]
    L.dmrgx_superblock_free.argtypes = [vp]
    L.dmrgx_record_begin.argtypes = []
    L.dmrgx_record_end.restype = vp
    L.dmrgx_terms_free.argtypes = [vp]
    L.dmrgx_terms_count.argtypes = [vp]
    L.dmrgx_get_ham_lllrrrrl.argtypes = [C.POINTER(DMRG_BASIS), vp]
    L.dmrgx_get_ham_lllrrrrl.restype = C.POINTER(CRS1)
    L.DMRG_MATRIX_VECTOR_PRODUCT_OBC.argtypes = [C.POINTER(CRS1)] * 4 + [vp, vp, _I, C.c_int, C.POINTER(DMRG_BASIS)]
    L.DMRG_CONVERGE_CHECK_OBC.argtypes = [C.c_double] + [C.POINTER(CRS1)] * 4 + [vp, vp, _I, C.c_int,
                                                                                C.POINTER(DMRG_BASIS)]
    L.DMRG_CONVERGE_CHECK_OBC.restype = C.c_double
    L.DMRG_LANCZOS_SLOW_OBC.argtypes = [C.POINTER(DMRG_BOX_LAN)]
    L.DMRG_CONJUGATE_GRADIENT_OBC.argtypes = [C.POINTER(DMRG_BOX_CG)]
    L.DMRG_INVERSE_ITERATION_OBC.argtypes = [C.POINTER(DMRG_BOX_II)]
    elem1 = [C.POINTER(DMRG_BASIS_ONSITE), C.POINTER(DMRG_A_BASIS), _I, C.POINTER(CRS1), C.c_double, _I]
    for k in ("LL", "LR", "RR", "RL"):
        getattr(L, f"DMRG_MAKE_ELEM_{k}_LLLRRRRL").argtypes = elem1
    two = [C.POINTER(DMRG_BASIS_ONSITE), C.POINTER(DMRG_A_BASIS), _I, C.POINTER(CRS1), C.POINTER(CRS1), C.c_double]
    L.DMRG_MAKE_ELEM_LLLR_LLLRRRRL.argtypes = two + [_I, _I, C.c_char_p, C.c_char_p]
    L.DMRG_MAKE_ELEM_LLRL_LLLRRRRL.argtypes = two + [_I, _I, _I, _I, C.c_char_p, C.c_char_p]
    L.DMRG_MAKE_ELEM_LLRR_LLLRRRRL.argtypes = two + [_I, _I, _I, C.c_char_p, C.c_char_p]
    L.DMRG_MAKE_ELEM_LRRL_LLLRRRRL.argtypes = two + [_I, _I, _I, C.c_char_p, C.c_char_p]
    L.DMRG_MAKE_ELEM_LRRR_LLLRRRRL.argtypes = two + [_I, _I, C.c_char_p, C.c_char_p]
    L.DMRG_MAKE_ELEM_RRRL_LLLRRRRL.argtypes = two + [_I, _I, C.c_char_p, C.c_char_p]
    L.DMRG_MAKE_ELEM_ZERO_LLLRRRRL.argtypes = [C.POINTER(DMRG_BASIS_ONSITE), C.POINTER(DMRG_A_BASIS), _I, _I]


_orig_lib = lib


def lib() -> C.CDLL:  # noqa: F811  (extends the loader with the superblock bindings)
    L = _orig_lib()
    if not getattr(L, "_sb_bound", False):
        _bind_superblock(L)
        L._sb_bound = True
    return L


class BasisQN(C.Structure):
    """dmrgx_basis_qn: the blocks' quantum numbers a superblock basis is built from."""
    _fields_ = [(k, C.c_int32) for k in ("n_qn", "dim_LL", "dim_RR", "dim_onsite", "target_sz", "target_ele")] + \
               [(k, C.c_void_p) for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR", "sz_site", "ele_site")]


def _basis_qn(q) -> Tuple[BasisQN, list]:
    """BasisQN from a dict (n_qn, sz_LL, ele_LL, sz_RR, ele_RR, sz_site, ele_site, target_sz,
    target_ele); returns the struct and the arrays it points to (keep them alive)."""
    arrs = {k: np.ascontiguousarray(q[k], np.int32) for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR", "sz_site",
                                                            "ele_site")}
    b = BasisQN(int(q["n_qn"]), len(arrs["sz_LL"]), len(arrs["sz_RR"]), len(arrs["sz_site"]), int(q["target_sz"]),
                int(q["target_ele"]), *[arrs[k].ctypes.data for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR",
                                                                         "sz_site", "ele_site")])
    return b, list(arrs.values())


def superblock_basis(q) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """dmrgx_superblock_basis: (LL, LR, RR, RL) of the target sector, built on the device."""
    L = lib()
    b, keep = _basis_qn(q)
    n = C.c_int64()
    check(L.dmrgx_superblock_basis(C.byref(b), C.byref(n), None, None, None, None))
    out = [np.empty(n.value, np.int16) for _ in range(4)]
    check(L.dmrgx_superblock_basis(C.byref(b), C.byref(n), *[o.ctypes.data for o in out]))
    return tuple(out)


def _host_crs1(row, col, val) -> Tuple[CRS1, list]:
    """CRS1 over numpy arrays (kept alive by the returned list)."""
    r = np.ascontiguousarray(row, np.int64)
    c = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    m = CRS1(len(v), len(r) - 1, len(r) - 1, len(r) - 1, c.ctypes.data_as(C.POINTER(C.c_int)),
             r.ctypes.data_as(C.POINTER(C.c_long)), v.ctypes.data_as(C.POINTER(C.c_double)))
    return m, [r, c, v]


def density_matrices(B: "DMRG_BASIS", dim_block, vec: np.ndarray) -> List[np.ndarray]:
    """dmrgx_density_matrices: the reduced density matrix of every LLLR block of `B`
    (DMRG_GET_DENSITY_MATRIX for blocks 0, 1, ...), formed on the device."""
    L = lib()
    if not getattr(L, "_dm_bound", False):
        L.dmrgx_density_matrices.argtypes = [C.POINTER(DMRG_BASIS), C.c_int32, C.c_void_p, C.c_void_p,
                                             C.POINTER(C.c_int64), C.c_void_p]
        L._dm_bound = True
    db = np.ascontiguousarray(dim_block, np.int32)
    v = np.ascontiguousarray(vec, np.float64)
    n = C.c_int64()
    check(L.dmrgx_density_matrices(C.byref(B), len(db), db.ctypes.data, None, C.byref(n), None))
    out = np.empty(n.value)
    check(L.dmrgx_density_matrices(C.byref(B), len(db), db.ctypes.data, v.ctypes.data, C.byref(n),
                                   out.ctypes.data))
    mats, off = [], 0
    for d in db:
        mats.append(out[off:off + d * d].reshape(d, d))
        off += d * d
    return mats


class CCS1(C.Structure):
    """include/SML.h:43-54"""
    _fields_ = [("max_val", C.c_long), ("max_col", C.c_int), ("row_dim", C.c_int), ("col_dim", C.c_int),
                ("Col", C.POINTER(C.c_long)), ("Row", C.POINTER(C.c_int)), ("Val", C.POINTER(C.c_double))]


class DMRG_SYSTEM_INFO(C.Structure):
    """include/dmrg.h:220-234"""
    _fields_ = [("Trans_Matrix", C.POINTER(CCS1)), ("Trans_Matrix_Dagger", C.POINTER(CRS1)),
                ("tr_error", C.c_double), ("Val_DM_Dist", C.POINTER(C.c_double)), ("sum_val_dm", C.c_double),
                ("dim_LLLR", C.c_int), ("dim_renorm", C.c_int), ("Q_Number1", C.POINTER(C.c_int)),
                ("Q_Number2", C.POINTER(C.c_int)), ("Q_Number3", C.POINTER(C.c_int))]


def system_info_dict(S: DMRG_SYSTEM_INFO, n_qn: int) -> dict:
    """Host copy of a DMRG_SYSTEM_INFO: the dense transformation matrix T (dim_LLLR x
    dim_renorm) from its CCS1, and its CRS1 transpose checked against it."""
    m, k = S.dim_LLLR, S.dim_renorm
    T = np.zeros((m, k))
    A = S.Trans_Matrix.contents
    for c in range(k):
        for e in range(A.Col[c], A.Col[c + 1]):
            T[A.Row[e], c] = A.Val[e]
    Ad = S.Trans_Matrix_Dagger.contents
    Td = np.zeros((k, m))
    for r in range(k):
        for e in range(Ad.Row[r], Ad.Row[r + 1]):
            Td[r, Ad.Col[e]] = Ad.Val[e]
    return {"dim_LLLR": m, "dim_renorm": k, "tr_error": S.tr_error, "sum_val_dm": S.sum_val_dm,
            "val_dist": np.array(S.Val_DM_Dist[:m]), "q1": np.array(S.Q_Number1[:k]),
            "q2": np.array(S.Q_Number2[:k]) if n_qn == 2 else None, "T": T, "T_dagger_is_transpose":
            bool(np.array_equal(Td, T.T)), "shapes": (A.row_dim, A.col_dim, Ad.row_dim, Ad.col_dim)}


def system_info(B: "DMRG_BASIS", n_qn: int, q1, q2, vec: np.ndarray, max_dim: int) -> dict:
    """dmrgx_get_system_info (DMRG_GET_SYSTEM_INFO_Q1/Q2 on the device) as a dict."""
    L = lib()
    if not getattr(L, "_si_bound", False):
        L.dmrgx_get_system_info.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(DMRG_BASIS), C.c_void_p,
                                            C.c_int32, C.POINTER(C.POINTER(DMRG_SYSTEM_INFO))]
        L.dmrgx_free_system_info.argtypes = [C.POINTER(DMRG_SYSTEM_INFO)]
        L._si_bound = True
    a1 = np.ascontiguousarray(q1, np.int32)
    a2 = np.ascontiguousarray(q2 if q2 is not None else q1, np.int32)
    v = np.ascontiguousarray(vec, np.float64)
    out = C.POINTER(DMRG_SYSTEM_INFO)()
    check(L.dmrgx_get_system_info(n_qn, a1.ctypes.data, a2.ctypes.data, C.byref(B), v.ctypes.data, max_dim,
                                  C.byref(out)))
    try:
        return system_info_dict(out.contents, n_qn)
    finally:
        L.dmrgx_free_system_info(out)


def set_block_eigensolver(mode: int) -> None:
    """dmrgx_set_block_eigensolver: the truncation's eigensolver for blocks up to 192 states,
    0 the library's own (default), 1 rocSOLVER dsyevd.  Process-wide."""
    L = lib()
    L.dmrgx_set_block_eigensolver.argtypes = [C.c_int]
    check(L.dmrgx_set_block_eigensolver(mode))


class WholeBasis:
    """dmrgx_whole_basis: the whole-space inverse table of DMRG_WHOLE_BASIS_Q1..Q4 on the
    device plus the sectors registered with add_sector(); v_m() is DMRG_V_M_*_Q1..Q4."""

    def __init__(self, dim_LL: int, dim_RR: int, dim_onsite: int, inv: np.ndarray):
        L = lib()
        if not getattr(L, "_wb_bound", False):
            vp = C.c_void_p
            L.dmrgx_whole_basis_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, C.POINTER(vp)]
            L.dmrgx_whole_basis_add_sector.argtypes = [vp, C.c_int32, vp, vp, vp, vp, C.POINTER(C.c_int32)]
            L.dmrgx_v_m_sector.argtypes = [vp, C.c_int32, C.POINTER(CRS1), C.c_int32, C.c_int32, vp, vp, vp, vp,
                                           C.c_int32, vp]
            L.dmrgx_whole_basis_free.argtypes = [vp]
            L._wb_bound = True
        inv = np.ascontiguousarray(inv, np.int32)
        if inv.size != dim_LL * dim_onsite * dim_RR * dim_onsite:
            raise ValueError("inverse table size is not dim_LL * d * dim_RR * d")
        h = C.c_void_p()
        check(L.dmrgx_whole_basis_create(dim_LL, dim_RR, dim_onsite, inv.ctypes.data, C.byref(h)))
        self.handle = h
        self.dims = (dim_LL, dim_onsite, dim_RR, dim_onsite)
        self.sector_dims: List[int] = []

    def add_sector(self, LL, LR, RR, RL) -> int:
        arrs = [np.ascontiguousarray(a, np.int16) for a in (LL, LR, RR, RL)]
        sid = C.c_int32()
        check(lib().dmrgx_whole_basis_add_sector(self.handle, len(arrs[0]), *[a.ctypes.data for a in arrs],
                                                 C.byref(sid)))
        self.sector_dims.append(len(arrs[0]))
        return sid.value

    def v_m(self, slot: int, op, sector_out: int, x: np.ndarray, ele_LL=None, ele_LR=None, ele_RR=None,
            sign: bool = False) -> np.ndarray:
        """op = (row, col, val) of the slot operator; x: the input vector (its sector's length)."""
        m, keep = _host_crs1(*op)
        x = np.ascontiguousarray(x, np.float64)
        eles = [None if e is None else np.ascontiguousarray(e, np.int32) for e in (ele_LL, ele_LR, ele_RR)]
        out = np.empty(self.sector_dims[sector_out] if 0 <= sector_out < len(self.sector_dims) else 0)
        check(lib().dmrgx_v_m_sector(self.handle, slot, C.byref(m), sector_out, len(x), x.ctypes.data,
                                     *[None if e is None else e.ctypes.data for e in eles], 1 if sign else 0,
                                     out.ctypes.data))
        return out

    def intersite(self, sector_in: int, sector_out1: int, sector_out2: int, M_CF, M_LL, M_On, M_RR, origin: int,
                  LL_site: int, RR_site: int, x: np.ndarray) -> np.ndarray:
        """dmrgx_expectation_intersite_sector; operators as (row, col, val) tuples, M_LL a
        dict or list indexed by site (only M_LL[origin] is used)."""
        L = lib()
        P = C.POINTER(CRS1)
        L.dmrgx_expectation_intersite_sector.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                                         C.c_void_p, P, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                                         C.c_void_p, C.c_void_p]
        keep = []

        def arr(ops):
            ms = []
            for o in ops:
                m, k = _host_crs1(*o)
                keep.append(k)
                ms.append(m)
            a = (CRS1 * max(len(ms), 1))(*ms)
            keep.append(a)
            return (P * max(len(ms), 1))(*[C.pointer(a[i]) for i in range(len(ms))])
        cf = arr(M_CF)
        ll = None
        if origin <= LL_site:
            ll = arr([M_LL[origin] if i == origin else M_LL[origin] for i in range(origin + 1)])
        on, kon = _host_crs1(*M_On)
        rr = arr(M_RR)
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(max(RR_site + LL_site + 4 - origin, 1))
        check(L.dmrgx_expectation_intersite_sector(self.handle, sector_in, sector_out1, sector_out2, C.cast(cf, C.c_void_p),
                                                   None if ll is None else C.cast(ll, C.c_void_p), C.byref(on),
                                                   C.cast(rr, C.c_void_p), origin, LL_site, RR_site, x.ctypes.data,
                                                   out.ctypes.data))
        return out

    def free(self):
        if getattr(self, "handle", None):
            lib().dmrgx_whole_basis_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Superblock:
    """Device-resident superblock inputs (basis, operators, term list) from a Snapshot; with
    qn given, the basis is built on the device from the blocks' quantum numbers instead of
    taken from the snapshot."""

    def __init__(self, snap, qn=None):
        L = lib()
        self.snap = snap
        self._ops = [HostCRS(o.row, o.col, o.val, o.row_dim) for o in snap.ops]
        for h, o in zip(self._ops, snap.ops):
            h.crs.col_dim = o.col_dim
        ops = (CRS1 * max(1, len(self._ops)))(*[h.crs for h in self._ops])
        self._eles = [np.ascontiguousarray(e, np.int32) for e in snap.eles]
        eptr = (C.c_void_p * max(1, len(self._eles)))(*[e.ctypes.data for e in self._eles])
        elen = np.array([len(e) for e in self._eles] or [0], np.int32)
        terms = (DMRG_TERM * max(1, len(snap.terms)))()
        for i, t in enumerate(snap.terms):
            terms[i] = DMRG_TERM(int(t["kind"]), int(t["op1"]), int(t["op2"]), int(t["ele0"]), int(t["ele1"]),
                                 int(t["ele2"]), int(t["sign_mode"]), 0, float(t["coeff"]))
        h = C.c_void_p()
        if qn is None:
            self._basis = [np.ascontiguousarray(a, np.int16) for a in (snap.LL, snap.LR, snap.RR, snap.RL)]
            check(L.dmrgx_superblock_create_arrays(snap.n, snap.dim_LL, snap.dim_RR, snap.dim_onsite,
                                                   *[a.ctypes.data for a in self._basis], len(self._ops),
                                                   C.cast(ops, C.c_void_p), len(self._eles),
                                                   C.cast(eptr, C.c_void_p), elen.ctypes.data, len(snap.terms),
                                                   C.cast(terms, C.c_void_p), C.byref(h)))
            self.n = snap.n
        else:
            b, self._qn_keep = _basis_qn(qn)
            check(L.dmrgx_superblock_create_qn(C.byref(b), len(self._ops), C.cast(ops, C.c_void_p), len(self._eles),
                                               C.cast(eptr, C.c_void_p), elen.ctypes.data, len(snap.terms),
                                               C.cast(terms, C.c_void_p), C.byref(h)))
            n = C.c_int32()
            check(L.dmrgx_superblock_n(h, C.byref(n)))
            self.n = n.value
        self.handle = h
        self.nnz_lower = -1

    def assemble(self, full: bool = True) -> Optional[DeviceMatrix]:
        nnz = C.c_int64()
        A = C.c_void_p()
        check(lib().dmrgx_superblock_assemble(self.handle, C.byref(nnz), C.byref(A) if full else None))
        self.nnz_lower = nnz.value
        return DeviceMatrix._wrap(A, self.n) if full else None

    def kron(self) -> "KronOperator":
        """The Kronecker form of this superblock's Hamiltonian (dmrgx_superblock_kron); the
        superblock must outlive it."""
        return KronOperator(self)

    def download_lower(self):
        row = np.empty(self.n + 1, np.int64)
        col = np.empty(max(self.nnz_lower, 1), np.int32)
        val = np.empty(max(self.nnz_lower, 1))
        check(lib().dmrgx_superblock_download_lower(self.handle, row.ctypes.data, col.ctypes.data, val.ctypes.data))
        return row, col[:self.nnz_lower], val[:self.nnz_lower]

    def free(self):
        if getattr(self, "handle", None):
            lib().dmrgx_superblock_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class KronOperator:
    """dmrgx_kron: the superblock Hamiltonian applied per sector as M_L X + X M_R^T plus the
    coupling terms, without an assembled CRS."""

    def __init__(self, sb: "Superblock"):
        L = lib()
        vp = C.c_void_p
        L.dmrgx_superblock_kron.argtypes = [vp, C.POINTER(vp)]
        L.dmrgx_kron_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_double)]
        L.dmrgx_kron_spmv.argtypes = [vp, vp, vp]
        L.dmrgx_kron_set_kernel.argtypes = [vp, C.c_int32]
        L.dmrgx_kron_autotune.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.dmrgx_kron_ground_state.argtypes = [vp, C.POINTER(SolveParams), vp, C.POINTER(SolveStats)]
        L.dmrgx_kron_free.argtypes = [vp]
        L.dmrgx_kron_free.restype = None
        self._sb = sb   # the superblock outlives the form
        h = vp()
        check(L.dmrgx_superblock_kron(sb.handle, C.byref(h)))
        self.handle = h
        n, ns, nc, fl = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        check(L.dmrgx_kron_info(h, C.byref(n), C.byref(ns), C.byref(nc), C.byref(fl)))
        self.n, self.sectors, self.coupling_terms, self.flops = n.value, ns.value, nc.value, fl.value

    VARIANTS = 16

    def set_kernel(self, variant: int):
        check(lib().dmrgx_kron_set_kernel(self.handle, variant))

    def autotune(self):
        v, us = C.c_int32(), C.c_double()
        check(lib().dmrgx_kron_autotune(self.handle, C.byref(v), C.byref(us)))
        return v.value, us.value

    def matvec(self, x: np.ndarray) -> np.ndarray:
        xx = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n)
        check(lib().dmrgx_kron_spmv(self.handle, xx.ctypes.data, y.ctypes.data))
        return y

    def ground_state(self, start: Optional[np.ndarray] = None, **kw):
        p = SolveParams()
        lib().dmrgx_default_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        vec = np.zeros(self.n)
        if start is not None:
            p.guess = 1
            vec[:] = start
        st = SolveStats()
        check(lib().dmrgx_kron_ground_state(self.handle, C.byref(p), vec.ctypes.data, C.byref(st)))
        return st.eig_val, vec, st

    def free(self):
        if getattr(self, "handle", None):
            lib().dmrgx_kron_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
