"""Batched / continuously batched ground-state solves (dmrgx_engine_* of libdmrg_c_amd.so)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import DeviceMatrix, SolveParams, SolveStats, check, lib as _lib


class EngineOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active", C.c_int32), ("chunk_ticks", C.c_int32),
                ("basis_steps_estimate", C.c_int32), ("use_graphs", C.c_int32)]


class EngineResult(C.Structure):
    _fields_ = [("request_id", C.c_int64), ("tag", C.c_uint64), ("status", C.c_int32), ("n", C.c_int32),
                ("eig_val", C.c_double), ("gs_error", C.c_double), ("lanczos_steps", C.c_int32),
                ("ii_steps", C.c_int32), ("cg_steps", C.c_int32), ("spmv_calls", C.c_int32),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


_bound = False


def lib() -> C.CDLL:
    global _bound
    L = _lib()
    if not _bound:
        vp = C.c_void_p
        L.dmrgx_engine_default_opts.argtypes = [C.POINTER(EngineOpts)]
        L.dmrgx_engine_create.argtypes = [C.POINTER(EngineOpts), C.POINTER(vp)]
        L.dmrgx_engine_submit.argtypes = [vp, vp, C.c_int, C.POINTER(SolveParams), C.c_uint64, C.POINTER(C.c_int64)]
        L.dmrgx_engine_run.argtypes = [vp, C.c_int32]
        L.dmrgx_engine_pending.argtypes = [vp]
        L.dmrgx_engine_pending.restype = C.c_int32
        L.dmrgx_engine_done.argtypes = [vp]
        L.dmrgx_engine_done.restype = C.c_int32
        L.dmrgx_engine_peek.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_int32)]
        L.dmrgx_engine_poll.argtypes = [vp, C.POINTER(EngineResult), vp]
        L.dmrgx_engine_free.argtypes = [vp]
        L.dmrgx_ground_state_batch.argtypes = [C.c_int32, vp, C.POINTER(SolveParams), vp, C.POINTER(SolveStats)]
        L.dmrgx_engine_host_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        i64p = C.POINTER(C.c_int64)
        L.dmrgx_engine_graph_stats.argtypes = [vp, i64p, i64p, i64p]
        _bound = True
    return L


def default_params(**kw) -> SolveParams:
    p = SolveParams()
    _lib().dmrgx_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@dataclass
class Result:
    tag: int
    status: int
    energy: float
    gs_error: float
    lanczos_steps: int
    ii_steps: int
    cg_steps: int
    spmv_calls: int
    t_submit: float
    t_start: float
    t_done: float
    vec: Optional[np.ndarray]

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class Engine:
    def __init__(self, hbm_budget_bytes: Optional[int] = None, max_active: Optional[int] = None,
                 chunk_ticks: Optional[int] = None, basis_steps_estimate: Optional[int] = None,
                 use_graphs: Optional[bool] = None):
        L = lib()
        o = EngineOpts()
        L.dmrgx_engine_default_opts(C.byref(o))
        if hbm_budget_bytes is not None:
            o.hbm_budget_bytes = hbm_budget_bytes
        if max_active is not None:
            o.max_active = max_active
        if chunk_ticks is not None:
            o.chunk_ticks = chunk_ticks
        if basis_steps_estimate is not None:
            o.basis_steps_estimate = basis_steps_estimate
        if use_graphs is not None:
            o.use_graphs = 1 if use_graphs else 0
        h = C.c_void_p()
        check(L.dmrgx_engine_create(C.byref(o), C.byref(h)))
        self.handle = h

    def submit(self, A: DeviceMatrix, params: Optional[SolveParams] = None, tag: int = 0, own: bool = False) -> int:
        p = params if params is not None else default_params()
        rid = C.c_int64()
        check(lib().dmrgx_engine_submit(self.handle, A.handle, 1 if own else 0, C.byref(p), tag, C.byref(rid)))
        if own:
            A.handle = None   # the engine frees it with the request
        return rid.value

    def run(self, max_chunks: int = 1) -> int:
        check(lib().dmrgx_engine_run(self.handle, max_chunks))
        return int(lib().dmrgx_engine_pending(self.handle))

    @property
    def pending(self) -> int:
        return int(lib().dmrgx_engine_pending(self.handle))

    def poll(self, want_vec: bool = True) -> List[Result]:
        out = []
        L = lib()
        tag, n = C.c_uint64(), C.c_int32()
        while L.dmrgx_engine_peek(self.handle, C.byref(tag), C.byref(n)) == 1:
            r = EngineResult()
            vec = np.empty(n.value) if want_vec else None
            got = L.dmrgx_engine_poll(self.handle, C.byref(r), vec.ctypes.data if want_vec else None)
            if got < 0:
                check(got)
            out.append(Result(r.tag, r.status, r.eig_val, r.gs_error, r.lanczos_steps, r.ii_steps, r.cg_steps,
                              r.spmv_calls, r.t_submit, r.t_start, r.t_done, vec))
        return out

    def run_until_done(self, chunks_per_call: int = 64, want_vec: bool = True) -> List[Result]:
        res = []
        while self.pending > 0:
            self.run(chunks_per_call)
            res += self.poll(want_vec)
        res += self.poll(want_vec)
        return res

    def host_times(self) -> dict:
        """Host seconds by activity (dmrgx_engine_host_times) and the number of chunks run."""
        t = (C.c_double * 6)()
        c = C.c_int64()
        check(lib().dmrgx_engine_host_times(self.handle, t, C.byref(c)))
        keys = ("descriptors", "launch", "wait", "lanczos_checks", "ii_cg_phase_changes", "admission")
        out = {k: round(t[i], 4) for i, k in enumerate(keys)}
        out["chunks"] = c.value
        return out

    def graph_stats(self) -> dict:
        """Chunks launched eagerly, hipGraphs captured, graph replays."""
        e, c, r = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().dmrgx_engine_graph_stats(self.handle, C.byref(e), C.byref(c), C.byref(r)))
        return {"eager": e.value, "captures": c.value, "replays": r.value}

    def free(self):
        if getattr(self, "handle", None):
            lib().dmrgx_engine_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ground_state_batch(mats: Sequence[DeviceMatrix], params: Optional[SolveParams] = None, want_vec: bool = True):
    """dmrgx_ground_state_batch: [(energy, vec, SolveStats)] in input order."""
    L = lib()
    p = params if params is not None else default_params()
    hs = (C.c_void_p * len(mats))(*[m.handle for m in mats])
    vecs = [np.zeros(m.n) for m in mats]
    vp = (C.c_void_p * len(mats))(*[v.ctypes.data for v in vecs])
    stats = (SolveStats * len(mats))()
    check(L.dmrgx_ground_state_batch(len(mats), C.cast(hs, C.c_void_p), C.byref(p),
                                     C.cast(vp, C.c_void_p) if want_vec else None, stats))
    return [(stats[i].eig_val, vecs[i], stats[i]) for i in range(len(mats))]
