"""Superblock snapshot: the inputs of one superblock eigensolve, as plain arrays.

A snapshot holds what ``GET_HAM_LLLRRRRL`` consumes for one DMRG step
(``main/dmrg/<model>_vf/src/allocations/GET_HAM_LLLRRRRL.c:11-123``):

* the sorted superblock basis ``LL/LR/RR/RL`` (``DMRG_BASIS``, ``include/dmrg.h:54-99``),
  ordered as ``GET_BASIS_SUPERBLOCK`` leaves it;
* the block / on-site operator CRS matrices that the model's
  ``MAKE_ELEMENT_HAM_LLLRRRRL`` passes to ``DMRG_MAKE_ELEM_*_LLLRRRRL``;
* the electron-number parity arrays (``Ele_LL``, ``Ele_RR``, ``Ele_On``) used by the
  fermion sign rule of the two-operator terms;
* the ordered term list ``(kind, op1, op2, ele..., sign_mode, coeff)``, in call
  order, which fixes the floating-point accumulation order of every matrix element.

Optionally it also holds the lower-triangle CRS the reference assembled from those
inputs, and (XXZ only) the three sub-Hamiltonians of the matrix-free path
(``main/dmrg/xxz_vf/src/allocations/DIAGONALIZE_SUPERBLOCK.c:15-17``).

Two encodings exist: the raw little-endian ``.bin`` stream (magic ``DMRGSNP1``) and
an ``.npz`` archive with the same fields under fixed names (``save_npz``/``load``).
"""
from __future__ import annotations

import io
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

# term kinds, in the order of include/dmrg.h:330-340
KIND_NAMES = ["LL", "LR", "RR", "RL", "LLLR", "LLRL", "LLRR", "LRRL", "LRRR", "RRRL", "ZERO"]
K_LL, K_LR, K_RR, K_RL, K_LLLR, K_LLRL, K_LLRR, K_LRRL, K_LRRR, K_RRRL, K_ZERO = range(11)

TERM_DTYPE = np.dtype([
    ("kind", "<i4"), ("op1", "<i4"), ("op2", "<i4"),
    ("ele0", "<i4"), ("ele1", "<i4"), ("ele2", "<i4"),
    ("sign_mode", "<i4"), ("coeff", "<f8"),
])


@dataclass
class Crs:
    """CRS1 content (include/SML.h:17-28): int64 Row[row_dim+1], int32 Col, f64 Val."""
    row_dim: int
    col_dim: int
    row: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row[self.row_dim])

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.row_dim, self.col_dim))
        for i in range(self.row_dim):
            s, e = int(self.row[i]), int(self.row[i + 1])
            d[i, self.col[s:e]] += self.val[s:e]
        return d


@dataclass
class Snapshot:
    n: int
    dim_LL: int
    dim_RR: int
    dim_onsite: int
    LL_site: int
    RR_site: int
    step: int
    LL: np.ndarray
    LR: np.ndarray
    RR: np.ndarray
    RL: np.ndarray
    ops: List[Crs]
    eles: List[np.ndarray]
    terms: np.ndarray
    ham: Optional[Crs] = None
    obc: Optional[Dict[str, object]] = None
    meta: Dict[str, object] = field(default_factory=dict)

    @property
    def dim_whole(self) -> int:
        return self.dim_LL * self.dim_onsite * self.dim_RR * self.dim_onsite

    def inv_table(self) -> np.ndarray:
        """Dense superblock inverse, -1 where absent
        (GET_BASIS_SUPERBLOCK.c: Inv fill loop, xxz_vf :79-86)."""
        d, dr = self.dim_onsite, self.dim_RR
        idx = ((self.LL.astype(np.int64) * d + self.LR) * dr + self.RR) * d + self.RL
        inv = np.full(self.dim_whole, -1, dtype=np.int32)
        inv[idx] = np.arange(self.n, dtype=np.int32)
        return inv


def _read(buf: io.BufferedReader, dtype, count: int) -> np.ndarray:
    dt = np.dtype(dtype)
    raw = buf.read(dt.itemsize * count)
    if len(raw) != dt.itemsize * count:
        raise ValueError("truncated snapshot")
    return np.frombuffer(raw, dtype=dt, count=count).copy()


def _read_crs(buf) -> Crs:
    row_dim, col_dim = (int(v) for v in _read(buf, "<i4", 2))
    nnz = int(_read(buf, "<i8", 1)[0])
    row = _read(buf, "<i8", row_dim + 1)
    col = _read(buf, "<i4", nnz)
    val = _read(buf, "<f8", nnz)
    return Crs(row_dim, col_dim, row, col, val)


def load_bin(path: str) -> Snapshot:
    with open(path, "rb") as f:
        if f.read(8) != b"DMRGSNP1":
            raise ValueError(f"{path}: not a DMRGSNP1 snapshot")
        n, dim_LL, dim_RR, d, LL_site, RR_site, step = (int(v) for v in _read(f, "<i4", 7))
        LL = _read(f, "<i2", n)
        LR = _read(f, "<i2", n)
        RR = _read(f, "<i2", n)
        RL = _read(f, "<i2", n)
        ops = [_read_crs(f) for _ in range(int(_read(f, "<i4", 1)[0]))]
        eles = []
        for _ in range(int(_read(f, "<i4", 1)[0])):
            ln = int(_read(f, "<i4", 1)[0])
            eles.append(_read(f, "<i4", ln))
        nt = int(_read(f, "<i4", 1)[0])
        terms = np.zeros(nt, dtype=TERM_DTYPE)
        for t in range(nt):
            ints = _read(f, "<i4", 7)
            c = _read(f, "<f8", 1)[0]
            terms[t] = (*[int(x) for x in ints], c)
        ham = _read_crs(f) if int(_read(f, "<i4", 1)[0]) else None
        obc = None
        tail = f.read(4)
        if len(tail) == 4 and int(np.frombuffer(tail, "<i4")[0]) == 1:
            obc = {}
            m = int(_read(f, "<i4", 1)[0])
            obc["LL_LLLR"], obc["LR_LLLR"] = _read(f, "<i2", m), _read(f, "<i2", m)
            m = int(_read(f, "<i4", 1)[0])
            obc["LR_LRRL"], obc["RL_LRRL"] = _read(f, "<i2", m), _read(f, "<i2", m)
            m = int(_read(f, "<i4", 1)[0])
            obc["RR_RRRL"], obc["RL_RRRL"] = _read(f, "<i2", m), _read(f, "<i2", m)
            obc["M_LLLR"], obc["M_LRRL"], obc["M_RRRL"] = _read_crs(f), _read_crs(f), _read_crs(f)
    return Snapshot(n, dim_LL, dim_RR, d, LL_site, RR_site, step, LL, LR, RR, RL,
                    ops, eles, terms, ham, obc)


def _crs_to_dict(prefix: str, c: Crs, out: dict) -> None:
    out[prefix + "_dims"] = np.array([c.row_dim, c.col_dim], dtype=np.int64)
    out[prefix + "_row"] = c.row
    out[prefix + "_col"] = c.col
    out[prefix + "_val"] = c.val


def _crs_from_dict(prefix: str, z) -> Crs:
    rd, cd = (int(v) for v in z[prefix + "_dims"])
    return Crs(rd, cd, z[prefix + "_row"], z[prefix + "_col"], z[prefix + "_val"])


def save_npz(snap: Snapshot, path: str, **extra: np.ndarray) -> None:
    out: Dict[str, np.ndarray] = {
        "header": np.array([snap.n, snap.dim_LL, snap.dim_RR, snap.dim_onsite,
                            snap.LL_site, snap.RR_site, snap.step], dtype=np.int64),
        "LL": snap.LL, "LR": snap.LR, "RR": snap.RR, "RL": snap.RL,
        "terms": snap.terms,
        "n_ops": np.array([len(snap.ops)]), "n_eles": np.array([len(snap.eles)]),
    }
    for i, c in enumerate(snap.ops):
        _crs_to_dict(f"op{i}", c, out)
    for i, e in enumerate(snap.eles):
        out[f"ele{i}"] = e
    if snap.ham is not None:
        _crs_to_dict("ham", snap.ham, out)
    if snap.obc is not None:
        for k, v in snap.obc.items():
            if isinstance(v, Crs):
                _crs_to_dict("obc_" + k, v, out)
            else:
                out["obc_" + k] = v
    out.update(extra)
    np.savez_compressed(path, **out)


def load(path: str) -> Snapshot:
    if path.endswith(".bin"):
        return load_bin(path)
    z = np.load(path, allow_pickle=False)
    n, dim_LL, dim_RR, d, LL_site, RR_site, step = (int(v) for v in z["header"])
    ops = [_crs_from_dict(f"op{i}", z) for i in range(int(z["n_ops"][0]))]
    eles = [z[f"ele{i}"] for i in range(int(z["n_eles"][0]))]
    ham = _crs_from_dict("ham", z) if "ham_dims" in z else None
    obc = None
    if "obc_M_LLLR_dims" in z:
        obc = {k: z["obc_" + k] for k in ("LL_LLLR", "LR_LLLR", "LR_LRRL", "RL_LRRL", "RR_RRRL", "RL_RRRL")}
        for k in ("M_LLLR", "M_LRRL", "M_RRRL"):
            obc[k] = _crs_from_dict("obc_" + k, z)
    meta = {k: z[k] for k in z.files if k.startswith("meta_") or k.startswith("gs_")}
    return Snapshot(n, dim_LL, dim_RR, d, LL_site, RR_site, step,
                    z["LL"], z["LR"], z["RR"], z["RL"], ops, eles, z["terms"].astype(TERM_DTYPE),
                    ham, obc, meta)


def load_gs_bin(path: str):
    """Reference ground state written next to a snapshot: (E, gs_error, vector)."""
    with open(path, "rb") as f:
        if f.read(8) != b"DMRGGSV1":
            raise ValueError(f"{path}: not a DMRGGSV1 file")
        n = int(_read(f, "<i4", 1)[0])
        e, err = (float(v) for v in _read(f, "<f8", 2))
        vec = _read(f, "<f8", n)
    return e, err, vec


@dataclass
class ExpectationCall:
    """Inputs and outputs of one DMRG_EXPECTATION_ONSITE (kind 0) or
    DMRG_EXPECTATION_INTERSITE_Q0 (kind 1) call (dmrg/DMRG_EXPECTATION_ONSITE.c:3-28,
    dmrg/DMRG_EXPECTATION_INTERSITE_Q0.c:3-43)."""
    kind: int
    n: int
    dim_LL: int
    dim_RR: int
    dim_onsite: int
    LL_site: int
    RR_site: int
    origin: int
    LL: np.ndarray
    LR: np.ndarray
    RR: np.ndarray
    RL: np.ndarray
    M_LL: List[Crs]
    M_On: Crs
    M_RR: List[Crs]
    M_CF: List[Crs]
    vec: np.ndarray
    out: np.ndarray


def load_expect_bin(path: str) -> ExpectationCall:
    with open(path, "rb") as f:
        if f.read(8) != b"DMRGEXP1":
            raise ValueError(f"{path}: not a DMRGEXP1 file")
        kind, n, dl, dr, d, lls, rrs, origin = (int(v) for v in _read(f, "<i4", 8))
        LL, LR, RR, RL = (_read(f, "<i2", n) for _ in range(4))
        mll = [_read_crs(f) for _ in range(int(_read(f, "<i4", 1)[0]))]
        mon = _read_crs(f)
        mrr = [_read_crs(f) for _ in range(int(_read(f, "<i4", 1)[0]))]
        mcf = [_read_crs(f) for _ in range(int(_read(f, "<i4", 1)[0]))]
        vec = _read(f, "<f8", n)
        out = _read(f, "<f8", int(_read(f, "<i4", 1)[0]))
    return ExpectationCall(kind, n, dl, dr, d, lls, rrs, origin, LL, LR, RR, RL, mll, mon, mrr, mcf, vec, out)


def save_expect_npz(e: ExpectationCall, path: str) -> None:
    out: Dict[str, np.ndarray] = {
        "header": np.array([e.kind, e.n, e.dim_LL, e.dim_RR, e.dim_onsite, e.LL_site, e.RR_site, e.origin]),
        "LL": e.LL, "LR": e.LR, "RR": e.RR, "RL": e.RL, "vec": e.vec, "out": e.out,
        "counts": np.array([len(e.M_LL), len(e.M_RR), len(e.M_CF)]),
    }
    for i, c in enumerate(e.M_LL):
        _crs_to_dict(f"ll{i}", c, out)
    for i, c in enumerate(e.M_RR):
        _crs_to_dict(f"rr{i}", c, out)
    for i, c in enumerate(e.M_CF):
        _crs_to_dict(f"cf{i}", c, out)
    _crs_to_dict("on", e.M_On, out)
    np.savez_compressed(path, **out)


def load_expect(path: str) -> ExpectationCall:
    if path.endswith(".bin"):
        return load_expect_bin(path)
    z = np.load(path, allow_pickle=False)
    kind, n, dl, dr, d, lls, rrs, origin = (int(v) for v in z["header"])
    nll, nrr, ncf = (int(v) for v in z["counts"])
    return ExpectationCall(kind, n, dl, dr, d, lls, rrs, origin, z["LL"], z["LR"], z["RR"], z["RL"],
                           [_crs_from_dict(f"ll{i}", z) for i in range(nll)], _crs_from_dict("on", z),
                           [_crs_from_dict(f"rr{i}", z) for i in range(nrr)],
                           [_crs_from_dict(f"cf{i}", z) for i in range(ncf)], z["vec"], z["out"])
