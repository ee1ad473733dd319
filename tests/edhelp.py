"""Exact-diagonalization fixtures (test helper).

``load_ed_bin`` reads what oracle/harness/ed_ref.c writes (the reference's element makers,
EXACT_DIAGONALIZE_HAMILTONIAN, EXACT_EXPECTATION_* and EXACT_V_M_Q1/Q2/Q3 on one model);
``save_ed_npz`` / ``load_ed`` keep it as a compact fixture under tests/golden/.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from dmrg_c_amd.snapshot import Crs


@dataclass
class EdObs:
    kind: int          # 0 onsite over all sites, 1 intersite Q0 from `origin`
    op1: int
    op2: int
    origin: int
    out: np.ndarray


@dataclass
class EdVm:
    q: int             # EXACT_V_M_Q<q>
    op: int
    site: int
    sign_flag: int
    basis_out: np.ndarray
    out: np.ndarray


@dataclass
class EdCase:
    model: int
    L: int
    d: int
    n: int
    ops: List[Crs]
    n_ele: np.ndarray
    sz2: np.ndarray
    basis: np.ndarray
    ham: Crs
    energy: float
    error: float
    vec: np.ndarray
    obs: List[EdObs] = field(default_factory=list)
    vm: List[EdVm] = field(default_factory=list)
    args: Dict[str, str] = field(default_factory=dict)


def _rd(f, dt, n):
    b = f.read(np.dtype(dt).itemsize * n)
    return np.frombuffer(b, dtype=dt, count=n).copy()


def _crs(f) -> Crs:
    rd, cd = _rd(f, "<i4", 2)
    nnz = int(_rd(f, "<i8", 1)[0])
    row = _rd(f, "<i8", rd + 1)
    col = _rd(f, "<i4", nnz)
    val = _rd(f, "<f8", nnz)
    return Crs(int(rd), int(cd), row, col, val)


def load_ed_bin(path: str, args: Dict[str, str]) -> EdCase:
    with open(path, "rb") as f:
        assert f.read(8) == b"DMRGED01"
        model, L, d, n, n_ops = (int(x) for x in _rd(f, "<i4", 5))
        ops = [_crs(f) for _ in range(n_ops)]
        n_ele = _rd(f, "<i4", d)
        sz2 = _rd(f, "<i4", d)
        basis = _rd(f, "<i8", n)
        ham = _crs(f)
        e, err = _rd(f, "<f8", 2)
        vec = _rd(f, "<f8", n)
        obs = []
        for _ in range(int(_rd(f, "<i4", 1)[0])):
            kind, op1, op2, _site, origin, ln = (int(x) for x in _rd(f, "<i4", 6))
            obs.append(EdObs(kind, op1, op2, origin, _rd(f, "<f8", ln)))
        vm = []
        for _ in range(int(_rd(f, "<i4", 1)[0])):
            q, op, site, flag, nout = (int(x) for x in _rd(f, "<i4", 5))
            vm.append(EdVm(q, op, site, flag, _rd(f, "<i8", nout), _rd(f, "<f8", nout)))
    return EdCase(model, L, d, n, ops, n_ele, sz2, basis, ham, float(e), float(err), vec, obs, vm, args)


def save_ed_npz(c: EdCase, path: str) -> None:
    z = dict(meta=np.array([c.model, c.L, c.d, c.n, len(c.ops), len(c.obs), len(c.vm)], np.int64),
             n_ele=c.n_ele, sz2=c.sz2, basis=c.basis, ham_row=c.ham.row, ham_col=c.ham.col, ham_val=c.ham.val,
             energy=np.array([c.energy, c.error]), vec=c.vec,
             args=np.frombuffer(json.dumps(c.args).encode(), np.uint8))
    for i, o in enumerate(c.ops):
        z[f"op{i}_row"], z[f"op{i}_col"], z[f"op{i}_val"] = o.row, o.col, o.val
    for i, o in enumerate(c.obs):
        z[f"obs{i}_hdr"] = np.array([o.kind, o.op1, o.op2, o.origin], np.int64)
        z[f"obs{i}_out"] = o.out
    for i, v in enumerate(c.vm):
        z[f"vm{i}_hdr"] = np.array([v.q, v.op, v.site, v.sign_flag], np.int64)
        z[f"vm{i}_basis"] = v.basis_out
        z[f"vm{i}_out"] = v.out
    np.savez_compressed(path, **z)


def load_ed(path: str) -> EdCase:
    z = np.load(path)
    model, L, d, n, n_ops, n_obs, n_vm = (int(x) for x in z["meta"])
    ops = [Crs(d, d, z[f"op{i}_row"], z[f"op{i}_col"], z[f"op{i}_val"]) for i in range(n_ops)]
    ham = Crs(n, n, z["ham_row"], z["ham_col"], z["ham_val"])
    obs = []
    for i in range(n_obs):
        k, o1, o2, org = (int(x) for x in z[f"obs{i}_hdr"])
        obs.append(EdObs(k, o1, o2, org, z[f"obs{i}_out"]))
    vm = []
    for i in range(n_vm):
        q, op, site, flag = (int(x) for x in z[f"vm{i}_hdr"])
        vm.append(EdVm(q, op, site, flag, z[f"vm{i}_basis"], z[f"vm{i}_out"]))
    e, err = z["energy"]
    args = json.loads(bytes(z["args"]).decode())
    return EdCase(model, L, d, n, ops, z["n_ele"], z["sz2"], z["basis"], ham, float(e), float(err), z["vec"], obs,
                  vm, args)


def model_of(c: EdCase):
    """The dmrg_c_amd.ed.Model the harness arguments describe (same defaults as ed_ref.c)."""
    from dmrg_c_amd.ed import Model
    a = c.args
    m = Model(model=a.get("model", "xxz"), L=int(a.get("L", 8)), spin=int(a.get("spin", 1)),
              N=int(a.get("N", -1)), sz=int(a.get("sz", 0)), bc=a.get("BC", "OBC"),
              J_xy=float(a.get("J_xy", 1.0)), J_z=float(a.get("J_z", 1.0)), h_z=float(a.get("h_z", 0.0)),
              D_z=float(a.get("D_z", 0.0)), t=float(a.get("t", -1.0)), U=float(a.get("U", 4.0)),
              V=float(a.get("V", 0.0)), mu=float(a.get("mu", 0.0)))
    return m
