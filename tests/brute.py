"""Brute-force superblock Hamiltonian by Kronecker products (test helper).

Independent of the element-by-element assembly: every term of the snapshot's term
list becomes ``coeff * S * (X_LL (x) X_LR (x) X_RR (x) X_RL)`` over the whole product
space (X = the term's operator on its slot, identity elsewhere; S = the diagonal
fermion sign of the row state), and the sum is restricted to the sector basis.
Index convention of the product space: ((LL*d + LR)*dim_RR + RR)*d + RL, the inv_sup
of dmrg/DMRG_MAKE_ELEM_LL_LLLRRRRL.c:26.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from dmrg_c_amd.snapshot import (K_LL, K_LLLR, K_LLRL, K_LLRR, K_LR, K_LRRL, K_LRRR, K_RL,
                                 K_RR, K_RRRL, K_ZERO)

SLOTS1 = {K_LL: 0, K_LR: 1, K_RR: 2, K_RL: 3, K_LLLR: 0, K_LLRL: 0, K_LLRR: 0, K_LRRL: 1,
          K_LRRR: 1, K_RRRL: 2}
SLOTS2 = {K_LLLR: 1, K_LLRL: 3, K_LLRR: 2, K_LRRL: 3, K_LRRR: 2, K_RRRL: 3}


def _csr(c):
    return sp.csr_matrix((c.val, c.col, c.row), shape=(c.row_dim, c.col_dim))


def superblock_dense(snap) -> np.ndarray:
    d, dl, dr = snap.dim_onsite, snap.dim_LL, snap.dim_RR
    dims = [dl, d, dr, d]
    idx = ((snap.LL.astype(np.int64) * d + snap.LR) * dr + snap.RR) * d + snap.RL
    qs = [snap.LL.astype(np.int64), snap.LR.astype(np.int64), snap.RR.astype(np.int64),
          snap.RL.astype(np.int64)]
    H = np.zeros((snap.n, snap.n))
    for t in snap.terms:
        k = int(t["kind"])
        c = float(t["coeff"])
        if k == K_ZERO or abs(c) < 1e-15:
            continue
        mats = [sp.identity(x, format="csr") for x in dims]
        mats[SLOTS1[k]] = _csr(snap.ops[int(t["op1"])])
        if k in SLOTS2:
            mats[SLOTS2[k]] = _csr(snap.ops[int(t["op2"])])
        full = mats[0]
        for m in mats[1:]:
            full = sp.kron(full, m, format="csr")
        block = full[idx][:, idx].toarray()
        sign = np.ones(snap.n)
        mode = int(t["sign_mode"])
        if mode:
            e = [snap.eles[int(t[f"ele{i}"])] if int(t[f"ele{i}"]) >= 0 else None for i in range(3)]
            if k == K_LLLR:
                p = e[0][qs[0]]
            elif k == K_LLRL:
                p = e[0][qs[0]] + e[1][qs[2]] + e[2][qs[1]]
            elif k == K_LLRR:
                p = e[0][qs[0]] + e[1][qs[1]]
            elif k == K_LRRL:
                p = e[0][qs[2]] + e[1][qs[1]]
            elif k == K_LRRR:
                p = e[0][qs[1]]
            else:  # K_RRRL
                p = e[0][qs[2]]
            even = (p % 2) == 0
            sign = np.where(even, -1.0, 1.0) if mode == 1 else np.where(even, 1.0, -1.0)
        H += c * sign[:, None] * block
    return H
