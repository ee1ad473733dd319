"""The library's host tridiagonal routine (dmrgx_tridiag_lowest: the Lanczos convergence
checks' replacement for dstev_) against scipy.linalg.eigh_tridiagonal, on random,
clustered, decoupled and real Lanczos tridiagonals (the last with the ghost copies of the
lowest eigenvalue that a recurrence without reorthogonalisation produces)."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla


def _lib():
    import dmrg_c_amd as D
    L = D.lib()
    L.dmrgx_tridiag_lowest.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def _run(a, b, want_vec=False):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b if len(b) else np.zeros(1), np.float64)
    lam = np.empty(len(a))
    vec = np.empty(len(a)) if want_vec else None
    rc = _lib().dmrgx_tridiag_lowest(len(a), a.ctypes.data, b.ctypes.data, lam.ctypes.data,
                                     vec.ctypes.data if want_vec else None)
    assert rc == 0
    return lam, vec


def _check_sequence(a, b, rel=2e-15):
    lam, vec = _run(a, b, want_vec=True)
    scale = max(np.abs(a).max(), np.abs(b).max() if len(b) else 0.0, 1.0)
    for j in range(len(a)):
        ref = sla.eigh_tridiagonal(a[:j + 1], b[:j], eigvals_only=True, select="i", select_range=(0, 0))[0] \
            if j > 0 else a[0]
        assert abs(lam[j] - ref) <= 8 * rel * scale, (j, lam[j], ref)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    assert abs(np.linalg.norm(vec) - 1.0) < 1e-12
    assert np.linalg.norm(T @ vec - lam[-1] * vec) <= 1e-10 * scale
    return lam


@pytest.mark.parametrize("k", [1, 2, 3, 10, 57, 200])
def test_random_tridiagonals(k):
    rng = np.random.default_rng(k)
    for _ in range(3):
        _check_sequence(rng.standard_normal(k) * 3, rng.standard_normal(k - 1))


def test_clustered_and_decoupled():
    k = 60
    _check_sequence(np.ones(k), np.full(k - 1, 1e-9))                       # tight cluster
    b = np.random.default_rng(1).standard_normal(k - 1)
    b[::7] = 0.0                                                           # decoupled blocks
    _check_sequence(np.random.default_rng(2).standard_normal(k), b)
    _check_sequence(np.zeros(k), np.ones(k - 1))                           # free chain
    a = np.full(k, -2.0)
    a[0] = -7.0                                                            # isolated lowest
    _check_sequence(a, np.full(k - 1, 0.5))


def _lanczos_coeffs(H, k, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(H.shape[0])
    v /= np.linalg.norm(v)
    vp = np.zeros_like(v)
    a, b = [], []
    beta = 0.0
    for _ in range(k):
        w = H @ v
        al = w @ v
        w = w - al * v - beta * vp
        beta = np.linalg.norm(w)
        a.append(al)
        b.append(beta)
        vp, v = v, w / beta
    return np.array(a), np.array(b[:-1])


def test_lanczos_sequences_with_ghosts():
    """180 steps without reorthogonalisation on a 400-state matrix: the lowest Ritz value
    converges, then ghost copies of it appear (near-degenerate lowest eigenvalue of T)."""
    rng = np.random.default_rng(7)
    n = 400
    M = rng.standard_normal((n, n))
    H = (M + M.T) / 2
    H[0, 0] -= 40.0   # a separated ground state, converged early
    a, b = _lanczos_coeffs(H, 180)
    lam = _check_sequence(a, b, rel=4e-15)
    assert np.all(np.diff(lam) <= 1e-12 * np.abs(lam).max())   # interlacing: non-increasing
    w = np.linalg.eigvalsh(H)[0]
    assert abs(lam[-1] - w) < 1e-9 * abs(w)
