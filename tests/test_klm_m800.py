"""The KLM L=48, N=36, S^z=0, t=1, J_K=2, h_z=0.19109 (seed 0), m=800 superblock of the
reference's own run (main/dmrg/klm_vf, model terms main/dmrg/klm_vf/src/main/
MAKE_ELEMENT_HAM_LLLRRRRL.c:11-99, parameters as main.c:14-50 sets them), last infinite-system
step (48 sites), written by tests/golden/make_golden.py --big (klm_L48_m800: block operators,
term list, basis, the reference's energy; the blocks' quantum numbers in qn_klm_L48_m800_s23.npz).

* the superblock basis built on the device from the blocks' quantum numbers equals the
  basis the reference's GET_BASIS_SUPERBLOCK built (bit for bit);
* device assembly equals the oracle's (pinned bit for bit against the reference's
  GET_HAM_LLLRRRRL on the golden set) in Row, Col and Val, with int64 row pointers;
* FP64 ground state within 1e-10 of the reference's energy, gs_error below ii_acc 1e-9;
* FP32-refined ground state within 1e-9, gs_error below 1e-9.
The snapshot holds the model's block operators and term list (4.3 MB); the 683.6 M-entry
matrix is rebuilt from them on the device and by the oracle."""
import os

import numpy as np
import pytest

import dmrg_c_amd as D
import oracle_py as O
from conftest import GOLDEN
from dmrg_c_amd import snapshot

pytestmark = pytest.mark.gpu

SNAP = os.path.join(GOLDEN, "klm_L48_m800_s23.npz")
QN = os.path.join(GOLDEN, "qn_klm_L48_m800_s23.npz")


@pytest.fixture(scope="module")
def snap():
    if not os.path.exists(SNAP):
        pytest.skip("KLM L=48 m=800 snapshot not generated (python tests/golden/make_golden.py --big)")
    return snapshot.load(SNAP)


@pytest.fixture(scope="module")
def sb(snap):
    s = D.Superblock(snap)
    yield s
    s.free()


def test_basis_from_quantum_numbers(snap):
    if not os.path.exists(QN):
        pytest.skip("quantum numbers not generated")
    z = np.load(QN)
    n_qn, dim_LL, dim_RR, d, t_sz, t_ele = (int(v) for v in z["header"])
    q = dict(n_qn=n_qn, target_sz=t_sz, target_ele=t_ele,
             **{k: z[k] for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR", "sz_site", "ele_site")})
    print(f"klm m=800 basis n={snap.n}", flush=True)
    got = D.superblock_basis(q)
    for a, b in zip(got, (snap.LL, snap.LR, snap.RR, snap.RL)):
        assert np.array_equal(a, b)


def test_device_assembly_bit_exact(snap, sb):
    print(f"klm m=800 assembly n={snap.n}", flush=True)
    A = sb.assemble(full=True)
    row, col, val = sb.download_lower()
    A.free()
    orow, ocol, oval = O.assemble(snap)
    assert np.array_equal(row, orow)
    assert np.array_equal(col, ocol)
    assert np.array_equal(val.view(np.int64), oval.view(np.int64))


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-9), (2, 1e-9)],
                         ids=["fp64", "fp32_refined", "fp32_two_phase"])
def test_ground_state(snap, sb, precision, tol):
    e_ref = float(snap.meta["gs_energy"][0])
    A = sb.assemble(full=True)
    print(f"klm m=800 solve precision={precision} n={snap.n} nnz={A.nnz}", flush=True)
    E, vec, st = A.ground_state(precision=precision)
    A.free()
    assert abs(E - e_ref) <= tol * abs(e_ref), (E, e_ref)
    assert st.gs_error < 1e-9, st.gs_error
    assert abs(np.linalg.norm(vec) - 1.0) < 1e-12
