"""The reference's own XXZ DMRG program relinked against libdmrg_c_amd.so.

oracle/_ref/relink_xxz_{eig,all} are the program's sources (main/dmrg/xxz_vf/src minus main.c,
oracle/harness/snap.c as main) compiled from the reference tree and linked with the library
ahead of the reference archive, exactly as a maintainer relinks a dmrg_C program: every
reference symbol the library exports (superblock eigensolve, element makers, expectation
values) resolves to the library.  ``eig`` keeps the reference's own truncation objects
(DMRG_GET_DENSITY_MATRIX, DMRG_GET_SYSTEM_INFO_Q1/Q2, CRS_CRS_CCS_PRODUCT) in the program;
``all`` takes those from the library too.  ``obc`` is ``all`` with limit=0: every step takes the
program's matrix-free path (main/dmrg/xxz_vf/src/allocations/GET_GROUND_STATE.c:20-21 ->
DIAGONALIZE_SUPERBLOCK.c:11-79): the sub-Hamiltonians GET_HAM_LLLR/LRRL/RRRL built with the
library's DMRG_MAKE_ELEM_*_{LLLR,LRRL,RRRL}, then the library's DMRG_LANCZOS_SLOW_OBC,
DMRG_INVERSE_ITERATION_OBC, DMRG_CONJUGATE_GRADIENT_OBC, DMRG_MATRIX_VECTOR_PRODUCT_OBC and
DMRG_CONVERGE_CHECK_OBC.  The unmodified reference run through its own matrix-free path
(limit=0) gives the same 127 dimensions and energies within 6e-14 of the CRS path's, so all
variants are compared with the same reference output.

Workload: XXZ chain, L=32, J_xy=J_z=1, h_z=0, S^z=0, m=64, 4 sweeps, Lanczos_Slow with
diag_acc=1e-12, inverse iteration CG (main/dmrg/xxz_vf/src/main/DMRG.c:28-53, one superblock
solve per RENORMALIZE.c:20-38 step, 127 steps).  Compared against the unmodified reference
program's own output for the same arguments (tests/golden/xxz_L32_m64_sw4_energies.txt and
tests/golden/xxz_L32_m64_sw4_result/, written by tests/golden/make_golden.py):

* every step's (step, sweep, LL_site, RR_site, superblock dimension) identical -- the
  dimension is set by the previous step's truncation, so this pins the whole trajectory;
* every step's ground-state energy within 1e-10 relative, gs_error below ii_acc (1e-9);
* the program's result files: on-site values (Sz, SzSz, SxSx) and correlation functions
  (Sz_CF, Sx_CF) at the last two sweeps' centre within 1e-8, entanglement entropies within
  1e-8, truncation-error lines with identical site/step columns and values within 1e-10.

The reference is not bit-deterministic itself: run with 2 and 8 OpenMP threads its energies
differ by up to ~5e-13 (OpenMP reductions), while every dimension is the same.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO

pytestmark = pytest.mark.gpu

ARGS = "L=32 m=64 sweep=4 J_xy=1.0 J_z=1.0 h_z=0.0 diag_acc=1e-12 dump= dump_snap=0 dump_vec=0 threads=8"
GOLD_E = os.path.join(GOLDEN, "xxz_L32_m64_sw4_energies.txt")
GOLD_R = os.path.join(GOLDEN, "xxz_L32_m64_sw4_result")


def _energies(path):
    rows = [ln.split() for ln in open(path) if ln.strip()]
    key = np.array([[int(x) for x in r[:5]] for r in rows])
    e = np.array([float(r[5]) for r in rows])
    err = np.array([float(r[6]) for r in rows])
    return key, e, err


def _table(path):
    """Numeric rows of a result file (the reference's '###' header lines skipped)."""
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    return np.array([[float(x) for x in r] for r in rows])


# variant -> (program, extra arguments); "obc" takes every step through the matrix-free path
VARIANTS = {"eig": ("relink_xxz_eig", ""), "all": ("relink_xxz_all", ""), "obc": ("relink_xxz_all", "limit=0")}


@pytest.fixture(scope="module", params=sorted(VARIANTS))
def run(request, tmp_path_factory):
    prog, extra = VARIANTS[request.param]
    exe = os.path.join(REPO, "oracle", "_ref", prog)
    if not os.path.exists(exe):
        pytest.skip("relinked reference program not built (make -C oracle relink, needs the reference tree)")
    wd = tmp_path_factory.mktemp(f"relink_{request.param}")
    env = dict(os.environ, MKL_THREADING_LAYER="GNU")
    print(f"{prog}: {ARGS} {extra}", flush=True)
    with open(wd / "log.txt", "w") as log:
        r = subprocess.run([exe] + ARGS.split() + extra.split() + [f"out={wd}"], cwd=wd, env=env, stdout=log,
                           stderr=subprocess.STDOUT, timeout=240)
    assert r.returncode == 0, open(wd / "log.txt", errors="replace").read()[-2000:]
    return request.param, wd


def test_trajectory_and_energies(run):
    variant, wd = run
    key, e, err = _energies(wd / "energies.txt")
    gkey, ge, _ = _energies(GOLD_E)
    assert key.shape == gkey.shape == (127, 5)
    bad = np.nonzero((key != gkey).any(axis=1))[0]
    assert bad.size == 0, f"{variant}: steps whose (sweep, LL, RR, n) differ: {key[bad].tolist()} vs {gkey[bad].tolist()}"
    rel = np.abs(e - ge) / np.abs(ge)
    assert rel.max() <= 1e-10, (variant, int(np.argmax(rel)) + 1, rel.max())
    assert err.max() < 1e-9, (variant, err.max())


def _result_files():
    return sorted(os.path.relpath(p, GOLD_R) for p in glob.glob(os.path.join(GOLD_R, "**", "*.txt"), recursive=True))


def test_result_files(run):
    variant, wd = run
    files = _result_files()
    assert len(files) >= 20
    for rel in files:
        mine = os.path.join(wd, "result", rel)
        assert os.path.exists(mine), rel
        a, b = _table(mine), _table(os.path.join(GOLD_R, rel))
        assert a.shape == b.shape, (rel, a.shape, b.shape)
        name = os.path.basename(rel)
        if name.startswith("T_Error"):
            # size, site, step columns; total weight and truncation error
            assert np.array_equal(a[:, :3], b[:, :3]), rel
            assert np.abs(a[:, 3:] - b[:, 3:]).max() <= 1e-10, (rel, np.abs(a[:, 3:] - b[:, 3:]).max())
        elif name.startswith("Energy"):
            assert np.array_equal(a[:, :3], b[:, :3]), rel
            assert (np.abs(a[:, 3:5] - b[:, 3:5]) <= 1e-10 * np.abs(b[:, 3:5])).all(), rel
        elif name == "energy.txt":
            assert np.array_equal(a[:, [0, 1, 2, 3, 4, 5, 8, 9]], b[:, [0, 1, 2, 3, 4, 5, 8, 9]]), rel
            assert (np.abs(a[:, 6:8] - b[:, 6:8]) <= 1e-10 * np.abs(b[:, 6:8])).all(), rel
        else:
            # EE_*: size, site, step, entropies; ExpectationValues/*, CorrelationFunctions/*:
            # site (and partner) indices, values
            ncol = 3 if name.startswith("EE") else (3 if "CorrelationFunctions" in rel else 1)
            assert np.array_equal(a[:, :ncol], b[:, :ncol]), rel
            assert np.abs(a[:, ncol:] - b[:, ncol:]).max() <= 1e-8, (rel, np.abs(a[:, ncol:] - b[:, ncol:]).max())
