"""Regenerate the golden fixtures in tests/golden/ from the compiled reference.

Runs the reference's own DMRG programs (built by ``make -C oracle ref`` from the
sources under /root/reference into oracle/_ref/, with oracle/harness/snap.c as main)
and converts the superblock snapshots they dump into compact ``.npz`` files:

  <name>_s<step>.npz   inputs of GET_HAM_LLLRRRRL (basis, operators, term list),
                       the reference's assembled lower-triangle CRS (``ham_*``),
                       and the reference's ground state (``gs_energy``, ``gs_error``,
                       ``gs_vec``) from DMRG_DIAGONALIZE_SUPERBLOCK_SYM
  <name>_expect_s<step>_<k>.npz  one DMRG_EXPECTATION_ONSITE / _INTERSITE_Q0 call of the
                       reference: block/on-site operators, the GS vector, the reference's output
  <name>_energies.txt  per-step: step sweep LL_site RR_site n E gs_error seconds
  ed_<name>.npz        one exact diagonalization (oracle/harness/ed_ref.c): on-site operators,
                       sector basis, the full Hamiltonian CRS built by the reference's
                       EXACT_MAKE_ELEM_*, EXACT_DIAGONALIZE_HAMILTONIAN's ground state, and the
                       outputs of EXACT_EXPECTATION_* and EXACT_V_M_Q1/Q2/Q3 on it

This script is only run in the build container (the reference is not on the GPU box).
Usage:  python tests/golden/make_golden.py [--only NAME] [--big]
"""
from __future__ import annotations

import argparse
import glob
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(HERE))
from dmrg_c_amd import snapshot  # noqa: E402
import edhelp  # noqa: E402

REF_BIN = os.path.join(REPO, "oracle", "_ref")

# name -> (program, args, steps to keep as fixtures ("all" or list), keep vectors)
CASES = {
    "xxz_L8_m16": ("snap_xxz", "L=8 m=16 sweep=1 J_xy=1.0 J_z=1.0 h_z=0.0", "all", True),
    "xxz_L12_m24_jz05": ("snap_xxz", "L=12 m=24 sweep=1 J_xy=1.0 J_z=0.5 h_z=0.1", [1, 4, 6, 9, 13], True),
    "hubbard_L8_m32": ("snap_hubbard", "L=8 N=8 m=32 sweep=1 t1=-1 U=4 h_z=0.1", "all", True),
    "klm_L6_m32": ("snap_klm", "L=6 N=4 m=32 sweep=1 t=-1 J=2 h_z=0.1", "all", True),
    "klm_L6_m24_ixy": ("snap_klm", "L=6 N=6 m=24 sweep=1 t=-1 J=1.5 I_xy=0.5 I_z=0.3 h_z=0.05", [2, 4], True),
}
# per-step energies (energies.txt) and the program's own result files (result/: expectation
# values and correlation functions at the final sweeps' centre, entanglement entropy,
# truncation errors, energies), copied to <name>_result/ -- XXZ L=32 m=64, 4 sweeps
ENERGY_RUNS = {
    "xxz_L32_m64_sw4": ("snap_xxz", "L=32 m=64 sweep=4 J_xy=1.0 J_z=1.0 h_z=0.0 diag_acc=1e-12"),
}
# exact diagonalization through oracle/_ref/ed_ref (the reference's exact/ + onsite_op/):
# name -> harness arguments
ED_CASES = {
    "ed_xxz_L10": "model=xxz L=10 sz=0",
    "ed_xxz_L12_jz05_h": "model=xxz L=12 sz=2 J_z=0.5 h_z=0.3",
    "ed_xxz_L8_s1_dz": "model=xxz spin=2 L=8 sz=0 D_z=0.5 h_z=0.1",
    "ed_xxz_L10_pbc": "model=xxz L=10 sz=0 J_z=1.5 BC=PBC",
    "ed_xxz_L10_ssd": "model=xxz L=10 sz=0 BC=SSD",
    "ed_hubbard_L6": "model=hubbard L=6 N=6 sz=0 t=-1 U=4 h_z=0.1",
    "ed_hubbard_L6_v": "model=hubbard L=6 N=5 sz=1 t=-1 U=2 V=0.5 mu=0.3 h_z=0.2",
    "ed_hubbard_L6_pbc": "model=hubbard L=6 N=6 sz=0 t=-1 U=4 BC=PBC",
    "ed_hubbard_L5_pbc_odd": "model=hubbard L=5 N=3 sz=1 t=-1 U=3 BC=PBC",
    "ed_hubbard_L7": "model=hubbard L=7 N=7 sz=1 t=-1 U=4 h_z=0.1",
    "ed_hubbard_L6_ssd": "model=hubbard L=6 N=6 sz=0 t=-1 U=4 BC=SSD",
}
# superblock basis construction (GET_BASIS_SUPERBLOCK): the quantum numbers it starts from
# and the basis it builds, at a few steps per model (qn_<name>_s<step>.npz)
QN_CASES = {
    "xxz_L12_m24": ("snap_xxz", "L=12 m=24 sweep=1 J_xy=1.0 J_z=0.5 h_z=0.1", [1, 4, 9, 13]),
    "xxz_L16_s1_m32": ("snap_xxz", "L=16 m=32 sweep=1 spin=2 tot_sz=2 J_z=0.8", [3, 8, 14]),
    "hubbard_L8_m32": ("snap_hubbard", "L=8 N=8 m=32 sweep=1 t1=-1 U=4 h_z=0.1", [1, 3, 5, 7]),
    "hubbard_L12_N10_m48": ("snap_hubbard", "L=12 N=10 tot_sz=0 m=48 sweep=1 t1=-1 U=2", [2, 6, 11]),
    "klm_L6_m32": ("snap_klm", "L=6 N=4 m=32 sweep=1 t=-1 J=2 h_z=0.1", [1, 3, 5]),
}
# bench input (not a parity fixture): Hubbard L=64 N=64 m=400 U=4 h=0.1, last infinite step
# KLM L=48 N=36 m=800 (t=-1, J=2, h_z drawn from numpy default_rng(0).uniform(0, 0.3)), the
# last infinite step (48 sites); Lanczos_Slow keeps the reference's memory at 3 vectors
BIG = {
    "hubbard_L64_m400": ("snap_hubbard", "L=64 N=64 m=400 sweep=3 t1=-1 U=4 h_z=0.1", 31),
    "klm_L48_m800": ("snap_klm", "L=48 N=36 m=800 sweep=0 t=-1 J=2 h_z=0.19108850619643628 "
                     "method=Lanczos_Slow", 23),
    # served-stream superblock at the low end of m=200..1000: XXZ L=64 m=300, last infinite step
    "xxz_L64_m300": ("snap_xxz", "L=64 m=300 sweep=0 J_xy=1.0 J_z=1.0 h_z=0.0", 31),
}


def run(prog: str, args: str, workdir: str, extra: str, out: str = "") -> None:
    env = dict(os.environ, MKL_THREADING_LAYER="GNU", OMP_NUM_THREADS="8")
    cmd = [os.path.join(REF_BIN, prog)] + args.split() + extra.split() + [f"out={out or workdir}", "threads=8"]
    with open(os.path.join(workdir, "log.txt"), "w") as log:
        subprocess.run(cmd, cwd=workdir, env=env, stdout=log, stderr=subprocess.STDOUT, check=True)


def convert(workdir: str, name: str, steps, keep_vec: bool) -> None:
    for path in sorted(glob.glob(os.path.join(workdir, "snap_*.bin"))):
        step = int(os.path.basename(path)[5:9])
        if steps != "all" and step not in steps:
            continue
        snap = snapshot.load_bin(path)
        extra = {}
        gs = os.path.join(workdir, f"gs_{step:04d}.bin")
        if os.path.exists(gs):
            e, err, vec = snapshot.load_gs_bin(gs)
            extra["gs_energy"] = np.array([e])
            extra["gs_error"] = np.array([err])
            if keep_vec:
                extra["gs_vec"] = vec
        out = os.path.join(HERE, f"{name}_s{step:02d}.npz")
        snapshot.save_npz(snap, out, **extra)
        print("wrote", os.path.relpath(out, REPO), "n =", snap.n)
    for path in sorted(glob.glob(os.path.join(workdir, "expect_*.bin"))):
        step, k = int(os.path.basename(path)[7:11]), int(os.path.basename(path)[12:14])
        if steps != "all" and step not in steps:
            continue
        e = snapshot.load_expect_bin(path)
        out = os.path.join(HERE, f"{name}_expect_s{step:02d}_{k}.npz")
        snapshot.save_expect_npz(e, out)
        print("wrote", os.path.relpath(out, REPO), "kind =", e.kind)


def convert_qn(workdir: str, name: str, steps) -> None:
    """qn_<step>.bin (oracle/harness/snap.c: dump_qn) -> qn_<name>_s<step>.npz."""
    for path in sorted(glob.glob(os.path.join(workdir, "qn_*.bin"))):
        step = int(os.path.basename(path)[3:7])
        if step not in steps:
            continue
        with open(path, "rb") as f:
            assert f.read(8) == b"DMRGQN01"
            n_qn, dim_LL, dim_RR, d, t_sz, t_ele = np.frombuffer(f.read(24), "<i4")
            rd = lambda k, dt="<i4": np.frombuffer(f.read(k * np.dtype(dt).itemsize), dt).copy()
            out = dict(header=np.array([n_qn, dim_LL, dim_RR, d, t_sz, t_ele], np.int64),
                       sz_LL=rd(dim_LL), ele_LL=rd(dim_LL), sz_RR=rd(dim_RR), ele_RR=rd(dim_RR),
                       sz_site=rd(d), ele_site=rd(d))
            n = int(rd(1)[0])
            for k in ("LL", "LR", "RR", "RL"):
                out[k] = rd(n, "<i2")
        dst = os.path.join(HERE, f"qn_{name}_s{step:02d}.npz")
        np.savez_compressed(dst, **out)
        print("wrote", os.path.relpath(dst, REPO), "n =", n)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    ap.add_argument("--big", action="store_true")
    ap.add_argument("--big-only", default=None, help="with --big: only this BIG case")
    a = ap.parse_args()
    for name, (prog, args, steps, keep_vec) in CASES.items():
        if a.only and a.only != name:
            continue
        wd = tempfile.mkdtemp(prefix="golden_")
        dump = "all" if steps == "all" else ",".join(str(s) for s in steps)
        run(prog, args, wd, f"dump={dump}")
        convert(wd, name, steps, keep_vec)
        shutil.copy(os.path.join(wd, "energies.txt"), os.path.join(HERE, f"{name}_energies.txt"))
        shutil.rmtree(wd)
    for name, (prog, args, steps) in QN_CASES.items():
        if a.only and a.only != "qn_" + name:
            continue
        wd = tempfile.mkdtemp(prefix="golden_")
        run(prog, args, wd, "dump=" + ",".join(str(s) for s in steps) + " dump_qn=1 dump_snap=0 dump_vec=0")
        convert_qn(wd, name, steps)
        shutil.rmtree(wd)
    for name, (prog, args) in ENERGY_RUNS.items():
        if a.only and a.only != name:
            continue
        wd = tempfile.mkdtemp(prefix="golden_")
        run(prog, args, wd, "dump= dump_snap=0 dump_vec=0")
        shutil.copy(os.path.join(wd, "energies.txt"), os.path.join(HERE, f"{name}_energies.txt"))
        dst = os.path.join(HERE, f"{name}_result")
        shutil.rmtree(dst, ignore_errors=True)
        shutil.copytree(os.path.join(wd, "result"), dst, ignore=shutil.ignore_patterns("log.txt"))
        shutil.rmtree(wd)
        print("wrote", f"{name}_energies.txt", f"{name}_result/")
    for name, args in ED_CASES.items():
        if a.only and a.only != name:
            continue
        wd = tempfile.mkdtemp(prefix="golden_")
        path = os.path.join(wd, "ed.bin")
        run("ed_ref", args, wd, "", out=path)
        kv = dict(x.split("=", 1) for x in args.split())
        case = edhelp.load_ed_bin(path, kv)
        out = os.path.join(HERE, f"{name}.npz")
        edhelp.save_ed_npz(case, out)
        print("wrote", os.path.relpath(out, REPO), "n =", case.n, "E =", case.energy)
        shutil.rmtree(wd)
    if a.big:
        for name, (prog, args, step) in BIG.items():
            if a.big_only and a.big_only != name:
                continue
            wd = tempfile.mkdtemp(prefix="golden_")
            run(prog, args, wd, f"dump={step} dump_ham=0 dump_qn=1 stop_after={step}")
            convert_qn(wd, name, [step])   # qn_<name>_s<step>.npz: the same step's basis inputs
            snap = snapshot.load_bin(os.path.join(wd, f"snap_{step:04d}.bin"))
            e, err, _ = snapshot.load_gs_bin(os.path.join(wd, f"gs_{step:04d}.bin"))
            out = os.path.join(HERE, f"{name}_s{step:02d}.npz")
            snapshot.save_npz(snap, out, gs_energy=np.array([e]), gs_error=np.array([err]))
            print("wrote", os.path.relpath(out, REPO), "n =", snap.n)
            shutil.rmtree(wd)


if __name__ == "__main__":
    main()
