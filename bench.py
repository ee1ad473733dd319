#!/usr/bin/env python3
"""Ground-state solves/s of the superblock eigensolve on one MI355X.

Headline workload: the KLM superblock of the reference's own DMRG run at L=48, N=36, S^z=0,
t=1, J_K=2, h_z=0.19109 (drawn from seed 0 in [0, 0.3]), m=800 (last infinite-system step,
48 sites), stored as tests/golden/klm_L48_m800_s23.npz (written by tests/golden/make_golden.py
--big: block operators, term list, basis, the reference's energy); without that file the
Hubbard L=64 N=64 U=4 h=0.1 m=400 superblock (tests/golden/hubbard_L64_m400_s31.npz) is the
headline and otherwise a secondary line.  Inputs are uploaded to HBM before timing.

One step = one ground-state solve as GET_GROUND_STATE does it (Lanczos with the reference
defaults of main/dmrg/<model>_vf/src/main/main.c: "Lanczos", "Normal", diag_acc 1e-14, then
inverse iteration with CG, cg_acc 1e-7, ii_acc 1e-9; vector to host), applying the superblock
Hamiltonian in its Kronecker form: per sector of the superblock basis the LL/LR/LLLR terms as
a dense matrix on the rows of the sector's block of the vector and the RR/RL/RRRL terms as one
on its columns (FP64 matrix-core products), the coupling terms row by row; the form is built
on the device from the recorded term list in every step.  The same solve on the assembled CRS
(device assembly bit-identical to GET_HAM_LLLRRRRL, expansion, the CRS SpMV) is reported
beside it as forms.crs_fp64, with its SpMV roofline.

Reported beside it:
* roofline of the dominant kernel (the dense sector products), timed live with HIP events on
  the library's stream during the timed steps: floating-point operations per product
  sum over sectors of 2 da db (da + db) against the MI355X dense FP64 matrix peak (78.6
  TFLOP/s); spmv_roofline: the CRS SpMV of forms.crs_fp64, algorithmic bytes per launch
  against the 8 TB/s HBM peak, and its HBM traffic per launch from the committed rocprofv3
  --pmc summary;
* cpu_baseline: the reference's own code (oracle/_ref/libdmrgref.so, built from the
  reference sources, OpenMP, pinned to the process's cores): for the Hubbard superblock its
  DMRG_DIAGONALIZE_SUPERBLOCK_SYM, for the KLM superblock (one reference solve takes minutes)
  its MATRIX_VECTOR_PRODUCT_SYM on a bounded sample scaled by the products of one solve;
* forms: FP32-refined solves of the same superblock; batched (256 seeded ED problems, plain
  FP64 one by one against FP32-refined through the batched engine); served (Poisson
  arrivals at 200/s of ED problems and superblock snapshots through the engine).
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")
# superblock workloads: snapshot of the reference's own run, description, CPU-baseline form
SUPERBLOCKS = {
    "klm_L48_m800": (os.path.join(GOLDEN, "klm_L48_m800_s23.npz"),
                     "KLM OBC L=48 N=36 Sz=0 t=1 J_K=2 h_z=0.19109 (seed 0) m=800, last infinite step (48 sites)",
                     "products"),
    "hubbard_L64_m400": (os.path.join(GOLDEN, "hubbard_L64_m400_s31.npz"),
                         "Hubbard OBC L=64 N=64 Sz=0 t=1 U=4 h=0.1 m=400, last infinite step (64 sites)",
                         "solves"),
}
WORKLOAD = SUPERBLOCKS["hubbard_L64_m400"][0]
PROFILE_EVERY = 16
PEAK_HBM_GBS = 8000.0
PEAK_FP64_MFMA_TFLOPS = 78.6   # MI355X dense FP64 matrix peak
KRON_VARIANTS = ["FP64 vector units, 64 x 64 tiles", "FP64 MFMA, 64 x 64 tiles", "FP64 MFMA, 128 x 128 tiles",
                 "FP64 MFMA, 64 x 128 tiles", "FP64 MFMA, 128 x 64 tiles", "FP64 MFMA, 64 x 64 tiles, k-slabs of 32",
                 "FP64 MFMA, 64 x 64 tiles, four wavefronts per SIMD", "FP64 MFMA, 128 x 128 tiles, two wavefronts per SIMD"]
SPMV_FORMATS = [("CSR", "k_apply<CsrRows<double>> (CSR SpMV + fused Lanczos/CG epilogue)"),
                ("value dictionary", "k_apply<DictRows> (value-dictionary CRS SpMV + fused Lanczos/CG epilogue)"),
                ("delta columns", "k_apply<DeltaRows<double>> (delta-column CRS SpMV + fused Lanczos/CG epilogue)"),
                ("sliced", "k_apply<P, SellRows<double, P>> (sliced SpMV: 64/P-row slices, P lanes per row, "
                           "+ fused Lanczos/CG epilogue)")]
SPMV_FORMATS.append(("sliced, value indices", "k_apply<P, SellxRows<double, P>> (sliced SpMV with 16-bit column "
                     "differences and 16-bit value indices: 64/P-row slices, P lanes per row, + fused Lanczos/CG "
                     "epilogue)"))
SELL16_KERNEL = ("k_apply<P, Sell16Rows<double, P>> (sliced SpMV with 16-bit column differences: 64/P-row "
                 "slices, P lanes per row, + fused Lanczos/CG epilogue)")


def kernel_name(sched):
    fmt = (sched[2] >> 3) & 7
    return SELL16_KERNEL if fmt == 3 and sched[2] & 1 else SPMV_FORMATS[fmt][1]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_solve(D, sb, params, sched=None):
    A = sb.assemble(full=True)
    if sched:
        D.check(D.lib().dmrgx_matrix_set_schedule(A.handle, *sched))
    E, vec, st = A.ground_state(**params)
    nnz = A.nnz
    A.free()
    return E, st, nnz


def autotune(D, sb):
    """SpMV schedule (lanes per row, workgroup cap, load variant) for this workload, picked on
    the device by dmrgx_matrix_autotune; the superblocks of one DMRG run share their row
    statistics."""
    L = D.lib()
    i32p = C.POINTER(C.c_int32)
    L.dmrgx_matrix_autotune.argtypes = [C.c_void_p, i32p, i32p, i32p, C.POINTER(C.c_double)]
    L.dmrgx_matrix_set_schedule.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
    A = sb.assemble(full=True)
    g, cap, nt, us = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
    D.check(L.dmrgx_matrix_autotune(A.handle, C.byref(g), C.byref(cap), C.byref(nt), C.byref(us)))
    A.free()
    return (g.value, cap.value, nt.value), us.value


def cpu_env():
    """OpenMP placement for the reference on the host: one thread per core of this process's
    CPU set (at most 16: the host cores that go with one GPU), bound close on cores.  Read by libgomp when it
    starts, so set before any OpenMP library is loaded."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, cores))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    return threads


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_products(sb, n, steps_per_solve, budget_s):
    """The reference's own MATRIX_VECTOR_PRODUCT_SYM (sml/MATRIX_VECTOR_PRODUCT_SYM.c, OpenMP
    with per-thread scatter) on the superblock's lower triangle as the reference assembles it
    (the device assembly, equal to GET_HAM_LLLRRRRL's bit for bit), timed for a bounded number
    of calls and scaled to one solve by the products the reference's solve makes: one per
    Lanczos step (LANCZOS_SYM, the program's default method) and one per CG iteration, with
    the Lanczos and CG step counts of the GPU solve of the same superblock."""
    import dmrg_c_amd as D
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    ref = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")
    if not os.path.exists(ref):
        return None
    row, col, val = sb.download_lower()
    H = D.HostCRS(row, col, val)
    R = C.CDLL(ref)
    R.MATRIX_VECTOR_PRODUCT_SYM.argtypes = [C.POINTER(D.CRS1), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    temp = np.zeros((threads, n))
    rows_p = (C.c_void_p * threads)(*[temp[i].ctypes.data for i in range(threads)])
    x = np.random.default_rng(4).standard_normal(n)
    y = np.empty(n)
    R.MATRIX_VECTOR_PRODUCT_SYM(H.ptr, x.ctypes.data, y.ctypes.data, C.cast(rows_p, C.c_void_p), threads)  # warm
    ts = []
    t0 = time.perf_counter()
    while len(ts) < 3 or (time.perf_counter() - t0 < budget_s and len(ts) < 200):
        ta = time.perf_counter()
        R.MATRIX_VECTOR_PRODUCT_SYM(H.ptr, x.ctypes.data, y.ctypes.data, C.cast(rows_p, C.c_void_p), threads)
        ts.append(time.perf_counter() - ta)
    med = float(np.median(ts))
    per_solve = med * steps_per_solve
    return {
        "value": round(1.0 / per_solve, 5),
        "unit": "solves/s",
        "cores": threads,
        "kind": "reference",
        "cpu_model": cpu_model(),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "sample": (f"{len(ts)} calls of the reference's MATRIX_VECTOR_PRODUCT_SYM on the n={n} superblock "
                   f"(median {med * 1e3:.1f} ms), scaled by the {steps_per_solve} products of one solve "
                   "(Lanczos steps + CG iterations of the GPU solve); the Lanczos tridiagonal solves, vector "
                   "updates and assembly are not counted, which favours the CPU"),
    }


def cpu_baseline(snap, e_ref, budget_s):
    """The reference's own eigensolve on the host cores (bounded sample)."""
    import dmrg_c_amd as D
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_py as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")
    row, col, val = O.assemble(snap)
    model = cpu_model()
    kind, times, energies = None, [], []
    if os.path.exists(ref):
        os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
        try:
            R = C.CDLL(ref)
            R.DMRG_DIAGONALIZE_SUPERBLOCK_SYM.argtypes = [C.POINTER(D.CRS1), C.POINTER(D.DMRG_TIME),
                                                          C.POINTER(D.DMRG_PARAMETER), C.POINTER(D.DMRG_STATUS),
                                                          C.c_int]
            kind = "reference"
        except OSError as e:
            log(f"cpu baseline: cannot load {ref}: {e}; using the oracle port")
    old = os.getcwd()
    os.chdir(tempfile.mkdtemp(prefix="cpu_baseline_"))  # the reference appends to ./SML_out
    # the reference prints progress on C stdout: keep it off the JSON line
    sys.stdout.flush()
    saved_fd = os.dup(1)
    sink = os.open(os.devnull, os.O_WRONLY)
    os.dup2(sink, 1)
    try:
        t_start = None
        warm = 1
        while True:
            if kind == "reference":
                H = D.HostCRS(row, col, val)
                P = D.default_dmrg_parameter()
                st = D.DMRG_STATUS()
                vec = np.zeros(snap.n)
                st.GS_Vec = vec.ctypes.data_as(C.POINTER(C.c_double))
                t0 = time.perf_counter()
                R.DMRG_DIAGONALIZE_SUPERBLOCK_SYM(H.ptr, C.byref(D.DMRG_TIME()), C.byref(P), C.byref(st), threads)
                times.append(time.perf_counter() - t0)
                energies.append(st.gs_val)
            else:
                kind, threads = "port", 1
                t0 = time.perf_counter()
                rc, e, err, vec, ls, cs = O.diagonalize_superblock_sym(row, col, val, O.start_vector(snap.n, 1))
                times.append(time.perf_counter() - t0)
                energies.append(e)
            if warm:  # first solve pays page faults and library start-up: not counted
                warm = 0
                times.pop()
                energies.pop()
                t_start = time.perf_counter()
                continue
            log(f"cpu baseline solve {len(times)}: {times[-1]:.2f} s  E={energies[-1]:.15f}")
            if time.perf_counter() - t_start > budget_s or len(times) >= 10:
                break
    finally:
        C.CDLL(None).fflush(None)
        os.dup2(saved_fd, 1)
        os.close(saved_fd)
        os.close(sink)
        os.chdir(old)
    med = float(np.median(times))
    return {
        "value": round(1.0 / med, 4),
        "unit": "solves/s",
        "cores": threads,
        "kind": kind,
        "cpu_model": model,
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "sample": (f"{len(times)} solve(s) of the same n={snap.n} superblock after 1 untimed, median {med:.2f} s; "
                   + ("DMRG_DIAGONALIZE_SUPERBLOCK_SYM of the reference (LANCZOS_SYM + INVERSE_ITERATION_SYM, "
                      "-Ofast -fopenmp, MKL dstev_) on the oracle-assembled CRS; assembly not counted"
                      if kind == "reference" else "oracle port of the eigensolve, 1 thread; assembly not counted")),
        "max_rel_energy_error": float(max(abs(e - e_ref) / abs(e_ref) for e in energies)),
    }


def spmv_sweep(D, sb, snap):
    """Every schedule the autotuner considers, timed with 20 launches each: lanes per row x
    workgroup cap x load variant -> us per product + fused dot, and algorithmic GB/s."""
    A = sb.assemble(full=True)
    n = snap.n
    alg = A.nnz * 12 + (n + 1) * 8 + 2 * n * 8
    rows = []
    for v in range(40):
        if (v >> 3) == 4 and v & 1:
            continue
        for g in ((1, 2, 4) if (v >> 3) >= 3 else (4, 8, 16, 32, 64)):
            for cap in (1024, 2048, 4096, 8192, 16384, 0):
                us = A.time_schedule(g, cap, v, 20)
                rows.append({"lanes": g, "cap": cap, "variant": v, "us": round(us, 2),
                             "alg_GBs": round(alg / us / 1e3, 1)})
        log(f"sweep: variant {v} done")
    A.free()
    rows.sort(key=lambda r: r["us"])
    best = {}
    for r in rows:
        best.setdefault(r["variant"], r)
    for v in sorted(best):
        log(f"sweep: best of variant {v:2d}: {best[v]}")
    print(json.dumps({"spmv_sweep": rows, "n": n, "alg_bytes": alg}))


def spmv_alg_bytes(D, A, sched, n):
    """Algorithmic bytes of one product in the format the schedule reads: CSR (8-byte
    values, 4-byte columns) or the value dictionary (8-byte packed entries + the table)."""
    if ((sched[2] >> 3) & 7) == 4:
        nsl = (n + 64 // sched[0] - 1) // (64 // sched[0])
        exc = A.sell16_exceptions(sched[0])
        L = D.lib()
        L.dmrgx_matrix_dict_values.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        nv = C.c_int64()
        D.check(L.dmrgx_matrix_dict_values(A.handle, C.byref(nv)))
        if 0 < nv.value <= 65536:
            return (A.nnz * 4 + n * 4 + (nsl + 1) * 8 + nsl * 64 * 12 + exc * 4 + nv.value * 8 + 2 * n * 8,
                    f"sliced, 16-bit columns and value indices ({sched[0]} lanes per row; {nv.value} distinct "
                    f"values, {exc} columns kept whole, padding not counted)")
        sched = (sched[0], sched[1], (sched[2] & 7) | 24 | 1)   # read as format 3 with 16-bit columns
    if ((sched[2] >> 3) & 7) == 3:
        nsl = (n + 64 // sched[0] - 1) // (64 // sched[0])
        if sched[2] & 1:
            exc = A.sell16_exceptions(sched[0])
            return (A.nnz * 10 + n * 4 + (nsl + 1) * 8 + nsl * 64 * 12 + exc * 4 + 2 * n * 8,
                    f"sliced, 16-bit columns ({sched[0]} lanes per row; 8-byte values, 2-byte column differences, "
                    f"{exc} columns kept whole, padding not counted)")
        return (A.nnz * 12 + n * 4 + (nsl + 1) * 8 + 2 * n * 8,
                f"sliced ({sched[0]} lanes per row; 8-byte values, 4-byte columns, padding not counted)")
    if ((sched[2] >> 3) & 7) == 2:
        return (A.nnz * 10 + (n + 1) * 8 + n * 4 + 2 * n * 8,
                "delta columns (8-byte values, 2-byte column differences, 4-byte first columns)")
    if ((sched[2] >> 3) & 7) == 1:
        L = D.lib()
        L.dmrgx_matrix_dict_values.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        nv = C.c_int64()
        D.check(L.dmrgx_matrix_dict_values(A.handle, C.byref(nv)))
        if nv.value > 0:
            return A.nnz * 8 + nv.value * 8 + (n + 1) * 8 + 2 * n * 8, f"value dictionary ({nv.value} values)"
    return A.nnz * 12 + (n + 1) * 8 + 2 * n * 8, "CSR"


def traffic_probe(D, L, sb, snap, fixed=None):
    L.dmrgx_device_alloc.argtypes = [C.c_int64, C.POINTER(C.c_void_p)]
    L.dmrgx_device_free.argtypes = [C.c_void_p]
    L.dmrgx_copy_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    sched = fixed if fixed else autotune(D, sb)[0]
    L.dmrgx_matrix_set_schedule.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
    A = sb.assemble(full=True)
    D.check(L.dmrgx_matrix_set_schedule(A.handle, *sched))
    n, ncopy = snap.n, 1 << 26
    bufs = []
    for nbytes in (8 * n, 8 * n, 8 * ncopy, 8 * ncopy):
        p = C.c_void_p()
        D.check(L.dmrgx_device_alloc(nbytes, C.byref(p)))
        bufs.append(p)
    x, y, src, dst = bufs
    D.check(L.dmrgx_copy_device(src, dst, ncopy))  # first touch
    for _ in range(5):
        D.check(L.dmrgx_spmv_device(A.handle, x, y))
    D.check(L.dmrgx_copy_device(src, dst, ncopy))
    alg, fmt = spmv_alg_bytes(D, A, sched, n)
    print(json.dumps({"spmv_launches": 5, "n": n, "nnz": A.nnz, "spmv_alg_bytes": alg, "format": fmt,
                      "schedule": sched,
                      "copy_doubles": ncopy, "copy_bytes_read": 8 * ncopy, "copy_bytes_written": 8 * ncopy}))
    for p in bufs:
        L.dmrgx_device_free(p)


# ------------------------------------------------------------------ batched ED workload
def ed_batch_models(count=256, seed=0):
    """The batched-ED mix: XXZ chains L in {12, 14, ..., 24} at S^z = 0 (odd L has no S^z = 0
    sector for spin 1/2) and Hubbard chains L in {6, ..., 10} at half filling (S^z = 0, or
    S^z = 1/2 for odd L); J_z ~ U[0.5, 1.5], h ~ U[0, 0.5], t = 1, U = 4; one seeded draw."""
    from dmrg_c_amd import ed
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if rng.random() < 0.5:
            L = int(rng.choice(np.arange(12, 25, 2)))
            out.append(ed.Model("xxz", L=L, sz=0, J_z=float(rng.uniform(0.5, 1.5)), h_z=float(rng.uniform(0, 0.5))))
        else:
            L = int(rng.integers(6, 11))
            out.append(ed.Model("hubbard", L=L, N=L, sz=L % 2, t=-1.0, U=4.0, h_z=float(rng.uniform(0, 0.5))))
    return out


def run_ed_batch(D, count, precision_forms=(0, 1), single=True):
    from dmrg_c_amd import ed, engine
    models = ed_batch_models(count)
    t0 = time.perf_counter()
    mats = []
    for m in models:
        P = ed.EdProblem.from_model(m)
        mats.append(P.assemble())
        P.free()
    t_asm = time.perf_counter() - t0
    # SpMV schedules: the problems of one model and length share their row statistics, so
    # the first large one of each is autotuned and the others take its schedule (setup, as
    # a caller sweeping couplings would do; not timed)
    tuned = {}
    for m, A in zip(models, mats):
        key = (m.model, m.L)
        if A.n < 50000:
            continue
        if key not in tuned:
            tuned[key] = A.autotune()[0]
            log(f"ed-batch: schedule for {key} (n={A.n}): {tuned[key]}")
        else:
            A.set_schedule(*tuned[key])
    dims = [A.n for A in mats]
    log(f"ed-batch: {count} problems, n from {min(dims)} to {max(dims)}, total rows {sum(dims)}, "
        f"device assembly {t_asm:.2f} s")
    params = dict(diag_acc=1e-14, diag_max_step=800, cg_acc=1e-7, cg_max_step=2000, ii_acc=1e-9,
                  ii_diag_add=1e-11, ii_max_step=3, seed=1)
    res = {"problems": count, "n_min": min(dims), "n_max": max(dims), "rows_total": int(sum(dims)),
           "device_assembly_s": round(t_asm, 3),
           "schedules": {f"{k[0]} L={k[1]}": list(v) for k, v in tuned.items()}}
    single_e = None
    for prec in precision_forms:
        p = dict(params, precision=prec)
        t0 = time.perf_counter()
        es = [A.ground_state(**p)[0] for A in mats] if single else None
        t_single = time.perf_counter() - t0
        ep = engine.default_params(**p)

        def batch(use_graphs):
            t0 = time.perf_counter()
            E = engine.Engine(use_graphs=use_graphs)
            for i, A in enumerate(mats):
                E.submit(A, ep, tag=i)
            out = sorted(E.run_until_done(want_vec=False), key=lambda r: r.tag)
            t = time.perf_counter() - t0
            host, graphs = E.host_times(), E.graph_stats()
            E.free()
            return out, t, host, graphs

        out_eager, t_eager, _, _ = batch(False)
        out, t_batch, host, graphs = batch(True)
        same_eager = all(a.energy == b.energy for a, b in zip(out, out_eager))
        host["graphs"] = graphs
        log(f"ed-batch precision {prec}: engine host seconds {host}; without graphs {count / t_eager:.1f}/s, "
            f"graph results identical to eager: {same_eager}")
        eb = [o.energy for o in out]
        tag = "fp64" if prec == 0 else "fp32_refined"
        g = {"batched_solves_per_s_without_graphs": round(count / t_eager, 2), "graphs_equal_eager": same_eager}
        if not single:
            res[tag] = {"batched_solves_per_s": round(count / t_batch, 2), "engine_host_s": host, **g}
            log(f"ed-batch {tag}: batched {count / t_batch:.1f}/s")
            continue
        same = all(a == b for a, b in zip(es, eb))
        if prec == 0:
            single_e = es
            err = 0.0
        else:
            err = max(abs(a - b) / abs(b) for a, b in zip(eb, single_e))
        res[tag] = {"single_solves_per_s": round(count / t_single, 2), "batched_solves_per_s": round(count / t_batch, 2),
                    "batched_over_single": round(t_single / t_batch, 2), "batched_equals_single": same,
                    "max_rel_energy_diff_vs_fp64": err, "batched_wall_s": round(t_batch, 3), "engine_host_s": host, **g}
        log(f"ed-batch {tag}: single {count / t_single:.1f}/s, batched {count / t_batch:.1f}/s, identical={same}")
    for A in mats:
        A.free()
    return res


# ------------------------------------------------------------------ basis construction
BASIS_QN = os.path.join(REPO, "tests", "golden", "qn_hubbard_L64_m400_s31.npz")


def run_basis(D, reps: int, cpu: bool):
    """Superblock basis of the Hubbard L=64 m=400 step (the bench superblock, n=144542) from
    the blocks' quantum numbers the reference run had: dmrgx_superblock_basis (sectors on the
    host, rows on the device, arrays back in host memory) against the reference's own arrays,
    and the oracle's restatement of the reference's enumerate-and-sort on one host core."""
    z = np.load(BASIS_QN)
    n_qn, dim_LL, dim_RR, d, t_sz, t_ele = (int(v) for v in z["header"])
    q = dict(n_qn=n_qn, target_sz=t_sz, target_ele=t_ele,
             **{k: z[k] for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR", "sz_site", "ele_site")})
    ref = tuple(z[k] for k in ("LL", "LR", "RR", "RL"))
    got = D.superblock_basis(q)   # warm-up
    same = all(np.array_equal(a, b) for a, b in zip(got, ref))
    ts = []
    for _ in range(max(reps, 5)):
        t0 = time.perf_counter()
        D.superblock_basis(q)
        ts.append(time.perf_counter() - t0)
    res = {"n": len(ref[0]), "dim_LL": dim_LL, "dim_RR": dim_RR, "dim_onsite": d,
           "product_space": dim_LL * d * dim_RR * d, "builds_per_s": round(1.0 / float(np.median(ts)), 2),
           "ms_per_build": round(1e3 * float(np.median(ts)), 3), "equals_reference_basis": same,
           "note": "quantum numbers dumped from the reference's own run (tests/golden/qn_hubbard_L64_m400_s31.npz)"}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_py as O
        t0 = time.perf_counter()
        k = 0
        while k < 3 or time.perf_counter() - t0 < 2.0:
            O.superblock_basis(q)
            k += 1
        res["cpu_baseline"] = {"builds_per_s": round(k / (time.perf_counter() - t0), 2), "cores": 1,
                               "kind": "port", "sample": f"{k} builds: the reference's enumeration of the "
                                                         "product space and quicksort order, restated in C"}
    return res


def run_v_m(D, reps: int, cpu: bool):
    """Single-slot product into a neighbouring sector (DMRG_V_M_LL_Q2 form: an operator on
    LL that raises S^z by one unit) on the Hubbard L=64 m=400 step: whole product space
    396*4*396*4, input the n=144542 target sector, output its S^z+2 sector.  The whole
    basis (sectors keyed by electrons and S^z, enumeration order, the target in the
    reference's order) is built from the quantum numbers of the reference's own run; the
    operator is a seeded random matrix with the S^+ sparsity pattern.  Times
    dmrgx_v_m_sector per call (operator and vectors from and to host memory, as the
    reference's signature has them) and the oracle's restatement on one host core."""
    z = np.load(BASIS_QN)
    n_qn, dim_LL, dim_RR, d, t_sz, t_ele = (int(v) for v in z["header"])
    qn = {k: z[k].astype(np.int64) for k in ("sz_LL", "ele_LL", "sz_RR", "ele_RR", "sz_site", "ele_site")}
    dims = (dim_LL, d, dim_RR, d)
    LL, LR, RR, RL = (a.ravel() for a in np.meshgrid(*[np.arange(k) for k in dims], indexing="ij"))
    ele = qn["ele_LL"][LL] + qn["ele_site"][LR] + qn["ele_RR"][RR] + qn["ele_site"][RL]
    sz = qn["sz_LL"][LL] + qn["sz_site"][LR] + qn["sz_RR"][RR] + qn["sz_site"][RL]
    inv = np.full(LL.size, -1, np.int32)
    sec = {}
    for key in ((t_ele, t_sz), (t_ele, t_sz + 2)):
        if key == (t_ele, t_sz):
            flat = np.ravel_multi_index(tuple(np.asarray(z[k], np.int64) for k in ("LL", "LR", "RR", "RL")), dims)
        else:
            flat = np.nonzero((ele == key[0]) & (sz == key[1]))[0]
        inv[flat] = np.arange(flat.size, dtype=np.int32)
        sec[key] = [a[flat].astype(np.int16) for a in (LL, LR, RR, RL)]
    rng = np.random.default_rng(5)
    e, s = qn["ele_LL"], qn["sz_LL"]
    mask = (e[:, None] == e[None, :]) & (s[:, None] == s[None, :] + 2) & (rng.random((dim_LL, dim_LL)) < 0.5)
    dense = np.where(mask, rng.uniform(-1, 1, mask.shape), 0.0)
    import scipy.sparse as sp
    m = sp.csr_matrix(dense)
    op = (m.indptr.astype(np.int64), m.indices.astype(np.int32), m.data.copy())
    n_in, n_out = sec[(t_ele, t_sz)][0].size, sec[(t_ele, t_sz + 2)][0].size
    x = rng.uniform(-1, 1, n_in)
    W = D.WholeBasis(dim_LL, dim_RR, d, inv)
    try:
        sid = W.add_sector(*sec[(t_ele, t_sz + 2)])
        y = W.v_m(0, op, sid, x)
        ts = []
        for _ in range(max(reps, 10)):
            t0 = time.perf_counter()
            W.v_m(0, op, sid, x)
            ts.append(time.perf_counter() - t0)
    finally:
        W.free()
    ob = sec[(t_ele, t_sz + 2)]
    entries = int(sum(op[0][r + 1] - op[0][r] for r in ob[0]))
    res = {"n_in": n_in, "n_out": n_out, "product_space": int(LL.size), "operator_nnz": len(op[2]),
           "entries_gathered": entries, "products_per_s": round(1.0 / float(np.median(ts)), 2),
           "ms_per_product": round(1e3 * float(np.median(ts)), 3),
           "note": "host-inclusive: operator upload, input vector up, output vector down; "
                   "quantum numbers and target basis of the reference's own run"}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import ctypes as C
        import oracle_py as O
        fn = O.lib().or_v_m_sector
        fn.restype = None
        fn.argtypes = [C.c_int, C.c_int32, C.c_int32, C.c_int32] + [C.c_void_p] * 11 + [C.c_int32, C.c_void_p,
                                                                                          C.c_void_p]
        eles = [np.ascontiguousarray(qn[k], np.int32) for k in ("ele_LL", "ele_site", "ele_RR")]
        yc = np.empty(n_out)
        args = [0, n_out, dim_RR, d, *[a.ctypes.data for a in ob], inv.ctypes.data,
                *[a.ctypes.data for a in op], *[a.ctypes.data for a in eles], 0, x.ctypes.data, yc.ctypes.data]
        t0 = time.perf_counter()
        k = 0
        while k < 3 or time.perf_counter() - t0 < 2.0:
            fn(*args)
            k += 1
        res["equals_oracle"] = bool(np.array_equal(y, yc))
        res["cpu_baseline"] = {"products_per_s": round(k / (time.perf_counter() - t0), 2), "cores": 1,
                               "kind": "port", "sample": f"{k} products: the reference's loop restated in C"}
    return res


def run_density(D, reps: int, cpu: bool):
    """Reduced density matrices of every LLLR block (DMRG_GET_DENSITY_MATRIX for all blocks)
    of the Hubbard L=64 m=400 step: the reference's own target-sector basis (n=144542), its
    LLLR states by first appearance, blocks of equal (electrons, S^z), a seeded unit
    vector.  dmrgx_density_matrices per call (vector up, matrices down) against the oracle's
    restatement of the reference's loop on one host core."""
    import ctypes as C
    z = np.load(BASIS_QN)
    n_qn, dim_LL, dim_RR, d, _, _ = (int(v) for v in z["header"])
    LL, LR, RR, RL = (np.ascontiguousarray(z[k], np.int16) for k in ("LL", "LR", "RR", "RL"))
    n = LL.size
    dims = (dim_LL, d, dim_RR, d)
    inv = np.full(int(np.prod(dims)), -1, np.int32)
    inv[np.ravel_multi_index(tuple(a.astype(np.int64) for a in (LL, LR, RR, RL)), dims)] = np.arange(n, dtype=np.int32)
    key = LL.astype(np.int64) * d + LR
    first = np.r_[True, key[1:] != key[:-1]]
    LL_LLLR, LR_LLLR = np.ascontiguousarray(LL[first]), np.ascontiguousarray(LR[first])
    Sum = np.diff(np.r_[np.nonzero(first)[0], n]).astype(np.int32)
    q = np.stack([z["ele_LL"][LL_LLLR] + z["ele_site"][LR_LLLR], z["sz_LL"][LL_LLLR] + z["sz_site"][LR_LLLR]], 1)
    starts = np.r_[0, np.nonzero(np.any(q[1:] != q[:-1], axis=1))[0] + 1, LL_LLLR.size]
    dim_block = np.diff(starts).astype(np.int32)
    v = np.random.default_rng(3).standard_normal(n)
    v /= np.linalg.norm(v)
    B = D.DMRG_BASIS()
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    B.LL_LLLRRRRL, B.LR_LLLRRRRL, B.RL_LLLRRRRL, B.RR_LLLRRRRL = (P(a, C.c_short) for a in (LL, LR, RL, RR))
    B.Inv_LLLRRRRL = P(inv, C.c_int)
    B.dim_LLLRRRRL, B.dim_LL, B.dim_RR, B.dim_onsite = n, dim_LL, dim_RR, d
    B.LL_LLLR, B.LR_LLLR, B.dim_LLLR, B.Sum_LLLR = P(LL_LLLR, C.c_short), P(LR_LLLR, C.c_short), LL_LLLR.size, \
        P(Sum, C.c_int)
    mats = D.density_matrices(B, dim_block, v)
    ts = []
    for _ in range(max(reps, 10)):
        t0 = time.perf_counter()
        D.density_matrices(B, dim_block, v)
        ts.append(time.perf_counter() - t0)
    flops = float(np.sum(dim_block.astype(np.float64) ** 2 * Sum[starts[:-1]]))
    res = {"n": n, "dim_LLLR": int(LL_LLLR.size), "blocks": len(dim_block), "max_block": int(dim_block.max()),
           "matrix_entries": int(np.sum(dim_block.astype(np.int64) ** 2)),
           "multiply_adds_full_blocks": flops, "calls_per_s": round(1.0 / float(np.median(ts)), 2),
           "ms_per_call": round(1e3 * float(np.median(ts)), 3),
           "note": "host-inclusive (vector up, matrices down); basis and quantum numbers of the reference's own run"}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_py as O
        fn = O.lib().or_density_matrix
        fn.restype = C.c_int
        fn.argtypes = [C.c_int32] * 4 + [C.c_void_p] * 7 + [C.c_int32, C.c_int32, C.c_void_p]
        outs = [np.empty((k, k)) for k in dim_block]
        t0 = time.perf_counter()
        k = 0
        while k < 2 or time.perf_counter() - t0 < 3.0:
            for b, o in enumerate(outs):
                fn(n, dim_RR, d, LL_LLLR.size, LL_LLLR.ctypes.data, LR_LLLR.ctypes.data, Sum.ctypes.data,
                   RR.ctypes.data, RL.ctypes.data, inv.ctypes.data, v.ctypes.data, int(starts[b]), int(dim_block[b]),
                   o.ctypes.data)
            k += 1
        res["equals_oracle"] = all(np.array_equal(a, b) for a, b in zip(mats, outs))
        res["cpu_baseline"] = {"calls_per_s": round(k / (time.perf_counter() - t0), 2), "cores": 1, "kind": "port",
                               "sample": f"{k} passes over all blocks: the reference's loop restated in C"}
    # the whole truncation step: DMRG_GET_SYSTEM_INFO_Q2 with max_dim 400 (the run's m)
    q1 = np.ascontiguousarray(q[:, 0], np.int32)
    q2 = np.ascontiguousarray(q[:, 1], np.int32)
    info = D.system_info(B, 2, q1, q2, v, 400)
    L = D.lib()
    ts = []
    for _ in range(max(reps // 2, 5)):
        out = C.POINTER(D.DMRG_SYSTEM_INFO)()
        t0 = time.perf_counter()
        D.check(L.dmrgx_get_system_info(2, q1.ctypes.data, q2.ctypes.data, C.byref(B), v.ctypes.data, 400,
                                        C.byref(out)))
        ts.append(time.perf_counter() - t0)
        L.dmrgx_free_system_info(out)
    tr = {"max_dim": 400, "dim_renorm": info["dim_renorm"], "tr_error": info["tr_error"],
          "calls_per_s": round(1.0 / float(np.median(ts)), 2), "ms_per_call": round(1e3 * float(np.median(ts)), 3),
          "note": "density matrices, block eigenpairs (blocks up to 192 states: the library's own "
                  "solver, eigenvalues of all states and eigenvectors of the kept ones; larger: "
                  "rocSOLVER dsyevd), selection and transformation matrices; host-inclusive, result back "
                  "in host memory"}
    ref_lib = os.path.join(REPO, "oracle", "_ref", "libdmrgref.so")
    if cpu and os.path.exists(ref_lib):
        os.environ["MKL_NUM_THREADS"] = "1"   # read by LAPACK at its first call: one core, as reported
        R = C.CDLL(ref_lib)
        R.DMRG_GET_SYSTEM_INFO_Q2.restype = C.POINTER(D.DMRG_SYSTEM_INFO)
        R.DMRG_GET_SYSTEM_INFO_Q2.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(D.DMRG_BASIS), C.c_void_p, C.c_int,
                                              C.c_int, C.POINTER(D.DMRG_TIME)]
        R.DMRG_FREE_SYSTEM_INFO.argtypes = [C.POINTER(D.DMRG_SYSTEM_INFO)]
        t = D.DMRG_TIME()
        t0 = time.perf_counter()
        k = 0
        pr = R.DMRG_GET_SYSTEM_INFO_Q2(q1.ctypes.data, q2.ctypes.data, C.byref(B), v.ctypes.data, 400, 1,
                                       C.byref(t))
        ref_info = D.system_info_dict(pr.contents, 2)
        R.DMRG_FREE_SYSTEM_INFO(pr)
        tr["dim_renorm_equals_reference"] = ref_info["dim_renorm"] == info["dim_renorm"]
        tr["max_abs_val_dm_diff_vs_reference"] = float(np.abs(ref_info["val_dist"] - info["val_dist"]).max())
        t0 = time.perf_counter()
        k = 0
        while k < 2 or time.perf_counter() - t0 < 3.0:
            pr = R.DMRG_GET_SYSTEM_INFO_Q2(q1.ctypes.data, q2.ctypes.data, C.byref(B), v.ctypes.data, 400, 1,
                                           C.byref(t))
            R.DMRG_FREE_SYSTEM_INFO(pr)
            k += 1
        tr["cpu_baseline"] = {"calls_per_s": round(k / (time.perf_counter() - t0), 2), "cores": 1,
                              "kind": "reference", "sample": f"{k} calls of the reference's own "
                              "DMRG_GET_SYSTEM_INFO_Q2 (LAPACK dsyev), p_threads=1"}
        t0 = time.perf_counter()
        k = 0
        while k < 2 or time.perf_counter() - t0 < 3.0:
            pr = R.DMRG_GET_SYSTEM_INFO_Q2(q1.ctypes.data, q2.ctypes.data, C.byref(B), v.ctypes.data, 400, 16,
                                           C.byref(t))
            R.DMRG_FREE_SYSTEM_INFO(pr)
            k += 1
        tr["cpu_baseline_16_threads"] = {"calls_per_s": round(k / (time.perf_counter() - t0), 2), "cores": 16,
                                         "kind": "reference", "sample": f"{k} calls, p_threads=16 (blocks in "
                                         "parallel, density matrices and dsyev), LAPACK single-threaded"}
    res["system_info"] = tr
    # renormalisation of one LLLR operator: CRS_CRS_CCS_PRODUCT(T^dagger, O, T) with the kept
    # transformation matrix and a seeded sparse symmetric O (1 % of the entries)
    T = info["T"]
    m, kk = T.shape
    rng = np.random.default_rng(12)
    Od = np.where(rng.random((m, m)) < 0.005, rng.uniform(-1, 1, (m, m)), 0.0)
    Od = Od + Od.T
    import scipy.sparse as sp
    Oc, Tt, Tc = sp.csr_matrix(Od), sp.csr_matrix(T.T), sp.csc_matrix(T)
    mk = lambda a: [np.ascontiguousarray(x) for x in (a.indptr.astype(np.int64), a.indices.astype(np.int32), a.data)]
    M1, M2, M3 = mk(Tt), mk(Oc), mk(Tc)

    def crs(a, n, cdim):
        return D.CRS1(len(a[2]), n + 1, n, cdim, a[1].ctypes.data_as(C.POINTER(C.c_int)),
                      a[0].ctypes.data_as(C.POINTER(C.c_long)), a[2].ctypes.data_as(C.POINTER(C.c_double)))
    m1, m2 = crs(M1, kk, m), crs(M2, m, m)
    m3 = D.CCS1(len(M3[2]), kk + 1, m, kk, M3[0].ctypes.data_as(C.POINTER(C.c_long)),
                M3[1].ctypes.data_as(C.POINTER(C.c_int)), M3[2].ctypes.data_as(C.POINTER(C.c_double)))
    bufs = [np.zeros(kk + 1, np.int64), np.zeros(kk * kk, np.int32), np.zeros(kk * kk),
            np.zeros(kk + 1, np.int64), np.zeros(m * kk, np.int32), np.zeros(m * kk), np.zeros(m)]
    out = D.CRS1(kk * kk, kk + 1, 0, 0, bufs[1].ctypes.data_as(C.POINTER(C.c_int)),
                 bufs[0].ctypes.data_as(C.POINTER(C.c_long)), bufs[2].ctypes.data_as(C.POINTER(C.c_double)))
    work = D.CCS1(m * kk, kk + 1, 0, 0, bufs[3].ctypes.data_as(C.POINTER(C.c_long)),
                  bufs[4].ctypes.data_as(C.POINTER(C.c_int)), bufs[5].ctypes.data_as(C.POINTER(C.c_double)))
    argt = [C.POINTER(D.CRS1), C.POINTER(D.CRS1), C.POINTER(D.CCS1), C.POINTER(D.CRS1), C.POINTER(D.CCS1),
            C.c_void_p]

    def run(fn):
        fn.argtypes = argt
        fn(C.byref(m1), C.byref(m2), C.byref(m3), C.byref(out), C.byref(work), bufs[6].ctypes.data)
        n = int(bufs[0][kk])
        return bufs[0].copy(), bufs[1][:n].copy(), bufs[2][:n].copy()
    got = run(L.CRS_CRS_CCS_PRODUCT)
    ts = []
    for _ in range(max(reps // 2, 5)):
        t0 = time.perf_counter()
        run(L.CRS_CRS_CCS_PRODUCT)
        ts.append(time.perf_counter() - t0)
    rn = {"dim_LLLR": m, "dim_renorm": kk, "operator_nnz": int(len(M2[2])), "out_nnz": int(got[0][kk]),
          "calls_per_s": round(1.0 / float(np.median(ts)), 2), "ms_per_call": round(1e3 * float(np.median(ts)), 3),
          "note": "drop-in CRS_CRS_CCS_PRODUCT, host arrays in and out"}
    if cpu and os.path.exists(ref_lib):
        Rf = C.CDLL(ref_lib)
        want = run(Rf.CRS_CRS_CCS_PRODUCT)
        rn["equals_reference"] = all(np.array_equal(a, b) for a, b in zip(got, want))
        t0 = time.perf_counter()
        k = 0
        while k < 2 or time.perf_counter() - t0 < 3.0:
            run(Rf.CRS_CRS_CCS_PRODUCT)
            k += 1
        rn["cpu_baseline"] = {"calls_per_s": round(k / (time.perf_counter() - t0), 2), "cores": 1,
                              "kind": "reference", "sample": f"{k} calls of the reference's own CRS_CRS_CCS_PRODUCT"}
    res["renormalise"] = rn
    return res


# ------------------------------------------------------------------ served stream
# superblock snapshots of the served stream: the reference's runs at m = 200..1000
SERVE_SNAPSHOTS = ["xxz_L64_m300_s31.npz", "hubbard_L64_m400_s31.npz", "klm_L48_m800_s23.npz"]


def run_serve(D, rate, duration, seed=0, chunk_ticks=1, precision=1, snapshots=None):
    """Poisson arrivals (rate/s, seeded) for `duration` s: 70 % ED (XXZ L 10..20, Hubbard
    L 6..10, random J_z, h), 30 % superblock snapshots of the reference's runs at m=300..800
    (those present in tests/golden); each request is assembled on the device when it
    arrives and joins the running engine (continuous batching).  Latency = completion -
    scheduled arrival.  precision 1: FP32-refined solves."""
    from dmrg_c_amd import ed, engine, snapshot
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=int(rate * duration * 2) + 10)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration]
    sbs, refs, sb_names = [], [], []
    for name in (snapshots or SERVE_SNAPSHOTS):
        path = os.path.join(REPO, "tests", "golden", name)
        if not os.path.exists(path):
            continue
        snap = snapshot.load(path)
        sbs.append(D.Superblock(snap))
        refs.append(float(snap.meta["gs_energy"][0]))
        sb_names.append(f"{name[:-4]} (n={snap.n})")
    reqs = []
    for _ in arrivals:
        if rng.random() < 0.7:
            if rng.random() < 0.5:
                L = int(rng.integers(10, 21))
                reqs.append(("ed", ed.Model("xxz", L=L, sz=L % 2, J_z=float(rng.uniform(0.5, 1.5)),
                                            h_z=float(rng.uniform(0, 0.5)))))
            else:
                L = int(rng.integers(6, 11))
                reqs.append(("ed", ed.Model("hubbard", L=L, N=L, sz=L % 2, t=-1.0, U=4.0,
                                            h_z=float(rng.uniform(0, 0.5)))))
        else:
            reqs.append(("sb", int(rng.integers(0, len(sbs)))))
    E = engine.Engine(chunk_ticks=chunk_ticks)
    p = engine.default_params(precision=precision)
    t0 = time.perf_counter()
    i, done = 0, {}
    errs = []
    host = {"assemble": 0.0, "engine_run": 0.0, "poll": 0.0}
    wait, service = {}, {}
    while len(done) < len(reqs):
        now = time.perf_counter() - t0
        ta = time.perf_counter()
        while i < len(reqs) and arrivals[i] <= now:
            kind, what = reqs[i]
            if kind == "ed":
                P = ed.EdProblem.from_model(what)
                A = P.assemble()
                P.free()
            else:
                A = sbs[what].assemble(full=True)
            E.submit(A, p, tag=i, own=True)
            i += 1
        tb = time.perf_counter()
        if E.pending > 0:
            E.run(1)
        elif i < len(reqs):
            time.sleep(max(0.0, arrivals[i] - (time.perf_counter() - t0)))
        tc = time.perf_counter()
        for r in E.poll(want_vec=False):
            if len(done) % 200 == 0:
                log(f"serve: {len(done)} of {len(reqs)} done, {i} arrived, t={time.perf_counter() - t0:.1f} s")
            done[r.tag] = r.t_done - t0 - arrivals[r.tag]
            wait[r.tag] = r.t_start - t0 - arrivals[r.tag]
            service[r.tag] = r.t_done - r.t_start
            kind, what = reqs[r.tag]
            if kind == "sb":
                errs.append(abs(r.energy - refs[what]) / abs(refs[what]))
        td = time.perf_counter()
        host["assemble"] += tb - ta
        host["engine_run"] += tc - tb
        host["poll"] += td - tc
    wall = time.perf_counter() - t0
    graphs = E.graph_stats()
    E.free()
    for sb in sbs:
        sb.free()
    lat = np.array([done[k] for k in sorted(done)])
    n_sb = sum(1 for k, _ in reqs if k == "sb")
    return {"requests": len(reqs), "rate_per_s": rate, "duration_s": duration, "ed_requests": len(reqs) - n_sb,
            "superblock_requests": n_sb, "solves_per_s": round(len(reqs) / wall, 2),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "max_rel_energy_error_superblocks": float(max(errs)) if errs else None,
            "engine_chunks": graphs,
            "host_s": {k: round(v, 3) for k, v in host.items()},
            "latency_split_ms": {
                part: {kind: {"p50": round(float(np.percentile(v, 50)) * 1e3, 2),
                              "p99": round(float(np.percentile(v, 99)) * 1e3, 2)}
                       for kind, v in (("ed", [d[k] for k in d if reqs[k][0] == "ed"]),
                                       ("superblock", [d[k] for k in d if reqs[k][0] == "sb"])) if v}
                for part, d in (("until_admitted", wait), ("in_engine", service))},
            "chunk_ticks": chunk_ticks,
            "precision": "fp32_refined" if precision else "fp64",
            "superblocks": sb_names,
            "note": "superblock requests drawn uniformly from the reference-run snapshots listed, "
                    "assembled on the device at arrival"}


def run_kron(D, L, name, steps, warmup, lan_type=0, profile=True):
    """Solves/s of the superblock workload on the Kronecker form of its Hamiltonian: per solve
    the form's build from the resident superblock (sector tables, the dense LLLR and RRRL
    sector matrices from the term list) + Lanczos + inverse iteration/CG (the reference's
    algorithm and parameters, as the CRS path) + vector to host; the live roofline of the
    dense sector products (HIP events on the library stream, one in PROFILE_EVERY)."""
    from dmrg_c_amd import snapshot
    path, desc, _ = SUPERBLOCKS[name]
    snap = snapshot.load(path)
    e_ref = float(snap.meta["gs_energy"][0])
    params = dict(diag_acc=1e-14, diag_max_step=800, cg_acc=1e-7, cg_max_step=2000, ii_acc=1e-9,
                  ii_diag_add=1e-11, ii_max_step=3, seed=1, lan_type=lan_type)
    sb = D.Superblock(snap)
    K = sb.kron()
    variant, tuned_us = K.autotune()
    K.free()
    log(f"kron {name}: {K.sectors} sectors, {K.coupling_terms} coupling terms, {K.flops / 1e9:.3f} GFLOP per "
        f"product; autotune: variant {variant} ({tuned_us:.1f} us per product)")

    def solve():
        Kf = sb.kron()
        Kf.set_kernel(variant)
        E, vec, st = Kf.ground_state(**params)
        Kf.free()
        return E, st

    for i in range(warmup):
        E, st = solve()
        log(f"kron warmup {i}: E={E:.15f} lanczos={st.lanczos_steps} cg={st.cg_steps}")
    if profile:
        D.check(L.dmrgx_profile_begin_sampled(PROFILE_EVERY))
    lat, errs, lan, cg, gerr = [], [], [], [], []
    t0 = time.perf_counter()
    for i in range(steps):
        ts = time.perf_counter()
        E, st = solve()
        lat.append(time.perf_counter() - ts)
        errs.append(abs(E - e_ref) / abs(e_ref))
        lan.append(st.lanczos_steps)
        cg.append(st.cg_steps)
        gerr.append(st.gs_error)
        log(f"kron step {i}: {lat[-1] * 1e3:.2f} ms  E={E:.15f}  dE/E={errs[-1]:.2e}")
    total = time.perf_counter() - t0
    ms, launches, flops = C.c_double(), C.c_int64(), C.c_double()
    if profile:
        L.dmrgx_profile_end.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        D.check(L.dmrgx_profile_end(C.byref(ms), C.byref(launches), C.byref(flops)))
    sb.free()
    if not profile:
        return {"solves_per_s": round(steps / total, 4), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                "max_rel_energy_error": float(max(errs)), "max_gs_error": float(max(gerr)),
                "lanczos_steps": int(np.median(lan)), "cg_steps": int(np.median(cg))}
    avg_s = ms.value / 1e3 / max(launches.value, 1)
    fl = flops.value / max(launches.value, 1)
    achieved = fl / avg_s / 1e12 if avg_s > 0 else 0.0
    v = variant % len(KRON_VARIANTS)
    return {
        "value": round(steps / total, 4),
        "ms_per_step": round(total / steps * 1e3, 3),
        "data": f"superblock snapshot from the reference's own DMRG run (tests/golden/{os.path.basename(path)})",
        "config": {
            "workload": f"{desc}, superblock n={snap.n}: Kronecker form of the superblock Hamiltonian built on the "
                        "device from the recorded term list (no CRS assembly) + Lanczos + inverse iteration/CG "
                        "per solve",
            "n": snap.n,
            "sectors": K.sectors,
            "coupling_terms": K.coupling_terms,
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "max_rel_energy_error": float(max(errs)),
            "max_gs_error": float(max(gerr)),
            "lanczos_steps": int(np.median(lan)),
            "cg_steps": int(np.median(cg)),
            "product_kernel": {"variant": variant, "form": KRON_VARIANTS[v],
                               "xcd_contiguous_tiles": variant >= len(KRON_VARIANTS)},
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "k_kron_tile<64, 64, true> (dense sector products M_L X and X M_R^T on v_mfma_f64_16x16x4)"
                      if v == 1 else f"k_kron_tile ({KRON_VARIANTS[v]})",
            "achieved": round(achieved, 2),
            "peak": PEAK_FP64_MFMA_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_MFMA_TFLOPS, 4),
            "traffic": None,
            "alg_flops_per_launch": int(fl),
            "avg_launch_us": round(avg_s * 1e6, 2),
            "launches": int(launches.value),
            "launches_sampled": f"1 in {PROFILE_EVERY} of the timed solves' products, HIP events on the library "
                                "stream around the dense-product kernel",
        },
    }


def run_superblock(D, L, name, steps, warmup, fixed, cpu_budget):
    """Solves/s of one superblock workload: device assembly + Lanczos + inverse iteration/CG +
    vector to host per solve, p50/p99, the live SpMV roofline (HIP events on the library
    stream around one in PROFILE_EVERY of the products that always run in full), the
    FP32-refined form and, with cpu_budget, the CPU baseline."""
    from dmrg_c_amd import snapshot
    path, desc, cpu_form = SUPERBLOCKS[name]
    snap = snapshot.load(path)
    e_ref = float(snap.meta["gs_energy"][0])
    params = dict(diag_acc=1e-14, diag_max_step=800, cg_acc=1e-7, cg_max_step=2000, ii_acc=1e-9,
                  ii_diag_add=1e-11, ii_max_step=3, seed=1)
    sb = D.Superblock(snap)  # inputs resident in HBM from here on
    log(f"workload {name}: n={snap.n}")
    if fixed:
        lanes, tuned_us = fixed, float("nan")
        log(f"schedule given: (lanes per row, workgroup cap, load variant) = {lanes}")
    else:
        lanes, tuned_us = autotune(D, sb)
        log(f"autotune: (lanes per row, workgroup cap, load variant) = {lanes}, {tuned_us:.1f} us per Lanczos "
            "step + CG iteration")
    for i in range(warmup):
        E, st, nnz = gpu_solve(D, sb, params, lanes)
        log(f"warmup {i}: E={E:.15f} lanczos={st.lanczos_steps} cg={st.cg_steps}")
    # events bracket one SpMV launch in PROFILE_EVERY: every bracketed launch costs the
    # stream a few microseconds of idle, which would slow the timed solves themselves
    D.check(L.dmrgx_profile_begin_sampled(PROFILE_EVERY))
    lat, errs, lan, cg, gerr = [], [], [], [], []
    t0 = time.perf_counter()
    for i in range(steps):
        ts = time.perf_counter()
        E, st, nnz = gpu_solve(D, sb, params, lanes)
        lat.append(time.perf_counter() - ts)
        errs.append(abs(E - e_ref) / abs(e_ref))
        lan.append(st.lanczos_steps)
        cg.append(st.cg_steps)
        gerr.append(st.gs_error)
        log(f"step {i}: {lat[-1] * 1e3:.2f} ms  E={E:.15f}  dE/E={errs[-1]:.2e}")
    t1 = time.perf_counter()
    spmv_ms, launches, alg_bytes = C.c_double(), C.c_int64(), C.c_double()
    L.dmrgx_profile_end.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    D.check(L.dmrgx_profile_end(C.byref(spmv_ms), C.byref(launches), C.byref(alg_bytes)))
    total = t1 - t0
    # the same solves in the two mixed-precision forms
    def mixed_form(precision, note):
        mlat, merrs, mgerr, mlan, m32, mcg = [], [], [], [], [], []
        for i in range(max(2, steps // 4)):
            ts = time.perf_counter()
            E, st, _ = gpu_solve(D, sb, dict(params, precision=precision), lanes)
            mlat.append(time.perf_counter() - ts)
            merrs.append(abs(E - e_ref) / abs(e_ref))
            mgerr.append(st.gs_error)
            mlan.append(st.lanczos_steps)
            m32.append(st.lanczos_fp32_steps)
            mcg.append(st.cg_steps)
        log(f"precision {precision}: {np.mean(mlat) * 1e3:.1f} ms, lanczos {mlan[-1]} (FP32 {m32[-1]}), cg {mcg[-1]}")
        return {"solves_per_s": round(len(mlat) / sum(mlat), 4),
                "p99_ms": round(float(np.percentile(mlat, 99)) * 1e3, 3),
                "max_rel_energy_error": float(max(merrs)),
                "max_gs_error": float(max(mgerr)),
                "lanczos_steps": int(np.median(mlan)),
                "lanczos_fp32_steps": int(np.median(m32)),
                "cg_steps": int(np.median(mcg)),
                "over_fp64": round((len(mlat) / sum(mlat)) / (steps / total), 3),
                "precision": precision,
                "note": note}

    two_phase = mixed_form(2, "precision 2: Lanczos on FP32 matrix values until the Ritz residual bound falls to "
                              "1e-3 max(1,|E|), then FP64 Lanczos from that Ritz vector until the bound is below "
                              "ii_acc/(2 sqrt(n)), then the FP64 L1 residual test (FP64 inverse iteration/CG only "
                              "if it fails); FP64 recurrence scalars throughout")
    ii_form = mixed_form(1, "precision 1 (the batched engine's form): FP32 matrix values in the Lanczos SpMV, "
                            "FP64 recurrence scalars, FP64 inverse iteration/CG (inner solves end once the "
                            "iterate passes half the residual test) and Rayleigh quotient")
    avg_launch_s = spmv_ms.value / 1e3 / max(launches.value, 1)
    bytes_per_launch = alg_bytes.value / max(launches.value, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    traffic, traffic_note = None, None
    tfile = os.path.join(REPO, "profiles", f"spmv_{name}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
        if list(tj.get("schedule", [])) != list(lanes):
            traffic_note = (f"rocprof counters of the schedule {tj.get('schedule')} "
                            f"(this run's autotuner picked {list(lanes)})")
    fmt = (lanes[2] >> 3) & 7
    res = {
        "value": round(steps / total, 4),
        "ms_per_step": round(total / steps * 1e3, 3),
        "data": f"superblock snapshot from the reference's own DMRG run (tests/golden/{os.path.basename(path)})",
        "config": {
            "workload": f"{desc}, superblock n={snap.n}: device assembly + Lanczos + inverse iteration/CG per solve",
            "n": snap.n,
            "nnz_full": int(nnz),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "max_rel_energy_error": float(max(errs)),
            "max_gs_error": float(max(gerr)),
            "lanczos_steps": int(np.median(lan)),
            "cg_steps": int(np.median(cg)),
            "spmv_schedule": {"lanes_per_row": lanes[0], "workgroup_cap": lanes[1],
                              "paired_entry_loads": bool(lanes[2] & 1) and fmt == 0,
                              "entries_in_flight_per_lane": 8 if (lanes[2] & 2 or fmt >= 3) else 4,
                              "non_temporal_matrix_loads": bool(lanes[2] & 2) and fmt >= 3,
                              "matrix_loads_one_group_ahead": bool(lanes[2] & 64),
                              "column_bits": 16 if ((fmt == 3 and lanes[2] & 1) or fmt == 4) else 32,
                              "xcd_contiguous_rows": bool(lanes[2] & 4),
                              "format": SPMV_FORMATS[fmt][0]},
        },
        "fp32_refined": two_phase,
        "fp32_refined_inverse_iteration": ii_form,
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name(lanes),
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            **({"traffic_note": traffic_note} if traffic_note else {}),
            "alg_bytes_per_launch": int(bytes_per_launch),
            "avg_launch_us": round(avg_launch_s * 1e6, 2),
            "launches": int(launches.value),
            "launches_sampled": f"1 in {PROFILE_EVERY} of the timed solves' SpMV launches that always run in full "
                                "(Lanczos products, start residuals; CG products can return at once past "
                                "convergence and are left out), HIP events on the library stream",
        },
        "cpu_baseline": None,
    }
    if cpu_budget is not None:
        if cpu_form == "products":
            res["cpu_baseline"] = cpu_baseline_products(sb, snap.n, int(np.median(lan)) + int(np.median(cg)),
                                                        cpu_budget)
        else:
            res["cpu_baseline"] = cpu_baseline(snap, e_ref, cpu_budget)
    sb.free()
    return res


def batched_form(D):
    """The batched form on the 256-problem ED mix: the plain FP64 path one solve at a time
    (dmrgx_ground_state per problem) against the FP32-refined solves batched through the engine
    (segmented launches for all active problems, hipGraph replays of repeated chunks)."""
    r = run_ed_batch(D, 256, precision_forms=(0, 1), single=True)
    f64, f32 = r["fp64"], r["fp32_refined"]
    return {"workload": f"256 seeded ED ground states (n {r['n_min']}..{r['n_max']}): XXZ L=12..24 S^z=0 "
                        "J_z~U[0.5,1.5] h~U[0,0.5]; Hubbard L=6..10 half filling t=1 U=4 h~U[0,0.5]",
            "plain_fp64_one_by_one_solves_per_s": f64["single_solves_per_s"],
            "batched_fp64_solves_per_s": f64["batched_solves_per_s"],
            "batched_fp32_refined_solves_per_s": f32["batched_solves_per_s"],
            "batched_fp32_refined_over_plain_fp64": round(f32["batched_solves_per_s"] / f64["single_solves_per_s"], 3),
            "max_rel_energy_error_vs_fp64": f32["max_rel_energy_diff_vs_fp64"],
            "batched_equals_single_fp64": f64["batched_equals_single"],
            "graphs": f32["engine_host_s"].get("graphs"),
            "detail": r}


def served_form(D, duration):
    """The served stream at 200 requests/s: the mix as configured (superblocks of the
    reference's runs at m=300..800, the KLM m=800 one with n ~ 2e6 included) and, beside it,
    the same stream without the m=800 superblock, which one GPU can keep up with."""
    keep = ("requests", "rate_per_s", "duration_s", "ed_requests", "superblock_requests", "solves_per_s",
            "p50_ms", "p99_ms", "max_rel_energy_error_superblocks", "superblocks", "precision",
            "engine_chunks", "host_s", "latency_split_ms")
    out = {}
    for tag, names, dur in (("c5_mix", SERVE_SNAPSHOTS, min(duration, 1.0)),
                            ("without_m800", SERVE_SNAPSHOTS[:2], duration)):
        r = run_serve(D, 200.0, dur, snapshots=names, precision=0)
        out[tag] = {k: r[k] for k in keep if k in r}
    out["c5_mix"]["note"] = ("the configured mix: one third of the superblock requests are the KLM m=800 "
                             "superblock (n=1993680, ~1.4 s of GPU time each), ~3 GPU-seconds of work per "
                             "second at 200 requests/s, so one GPU falls behind and latency grows with the "
                             "arrival window (1 s here)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run 5 plain SpMV launches on the workload and one copy kernel of 2^26 doubles "
                         "(traffic calibration under rocprofv3 --pmc), then exit")
    ap.add_argument("--kron-probe", type=int, default=-1, metavar="VARIANT",
                    help="build the workload's Kronecker form, apply product-kernel VARIANT to a seeded vector "
                         "5 times (for rocprofv3 --pmc), then exit")
    ap.add_argument("--schedule", default="",
                    help="SpMV schedule 'lanes,cap,variant' instead of autotuning (bench and --traffic-probe)")
    ap.add_argument("--spmv-sweep", action="store_true",
                    help="time every SpMV schedule on the workload superblock (JSON table), then exit")
    ap.add_argument("--workload", choices=["superblock", "ed-batch", "serve", "basis", "v-m", "density"], default="superblock",
                    help="superblock: the headline line; ed-batch: 256 seeded ED solves one by one and "
                         "through the batched engine; serve: Poisson request stream through the engine; "
                         "basis: superblock basis from quantum numbers; v-m: sector-changing "
                         "single-slot product; density: reduced density matrices of the left half "
                         "(each a separate JSON line)")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=10.0)
    ap.add_argument("--chunk-ticks", type=int, default=1,
                    help="serve: engine steps per launched chunk (1: lowest latency of 1/2/4/8/16 measured)")
    ap.add_argument("--ed-count", type=int, default=256)
    ap.add_argument("--ed-batched-only", action="store_true",
                    help="ed-batch: skip the one-by-one solves (for profiling the engine alone)")
    ap.add_argument("--superblock", choices=sorted(SUPERBLOCKS), default="",
                    help="headline superblock (default: KLM L=48 m=800 when its snapshot is present, "
                         "else Hubbard L=64 m=400)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline superblock (no secondary superblock, batched or served forms)")
    ap.add_argument("--serve-duration", type=float, default=2.0,
                    help="served form of the headline line: seconds of Poisson arrivals at 200/s")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the superblock eigensolve runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("DMRGX_SML_OUT", "0")
    cpu_env()   # OpenMP placement of the CPU baseline, before any OpenMP library loads
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot

    L = D.lib()
    D.check(L.dmrgx_init(0))
    if args.workload == "ed-batch":
        print(json.dumps({"metric": "batched ED ground-state solves/s", "workload": "ed-batch",
                          "result": run_ed_batch(D, args.ed_count, single=not args.ed_batched_only)}), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps({"metric": "served solves/s and p50/p99 latency", "workload": "serve",
                          "result": run_serve(D, args.rate, args.duration, chunk_ticks=args.chunk_ticks)}), flush=True)
        return
    if args.workload == "basis":
        print(json.dumps({"metric": "superblock basis builds/s", "workload": "basis",
                          "result": run_basis(D, args.steps, not args.no_cpu_baseline)}), flush=True)
        return
    if args.workload == "density":
        print(json.dumps({"metric": "reduced density matrix sets/s", "workload": "density",
                          "result": run_density(D, args.steps, not args.no_cpu_baseline)}), flush=True)
        return
    if args.workload == "v-m":
        print(json.dumps({"metric": "sector-changing single-slot products/s", "workload": "v-m",
                          "result": run_v_m(D, args.steps, not args.no_cpu_baseline)}), flush=True)
        return
    wl = args.superblock if args.superblock else ("klm_L48_m800" if os.path.exists(SUPERBLOCKS["klm_L48_m800"][0])
                                                   else "hubbard_L64_m400")
    path = SUPERBLOCKS[wl][0]
    fixed = tuple(int(v) for v in args.schedule.split(",")) if args.schedule else None
    if args.kron_probe >= 0:
        snap = snapshot.load(path)
        sb = D.Superblock(snap)
        K = sb.kron()
        K.set_kernel(args.kron_probe)
        x = np.random.default_rng(1).standard_normal(K.n)
        for _ in range(5):
            y = K.matvec(x)
        print(json.dumps({"kron_products": 5, "variant": args.kron_probe, "n": K.n, "flops_per_product": K.flops,
                          "y0": float(y[0])}), flush=True)
        K.free()
        sb.free()
        return
    if args.traffic_probe or args.spmv_sweep:
        snap = snapshot.load(path)
        sb = D.Superblock(snap)
        if args.traffic_probe:
            traffic_probe(D, L, sb, snap, fixed)
        else:
            spmv_sweep(D, sb, snap)
        return
    kron = run_kron(D, L, wl, args.steps, args.warmup)
    head = run_superblock(D, L, wl, max(2, args.steps // 2) if not args.traffic_probe else args.steps, 1, fixed,
                          None if args.no_cpu_baseline else args.cpu_budget)
    crs = {k: head[k] for k in ("value", "ms_per_step", "config", "roofline")}
    crs["note"] = ("the reference's own structure: device assembly of the CRS (bit-exact with GET_HAM_LLLRRRRL), "
                   "lower -> full expansion, sliced-form build and the CRS SpMV per product; the same Lanczos + "
                   "inverse iteration/CG")
    crs["over"] = "kron value / crs value = " + str(round(kron["value"] / head["value"], 2))
    crs["spmv_schedule_note"] = "the SpMV roofline of this form is spmv_roofline"
    line = {
        "metric": "ground-state solves/s + p99 latency at |dE/E|<=1e-9; SpMV HBM GB/s vs peak",
        "value": kron["value"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kron["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": kron["data"],
        "config": kron["config"],
        "forms": {"crs_fp64": crs,
                  "kron_residual_stop": dict(run_kron(D, L, wl, max(2, args.steps // 2), 1, lan_type=2, profile=False),
                                             note="the Kronecker form with lan_type 2 (no reference counterpart): "
                                                  "Lanczos until the Ritz pair's residual bound is below "
                                                  "ii_acc/(2 sqrt(n)), then the reference's L1 residual test (CG only "
                                                  "if it fails); FP64"),
                  "fp32_refined": dict(head["fp32_refined"], over="CRS FP64 path"),
                  "fp32_refined_inverse_iteration": dict(head["fp32_refined_inverse_iteration"],
                                                         over="CRS FP64 path")},
        "roofline": kron["roofline"],
        "spmv_roofline": head["roofline"],
        "cpu_baseline": head["cpu_baseline"],
    }
    if not args.headline_only:
        # the other superblock as a secondary measurement (same solve, its own roofline)
        for other in SUPERBLOCKS:
            if other != wl and os.path.exists(SUPERBLOCKS[other][0]):
                ksec = run_kron(D, L, other, max(4, args.steps // 2), 1)
                sec = run_superblock(D, L, other, max(4, args.steps // 2), 1, None, None)
                line.setdefault("secondary_workloads", {})[other] = {
                    "kron": {k: ksec[k] for k in ("value", "ms_per_step", "config", "roofline")},
                    "crs_fp64": {k: sec[k] for k in ("value", "ms_per_step", "config", "fp32_refined",
                                                     "fp32_refined_inverse_iteration", "roofline")}}
        log("forms: batched ED (256 seeded problems)")
        line["forms"]["batched"] = batched_form(D)
        log("forms: served stream")
        line["forms"]["served"] = served_form(D, args.serve_duration)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
