"""Batched and continuously batched solves (dmrgx_engine_*, dmrgx_ground_state_batch).

Every request keeps the blocks, lanes per row and summation order of dmrgx_ground_state,
so the bar here is bit-identical results: alone, in a batch, and admitted mid-run."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_snapshots

SNAPS = [p for p in golden_snapshots() if p.endswith(("_s03.npz", "_s05.npz", "_s07.npz", "_s13.npz"))]


def _superblock_matrices():
    import dmrg_c_amd as D
    from dmrg_c_amd import snapshot
    mats, refs = [], []
    for p in SNAPS:
        s = snapshot.load(p)
        mats.append(D.DeviceMatrix(s.ham.row, s.ham.col, s.ham.val, sym=True))
        refs.append(float(s.meta["gs_energy"][0]))
    return mats, refs


def _ed_matrices():
    from dmrg_c_amd import ed
    out = []
    for m in (ed.Model("xxz", L=12, sz=0, J_z=0.7, h_z=0.2), ed.Model("hubbard", L=6, N=6, sz=0, U=3.0),
              ed.Model("xxz", L=14, sz=2, J_z=1.3), ed.Model("hubbard", L=7, N=7, sz=1, U=4.0, bc="PBC")):
        P = ed.EdProblem.from_model(m)
        out.append(P.assemble())
        P.free()
    return out


def _single(A, **kw):
    return A.ground_state(**kw)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32_refined"])
def test_gpu_batch_bit_identical_to_single(precision):
    from dmrg_c_amd import engine
    mats, refs = _superblock_matrices()
    mats += _ed_matrices()
    singles = [_single(A, precision=precision) for A in mats]
    batch = engine.ground_state_batch(mats, engine.default_params(precision=precision))
    for (e1, v1, s1), (e2, v2, s2) in zip(singles, batch):
        assert e1 == e2
        assert np.array_equal(v1, v2)
        assert s1.lanczos_steps == s2.lanczos_steps and s1.cg_steps == s2.cg_steps
        assert s1.ii_steps == s2.ii_steps
    for (e, _, _), r in zip(batch, refs):
        assert abs(e - r) <= (1e-10 if precision == 0 else 1e-9) * abs(r)
    for A in mats:
        A.free()


@pytest.mark.gpu
def test_gpu_engine_continuous_admission_same_as_alone():
    """Requests admitted while others are mid-solve get the same bits as alone."""
    from dmrg_c_amd import engine
    mats = _ed_matrices()
    smats, _ = _superblock_matrices()
    mats += smats
    alone = {}
    for i, A in enumerate(mats):
        E = engine.Engine()
        E.submit(A, tag=i)
        (r,) = E.run_until_done()
        alone[i] = r
        E.free()
    E = engine.Engine(chunk_ticks=4)
    got = {}
    for i, A in enumerate(mats):
        E.submit(A, tag=i)
        E.run(1)                      # the earlier ones advance before the next arrives
        for r in E.poll():
            got[r.tag] = r
    for r in E.run_until_done():
        got[r.tag] = r
    E.free()
    assert sorted(got) == sorted(alone)
    for i in alone:
        assert got[i].energy == alone[i].energy
        assert np.array_equal(got[i].vec, alone[i].vec)
        assert got[i].status == 0
    for A in mats:
        A.free()


@pytest.mark.gpu
def test_gpu_engine_hbm_budget_serialises():
    """A budget that fits one request at a time still completes every request correctly."""
    from dmrg_c_amd import engine
    mats = _ed_matrices()
    E = engine.Engine(hbm_budget_bytes=1, max_active=64)
    for i, A in enumerate(mats):
        E.submit(A, tag=i)
    res = E.run_until_done()
    E.free()
    assert len(res) == len(mats)
    # one at a time: each starts after the previous finished
    res.sort(key=lambda r: r.t_start)
    for a, b in zip(res, res[1:]):
        assert b.t_start >= a.t_done
    for r, A in zip(sorted(res, key=lambda r: r.tag), mats):
        e, v, _ = A.ground_state()
        assert r.energy == e
    for A in mats:
        A.free()


@pytest.mark.gpu
def test_gpu_engine_owned_matrices_and_many_small():
    """64 small ED requests owned by the engine (freed with their results)."""
    from dmrg_c_amd import ed, engine
    rng = np.random.default_rng(0)
    E = engine.Engine()
    expect = {}
    for i in range(64):
        m = ed.Model("xxz", L=int(rng.choice([8, 10, 12])), sz=0, J_z=float(rng.uniform(0.5, 1.5)),
                     h_z=float(rng.uniform(0, 0.5)))
        P = ed.EdProblem.from_model(m)
        A = P.assemble()
        P.free()
        expect[i] = A.ground_state()[0]
        E.submit(A, tag=i, own=True)
    res = E.run_until_done(want_vec=False)
    E.free()
    assert len(res) == 64
    for r in res:
        assert r.energy == expect[r.tag]


@pytest.mark.gpu
def test_gpu_engine_graph_replays_bit_identical_to_eager():
    """Chunks whose launch sequence repeats run as a replayed hipGraph; results match the
    eager engine bit for bit, over many replays, and over two engines in a row."""
    from dmrg_c_amd import ed, engine
    mats = []
    for i in range(6):
        P = ed.EdProblem.from_model(ed.Model("xxz", L=14, sz=0, J_z=0.6 + 0.1 * i, h_z=0.05 * i))
        mats.append(P.assemble())
        P.free()
    mats += _ed_matrices()

    def run(use_graphs):
        E = engine.Engine(chunk_ticks=4, use_graphs=use_graphs)
        for i, A in enumerate(mats):
            E.submit(A, tag=i)
        res = {r.tag: r for r in E.run_until_done()}
        st = E.graph_stats()
        E.free()
        return res, st

    eager, st0 = run(False)
    assert st0["captures"] == 0 and st0["replays"] == 0 and st0["eager"] > 0
    for _ in range(2):
        got, st1 = run(True)
        assert st1["captures"] >= 1 and st1["replays"] >= 1, st1
        assert sorted(got) == sorted(eager)
        for i in eager:
            assert got[i].status == 0
            assert got[i].energy == eager[i].energy
            assert np.array_equal(got[i].vec, eager[i].vec)
            assert got[i].lanczos_steps == eager[i].lanczos_steps
            assert got[i].cg_steps == eager[i].cg_steps
    for A in mats:
        A.free()
