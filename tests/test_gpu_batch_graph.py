"""Batches as HIP graphs and with measured tile shapes: one graph launch per
execution of every chain of a PlanBatch equals the eager batch execution bit
for bit, replay after replay, with new psi values and eager work between the
replays; a batch retuned by timing every workgroup tile (sb_batch_autotune)
computes the same values; the batched eigensolve replaying its operator as a
graph equals the eager batched eigensolve exactly."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import krylov as K
from test_gpu_batch import built, outputs

pytestmark = pytest.mark.gpu


def test_batch_graph_replays():
    print("[batch graph] replays", flush=True)
    eager, graphed = built(), built()
    be = sb.PlanBatch([p.plans for p in eager])
    bg = sb.PlanBatch([p.plans for p in graphed])
    g = sb.PlanGraph(bg)
    rng = np.random.default_rng(11)
    for r in range(6):
        for pe, pg in zip(eager, graphed):
            base = pe.psi.parts()["data"]
            vals = np.where(base != 0, rng.standard_normal(base.shape), 0.0)
            pe.psi.set_data_(vals)
            pg.psi.set_data_(vals)
        be.execute()
        g.launch()
        sb.synchronize()
        for pe, pg in zip(eager, graphed):
            for a, b in zip(outputs(pe), outputs(pg)):
                assert np.array_equal(a, b), r
        assert all(p.psi.norm() > 0 for p in graphed)  # eager work between replays
        sb.synchronize()
    st = g.stats()
    assert st["eager"] == 1 and st["captures"] == 1 and st["replays"] == 4, st


def test_batch_graph_after_partial_execute():
    """an eager execution of some chains in between: the next launch runs
    eagerly (the uploaded active lists differ from the capture's), later ones
    replay again"""
    print("[batch graph] partial execute", flush=True)
    eager, graphed = built(), built()
    be = sb.PlanBatch([p.plans for p in eager])
    bg = sb.PlanBatch([p.plans for p in graphed])
    g = sb.PlanGraph(bg)
    for _ in range(3):
        g.launch()
    bg.execute([0, 2])
    be.execute()
    for _ in range(3):
        g.launch()
    be.execute()
    sb.synchronize()
    for pe, pg in zip(eager, graphed):
        for a, b in zip(outputs(pe), outputs(pg)):
            assert np.array_equal(a, b)
    st = g.stats()
    assert st["eager"] >= 2 and st["replays"] >= 3, st


def test_batch_autotune_same_values():
    print("[batch autotune]", flush=True)
    ps = built()
    batch = sb.PlanBatch([p.plans for p in ps])
    batch.execute()
    ref = [outputs(p) for p in ps]
    before = batch.stage_tiles()
    batch.autotune()
    after = batch.stage_tiles()
    assert all(0 <= t < 8 and ms > 0 for t, ms in after), (before, after)
    batch.execute()
    for p, r in zip(ps, ref):
        for a, b in zip(outputs(p), r):
            assert np.array_equal(a, b)


def test_eigsolve_batch_graph_equals_eager():
    print("[batch graph] eigsolve", flush=True)
    from sparsebackends_amd import heff as H
    res = []
    for mode in (0, 1):
        sb.set_graphs(mode)
        ps = [H.HeffProblem(H.j1j2_model(1.0, 0.2), H.heisenberg_bond(40, qmax=3, width=6.0),
                            H.heisenberg_bond(40, qmax=3, width=6.0), seed=4, hermitian=True),
              H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(64, n0=0, sigma=1.8),
                            H.hubbard_bond(64, n0=2, sigma=1.8), seed=5, hermitian=True),
              H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(96, qmax=4, width=8.0),
                            H.heisenberg_bond(96, qmax=4, width=8.0), seed=6, hermitian=True)]
        for p in ps:
            p.upload().outputs().make_plans(0)
        batch = sb.PlanBatch([p.plans for p in ps])
        vals, infos = K.eigsolve_batch(batch, [p.psi for p in ps], krylovdim=8, maxiter=3, tol=1e-10)
        res.append((vals, [i.numops for i in infos], [p.psi.parts()["data"] for p in ps]))
    sb.set_graphs(-1)
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
    for a, b in zip(res[0][2], res[1][2]):
        assert np.array_equal(a, b)
