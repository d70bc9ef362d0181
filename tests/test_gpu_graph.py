"""Plan chains replayed as HIP graphs (sb_graph_*): every replay equals the
eager execution of the same chain on the same inputs, bit for bit, with the
program's own eager work and synchronisations between the replays and new
operand values before each; a changed host-side decision (an operand whose
payload pointer was handed out, so its boxes are rescanned) is still handled
correctly; an eigensolve replaying its operator as a graph gives exactly the
eager eigensolve's result."""
import copy

import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from sparsebackends_amd import krylov as K

pytestmark = pytest.mark.gpu


def pair(seed=1):
    prob = H.heisenberg_chi256_problem(seed=seed)
    prob.upload().outputs()
    prob.make_plans(0)
    other = copy.copy(prob).upload().outputs()
    other.make_plans(0)
    return prob, other


def test_graph_replays_equal_eager():
    print("[graph] replays", flush=True)
    eager, graphed = pair()
    g = sb.PlanGraph(graphed.plans)
    rng = np.random.default_rng(3)
    base = eager.psi.parts()["data"]
    mask = base != 0
    for r in range(7):
        vals = np.where(mask, rng.standard_normal(base.shape), 0.0)
        eager.psi.set_data_(vals)
        graphed.psi.set_data_(vals)
        eager.execute()
        g.launch()
        sb.synchronize()
        a, b = eager.Hpsi.parts(), graphed.Hpsi.parts()
        assert np.array_equal(a["keys"], b["keys"]) and np.array_equal(a["ids"], b["ids"])
        assert np.array_equal(a["data"], b["data"]), r
        # the program's own eager work between replays
        n = graphed.psi.norm()
        assert n > 0
        sb.synchronize()
    st = g.stats()
    assert st["eager"] == 1 and st["captures"] == 1 and st["replays"] == 5, st
    assert st["nodes"] > 0


def test_graph_changed_decision():
    print("[graph] changed decision", flush=True)
    eager, graphed = pair(seed=2)
    g = sb.PlanGraph(graphed.plans)
    for _ in range(3):
        g.launch()
    sb.synchronize()
    # handing out the payload pointer marks L as written from outside: its
    # boxes are rescanned at every use, a decision the capture did not make
    graphed.L.device_ptrs()
    eager.L.device_ptrs()
    for _ in range(3):
        eager.execute()
        g.launch()
        sb.synchronize()
        assert np.array_equal(eager.Hpsi.parts()["data"], graphed.Hpsi.parts()["data"])
    st = g.stats()
    assert st["captures"] >= 1 and st["eager"] + st["captures"] + st["replays"] == 6, st


@pytest.mark.parametrize("kdim,maxiter", [(8, 3), (20, 2)])
def test_eigsolve_graph_equals_eager(kdim, maxiter):
    print(f"[graph] eigsolve {kdim} {maxiter}", flush=True)
    res = []
    for mode in (0, 1):
        sb.set_graphs(mode)
        prob = H.heisenberg_chi256_problem(seed=4, hermitian=True)
        prob.upload().outputs()
        plans = prob.make_plans(0)
        vals, _, info = K.eigsolve(plans, prob.psi, krylovdim=kdim, maxiter=maxiter, tol=1e-10)
        res.append((vals[0], info.numops, info.normres, prob.psi.parts()["data"]))
    sb.set_graphs(-1)
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1] and res[0][2] == res[1][2]
    assert np.array_equal(res[0][3], res[1][3])
