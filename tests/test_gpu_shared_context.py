"""Several host threads on one context (bind_context): the device-memory cache
is shared under a mutex and the work of all threads is ordered by the
context's one stream, so concurrent contractions, plans and eigensolves give
exactly what they give on one thread; a slot's context of the serving loop
holds no cached memory after a run."""
import gc
import threading

import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from sparsebackends_amd import krylov as K

pytestmark = pytest.mark.gpu


def solve(seed):
    prob = H.HeffProblem(H.j1j2_model(1.0, 0.25), H.heisenberg_bond(48, qmax=3, width=6.0),
                         H.heisenberg_bond(48, qmax=3, width=6.0), seed=seed, hermitian=True)
    prob.upload().outputs()
    plans = prob.make_plans(0)
    prob.execute()
    hp = prob.Hpsi.parts()["data"].copy()
    vals, _, info = K.eigsolve(plans, prob.psi, krylovdim=8, maxiter=2, tol=1e-10)
    sb.synchronize()
    return hp, vals[0], info.numops


def test_threads_sharing_one_context():
    print("[shared context]", flush=True)
    ref = {s: solve(s) for s in range(6)}
    shared = sb.context(0)
    got, errs = {}, []

    def work(seeds):
        try:
            sb.bind_context(shared, 0)
            for s in seeds:
                got[s] = solve(s)
        except Exception as e:  # reported below
            errs.append(e)
        finally:
            sb.bind_context(None, 0)

    ths = [threading.Thread(target=work, args=([s, s + 3],)) for s in range(3)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    sb.set_graphs(-1)  # binding turned the shared context's eigensolve graphs off
    assert not errs, errs
    for s in range(6):
        assert np.array_equal(got[s][0], ref[s][0])
        assert got[s][1] == ref[s][1] and got[s][2] == ref[s][2]


def test_serving_slots_release_cache():
    from sparsebackends_amd import serve as S
    print("[serve release]", flush=True)
    seen = []
    lock = threading.Lock()

    def release(dev):
        # every slot thread has a context of its own (made at its first call):
        # after its solves, collect its garbage and trim, with the collector
        # kept off while the figures are read
        with lock:
            gc.collect()
            gc.disable()
            try:
                before = sb.memory(dev)
                sb.trim(dev)
                seen.append((before, sb.memory(dev)))
            finally:
                gc.enable()

    srv = S.DMRGServer(slots=2, co_batch=False)
    srv.release = release
    srv.run(S.poisson_requests(3, float("inf"), seed=1, sites=(8,), chis=(16,), max_sweeps=3))
    assert len(seen) == 2, seen  # each slot released its context once
    for before, after in seen:
        # the solves' tensors are gone: nothing in use, the cache held blocks and the trim returned them all
        assert before[0] == 0 and after[0] == 0, seen
        assert before[1] > 0 and after[1] == 0, seen
