"""The serving loop's schedule on the CPU: request streams, footprint
estimates, admission by slots and memory budget, first-come-first-served
order, leader-based co-batching of eigensolves, error propagation.  The DMRG
solve and the eigensolvers are replaced by deterministic stand-ins that
sleep like a solve and record what was batched with what."""
import threading
import time

import numpy as np
import pytest

from sparsebackends_amd import serve as S


class FakeVec:
    device = 0

    def __init__(self, rid):
        self.rid = rid


def fake_value(plans):
    _, rid, sweep, b = plans
    return -(rid + 1e-3 * sweep + 1e-6 * b)


class Recorder:
    def __init__(self):
        self.lock = threading.Lock()
        self.batches = []
        self.solo = 0

    def solo_fn(self, plans, x, krylovdim=3, maxiter=1, tol=1e-14):
        with self.lock:
            self.solo += 1
        return [fake_value(plans)], [x], None

    def multi_fn(self, chains, xs, krylovdim=3, maxiter=1, tol=1e-14):
        rids = [x.rid for x in xs]
        with self.lock:
            self.batches.append(rids)
        time.sleep(0.0005)
        return [fake_value(p) for p in chains], [None] * len(chains)


def fake_solver(req, eig, device):
    es = []
    for sweep in range(3):
        e = None
        for b in range(max(1, req.sites // 8)):
            time.sleep(0.0002 * (1 + req.chi_max // 32))
            vals, _, _ = eig(("plans", req.rid, sweep, b), FakeVec(req.rid), krylovdim=3, maxiter=1, tol=1e-14)
            e = vals[0]
        es.append(e)
    return es[-1], es, True, req.chi_max, 0, 0.0


def reqs(n=12, rate=400.0, seed=0):
    return S.poisson_requests(n, rate, seed=seed, sites=(16, 24, 32), chis=(32, 64))


def test_poisson_requests():
    a, b = reqs(20, 5.0, 3), reqs(20, 5.0, 3)
    assert [(r.sites, r.chi_max, r.delta, r.arrival) for r in a] == [(r.sites, r.chi_max, r.delta, r.arrival)
                                                                      for r in b]
    t = np.array([r.arrival for r in a])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    assert all(r.sites in (16, 24, 32) and r.chi_max in (32, 64) and 0.5 <= r.delta <= 1.5 for r in a)
    assert all(r.arrival == 0.0 for r in reqs(5, float("inf")))
    # mean gap near 1/rate
    big = S.poisson_requests(4000, 10.0, seed=1)
    assert abs(np.mean(np.diff([r.arrival for r in big])) - 0.1) < 0.01


def test_footprint_grows_with_size():
    r = S.SolveRequest(0, 32, 128, 1.0)
    assert S.footprint_bytes(r) < S.footprint_bytes(S.SolveRequest(0, 64, 128, 1.0))
    assert S.footprint_bytes(r) < S.footprint_bytes(S.SolveRequest(0, 32, 256, 1.0))
    # at least the dense size of the sites' tensors
    assert S.footprint_bytes(r) >= 8 * 32 * 2 * 128 * 128


@pytest.mark.parametrize("co_batch", [False, True])
def test_server_results_and_bounds(co_batch):
    rq = reqs()
    rec = Recorder()
    plain = S.run_plain(rq, solver=fake_solver, eigsolver=rec.solo_fn)
    srv = S.DMRGServer(slots=4, co_batch=co_batch, window=0.002, solver=fake_solver,
                       batcher_args=dict(solo=rec.solo_fn, multi=rec.multi_fn, sync=lambda x: None))
    res = srv.run(rq)
    assert [r.rid for r in res] == list(range(len(rq)))
    # the same result alone and in the mix
    assert [r.energy for r in res] == [r.energy for r in plain]
    assert [r.energies for r in res] == [r.energies for r in plain]
    assert srv.max_inflight <= 4
    for r in res:
        assert r.admitted >= r.arrival - 1e-9 and r.finish >= r.admitted and r.latency > 0
    if co_batch:
        calls = sum(max(1, r.sites // 8) * 3 for r in rq)
        assert sum(srv.batch_sizes) == calls
        # a batch holds at most one job per slot, never two of one request
        assert all(len(set(b)) == len(b) and len(b) <= 4 for b in rec.batches)
        assert any(len(b) > 1 for b in rec.batches)


def test_fifo_admission_and_memory_budget():
    rq = S.poisson_requests(10, float("inf"), seed=2, sites=(24,), chis=(64,))
    fp = S.footprint_bytes(rq[0])
    srv = S.DMRGServer(slots=8, hbm_budget=2 * fp, co_batch=False, solver=fake_solver,
                       batcher_args=dict(solo=Recorder().solo_fn))
    res = srv.run(rq)
    assert srv.max_reserved <= 2 * fp
    assert srv.max_inflight <= 2
    # all arrive at once: admitted in request order
    adm = [r.admitted for r in sorted(res, key=lambda r: r.rid)]
    assert all(adm[i] <= adm[i + 1] for i in range(len(adm) - 1))


def test_duplicate_ids_rejected():
    rq = [S.SolveRequest(1, 16, 32, 1.0), S.SolveRequest(1, 16, 32, 0.5)]
    with pytest.raises(ValueError):
        S.DMRGServer(slots=2, solver=fake_solver).run(rq)


def test_results_in_given_order():
    rq = [S.SolveRequest(i, 16, 32, 1.0) for i in (5, 2, 9, 0)]
    solo = Recorder().solo_fn
    for res in (S.DMRGServer(slots=2, co_batch=False, solver=fake_solver, batcher_args={"solo": solo}).run(rq),
                S.run_plain(rq, solver=fake_solver, eigsolver=Recorder().solo_fn)):
        assert [r.rid for r in res] == [5, 2, 9, 0]
    with pytest.raises(ValueError):
        S.run_plain(rq + rq[:1], solver=fake_solver, eigsolver=Recorder().solo_fn)


def test_oversized_request_rejected():
    rq = [S.SolveRequest(0, 128, 1024, 1.0)]
    with pytest.raises(ValueError):
        S.DMRGServer(slots=2, hbm_budget=1 << 20, solver=fake_solver).run(rq)


def test_solver_error_propagates():
    def bad(req, eig, device):
        raise RuntimeError("solve failed")
    with pytest.raises(RuntimeError):
        S.DMRGServer(slots=2, co_batch=False, solver=bad).run(reqs(3, float("inf")))


def test_batch_error_reaches_every_slot():
    rec = Recorder()

    def multi_bad(chains, xs, **kw):
        raise RuntimeError("batch failed")
    srv = S.DMRGServer(slots=3, co_batch=True, window=0.05, solver=fake_solver,
                       batcher_args=dict(solo=rec.solo_fn, multi=multi_bad, sync=lambda x: None))
    with pytest.raises(RuntimeError):
        srv.run(reqs(6, float("inf")))


def test_summarize():
    res = [S.SolveResult(i, -1.0, 3, True, [], arrival=0.1 * i, admitted=0.1 * i, finish=0.1 * i + 1.0)
           for i in range(10)]
    s = S.summarize(res)
    assert s["solves"] == 10 and abs(s["span_s"] - 1.9) < 1e-12
    assert abs(s["p50_s"] - 1.0) < 1e-12 and abs(s["p99_s"] - 1.0) < 1e-12
