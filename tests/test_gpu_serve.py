"""The serving loop on the device: eigensolves of plan chains owned by other
threads' contexts batched on one stream (sb_eigsolve_lanczos_multi) equal the
solo eigensolve bit for bit; served DMRG solves (slots on their own contexts,
with and without co-batched Krylov steps) give each request the energy it
gets alone, bit for bit; energies match exact diagonalisation."""
import threading

import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import dmrg as D
from sparsebackends_amd import heff as H
from sparsebackends_amd import krylov as K
from sparsebackends_amd import serve as S

pytestmark = pytest.mark.gpu


def heff_at(n, delta, b, chi):
    """a DMRG state of an n-site XXZ chain after two sweeps, moved to bond b:
    (psi, H_eff plans) ready for an eigensolve"""
    dm = D.DMRG(H.heisenberg_model(1.0, delta), S.neel(n))
    for _ in range(2):
        dm.sweep(maxdim=chi, cutoff=1e-12, krylovdim=3, maxiter=1, tol=1e-14)
    for bb in range(1, b):
        dm.step(bb, +1, chi, 1e-12, 3, 1, 1e-14)
    psi = D.two_site(dm.A[b - 1], dm.A[b])
    plans, _ = H.heff_plans(psi, dm.L[b - 1], dm.W, dm.W, dm.R[b + 2])
    return dm, psi, plans


@pytest.mark.parametrize("kdim,maxiter", [(3, 1), (8, 3)])
def test_eigsolve_multi_across_contexts(kdim, maxiter):
    print(f"[eigsolve_multi] kdim={kdim} maxiter={maxiter}", flush=True)
    cases = [(12, 0.7, 5, 24), (14, 1.2, 6, 32), (10, 1.0, 4, 16)]
    made = {}

    def build(i, n, dlt, b, chi):  # each problem on its own thread's context
        dm, psi, plans = heff_at(n, dlt, b, chi)
        ref = D.two_site(dm.A[b - 1], dm.A[b])  # the same start vector, solved alone on this context
        rplans, _ = H.heff_plans(ref, dm.L[b - 1], dm.W, dm.W, dm.R[b + 2])
        vals, _, info = K.eigsolve(rplans, ref, krylovdim=kdim, maxiter=maxiter, tol=1e-14)
        sb.synchronize()
        made[i] = (dm, psi, plans, ref, rplans, vals[0], info)

    ths = [threading.Thread(target=build, args=(i,) + c) for i, c in enumerate(cases)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert len(made) == len(cases)
    chains = [made[i][2] for i in range(len(cases))]
    xs = [made[i][1] for i in range(len(cases))]
    vals, infos = K.eigsolve_multi(chains, xs, krylovdim=kdim, maxiter=maxiter, tol=1e-14)
    for i in range(len(cases)):
        _, psi, _, ref, _, v0, info0 = made[i]
        assert vals[i] == v0, (i, vals[i], v0)
        assert infos[i].numops == info0.numops and infos[i].normres == info0.normres
        assert np.array_equal(psi.parts()["data"], ref.parts()["data"])


def test_eigsolve_multi_errors():
    _, psi, plans = heff_at(10, 1.0, 4, 16)
    with pytest.raises(ValueError):
        K.eigsolve_multi([plans, plans[:3]], [psi, psi], krylovdim=3, maxiter=1)  # uneven chains
    with pytest.raises(sb.SBError):
        K.eigsolve_multi([plans, plans], [psi, psi], krylovdim=3, maxiter=1)  # one vector twice
    # the contexts got their streams back: a solo solve still works
    vals, _, info = K.eigsolve(plans, psi, krylovdim=3, maxiter=1, tol=1e-14)
    assert np.isfinite(vals[0])


def requests():
    rq = S.poisson_requests(6, float("inf"), seed=4, sites=(8, 10, 12), chis=(16, 32), max_sweeps=8)
    # exact at chi = 32; converged eigensolves
    rq.append(S.SolveRequest(len(rq), 10, 32, 0.8, 0.0, 1e-12, 10, cutoff=0.0, krylovdim=8, maxiter=3))
    return rq


def test_served_equals_alone():
    from chain_ed import ground_energy, spin_chain
    rq = requests()
    print("[serve] plain", flush=True)
    plain = S.run_plain(rq)
    for co_batch, window in [(False, 0.0), (True, 0.003)]:
        print(f"[serve] co_batch={co_batch}", flush=True)
        srv = S.DMRGServer(slots=4, co_batch=co_batch, window=window)
        res = srv.run(rq)
        for a, b in zip(plain, res):
            assert a.energy == b.energy and a.energies == b.energies, (a.rid, a.energy, b.energy)
        if co_batch:
            assert max(srv.batch_sizes) > 1
    # the admission estimate bounds what a solve holds
    assert all(0 < r.peak_bytes <= r.footprint for r in plain)
    # the N = 10, chi = 32 chain is exact: the ground energy of the Sz = 0 sector
    last = plain[-1]
    assert last.converged
    e_ed, _ = ground_energy(*spin_chain(10, [(1, 1.0, 0.8)]), (0,))
    assert abs(last.energy - e_ed) <= 1e-10
