"""GPU parity of the co-batched Lanczos (eigsolve_batch / sb_eigsolve_lanczos_batch):
several independent ground-state solves advanced in lock step, one batched
operator application per iteration for the solves still running.

Each chain's result must be bit for bit what eigsolve gives for it alone
(energy, residual, counters and Ritz vector) -- also when the solves finish
at different iterations, when their compact Krylov layouts widen, and in the
FP32 form; exact chains reach their exact-diagonalisation ground energy.
"""
import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from chain_ed import spin_chain, hubbard_chain, ground_energy

pytestmark = pytest.mark.gpu


def hermitian_problems():
    # different sizes, so that the solves converge after different numbers of steps
    return [
        H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0),
                      H.heisenberg_bond(20, qmax=3, width=4.0), seed=3, hermitian=True),
        H.HeffProblem(H.j1j2_model(1.0, 0.3), H.heisenberg_bond(40, qmax=3, width=6.0),
                      H.heisenberg_bond(40, qmax=3, width=6.0), seed=4, hermitian=True),
        H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(40, n0=0, sigma=1.2),
                      H.hubbard_bond(36, n0=2, sigma=1.2), seed=5, hermitian=True),
        H.HeffProblem(H.j1j2_model(1.0, 0.1), H.heisenberg_bond(64, qmax=4, width=8.0),
                      H.heisenberg_bond(64, qmax=4, width=8.0), seed=6, hermitian=True),
    ]


def alone_and_batched(probs, flags, **kw):
    x0 = [p.psi_parts["data"].copy() for p in probs]
    for p in probs:
        p.upload().outputs().make_plans(flags)
    alone = []
    for p, x in zip(probs, x0):
        p.psi.set_data_(x)
        vals, _, info = sb.eigsolve(p.plans, p.psi, **kw)
        alone.append((vals[0], info, p.psi.parts()["data"].copy()))
    for p, x in zip(probs, x0):
        p.psi.set_data_(x)
    batch = sb.PlanBatch([p.plans for p in probs])
    vals, infos = sb.eigsolve_batch(batch, [p.psi for p in probs], **kw)
    return alone, list(zip(vals, infos, [p.psi.parts()["data"].copy() for p in probs]))


def same(alone, batched):
    for (va, ia, xa), (vb, ib, xb) in zip(alone, batched):
        assert va == vb
        assert ia.normres == ib.normres and ia.numops == ib.numops and ia.numiter == ib.numiter
        assert ia.converged == ib.converged
        assert np.array_equal(xa, xb)


@pytest.mark.parametrize("flags", [sb.PREC_FP64, sb.PREC_FP32], ids=["fp64", "fp32"])
def test_batch_solves_equal_alone(flags):
    print(f"[eigsolve batch] converged solves flags={flags}", flush=True)
    probs = hermitian_problems()
    tol = 1e-10 if flags == sb.PREC_FP64 else 1e-5  # the FP32 products' floor is near 1e-6
    alone, batched = alone_and_batched(probs, flags, krylovdim=10, maxiter=30, tol=tol)
    same(alone, batched)
    if flags == sb.PREC_FP64:
        assert len({i.numops for _, i, _ in alone}) > 1  # the solves left the batch at different times
        assert all(i.converged == 1 for _, i, _ in batched)


def test_batch_dmrg_step_solves():
    """the DMRG call (krylovdim 3, maxiter 1): every chain does the same steps"""
    print("[eigsolve batch] dmrg step", flush=True)
    probs = hermitian_problems()
    alone, batched = alone_and_batched(probs, sb.PREC_FP64, krylovdim=3, maxiter=1, tol=1e-14)
    same(alone, batched)


CHAINS = [
    (lambda: H.heisenberg_model(1.0), 3, 3, (0,), lambda: spin_chain(8, [(1, 1.0, 1.0)])),
    (lambda: H.j1j2_model(1.0, 0.5), 3, 3, (0,), lambda: spin_chain(8, [(1, 1.0, 1.0), (2, 0.5, 0.5)])),
    (lambda: H.hubbard_model(1.0, 4.0), 2, 2, (6, 0), lambda: hubbard_chain(6, 1.0, 4.0)),
]


def test_batch_chains_reach_ed():
    print("[eigsolve batch] exact chains", flush=True)
    probs, refs = [], []
    for model, nl, nr, total, ed in CHAINS:
        Hm, q = ed()
        refs.append(ground_energy(Hm, q, total)[0])
        probs.append(H.chain_problem(model(), nl, nr, total, seed=1))
    alone, batched = alone_and_batched(probs, sb.PREC_FP64, krylovdim=30, maxiter=50, tol=1e-12)
    same(alone, batched)
    for (v, info, _), E0 in zip(batched, refs):
        assert info.converged == 1
        assert abs(v - E0) <= 1e-10, (v, E0)


def tridiag_operator(n, shift):
    Mat = np.diag(np.full(n, 2.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    Mat[shift, shift] = 0.5
    return Mat


LAB = lambda s: (int.from_bytes(("u:" + s).encode().ljust(8, b"\0"), "little"), 0)  # noqa: E731


def test_batch_widening():
    """operators whose reach grows one entry per application from a one-entry
    start vector: every chain's compact layout widens at every step"""
    print("[eigsolve batch] widening", flush=True)
    mats, xs, chains, single = [], [], [], []
    for n, sh in ((12, 3), (9, 5), (14, 1)):
        Mat = tridiag_operator(n, sh)
        x0 = np.zeros(n)
        x0[0] = 1.0
        X = sb.NewBlockSparseSorted.from_parts((1, n), 1, np.array([[1]]), np.array([1]), x0)
        Y = sb.NewBlockSparseSorted.new((1, n), 1)
        Mt = sb.NewBlockSparseSorted.from_parts((1, 1, n, n), 2, np.array([[1, 1]]), np.array([1]),
                                                Mat.reshape(-1, order="F"))
        plan = sb.Plan(Y, [LAB("Qp"), LAB("j")], X, [LAB("Q"), LAB("i")], Mt,
                       [LAB("Qp"), LAB("Q"), LAB("i"), LAB("j")])
        mats.append(Mat)
        xs.append((X, x0, Mt))
        chains.append([plan])
    for (X, x0, _), ch in zip(xs, chains):
        vals, _, info = sb.eigsolve(ch, X, krylovdim=14, maxiter=3, tol=1e-12)
        single.append((vals[0], info, X.parts()["data"].copy()))
        X.set_data_(x0)
    batch = sb.PlanBatch(chains)
    vals, infos = sb.eigsolve_batch(batch, [X for X, _, _ in xs], krylovdim=14, maxiter=3, tol=1e-12)
    same(single, list(zip(vals, infos, [X.parts()["data"].copy() for X, _, _ in xs])))
    for v, Mat in zip(vals, mats):
        assert abs(v - np.linalg.eigvalsh(Mat)[0]) <= 1e-12


def test_batch_eigsolve_errors():
    print("[eigsolve batch] errors", flush=True)
    probs = hermitian_problems()[:2]
    for p in probs:
        p.upload().outputs().make_plans(sb.PREC_FP64)
    batch = sb.PlanBatch([p.plans for p in probs])
    with pytest.raises(ValueError):
        sb.eigsolve_batch(batch, [probs[0].psi])
    with pytest.raises(sb.SBError):  # chain 0 does not read probs[1].psi
        sb.eigsolve_batch(batch, [probs[1].psi, probs[0].psi])
    with pytest.raises(sb.SBError):
        sb.eigsolve_batch(batch, [p.psi for p in probs], maxiter=0)
