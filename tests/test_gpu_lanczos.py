"""GPU eigensolver and Krylov vector operations through the C ABI.

Vector operations against the oracle's restatement and numpy (bit-exact on
dyadic data).  eigsolve: exact-chain H_eff ground energies against exact
diagonalisation (FP64 within 1e-10); the Lanczos trajectory against the
oracle's eigsolve run on the oracle's H_eff chain (same Ritz values to 1e-11
relative after a fixed number of applications, same application counts);
mixed-precision solves with an FP64 second stage within 1e-10 of ED."""
import copy

import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from oracle import lanczos_oracle as LO
from oracle import oracle as O
from chain_ed import spin_chain, hubbard_chain, ground_energy, OracleHeff

pytestmark = pytest.mark.gpu


def dyadic_bs(rng, dims, N2, fill=0.5):
    D = rng.integers(-8, 9, size=dims).astype(np.float64) / 8
    P = len(dims) - N2
    D[rng.random(dims[:P]) < fill] = 0.0
    return O.BS.from_dense(D, N2).parts()


def dev(p):
    return sb.NewBlockSparseSorted.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])


def test_vector_ops():
    rng = np.random.default_rng(1)
    dims, N2 = (3, 4, 2, 3), 2
    for _ in range(6):
        pa, pb = dyadic_bs(rng, dims, N2), dyadic_bs(rng, dims, N2)
        A, B = dev(pa), dev(pb)
        assert A.norm() == pytest.approx(np.linalg.norm(pa["data"]), rel=1e-15, abs=0)
        c = A.copy().parts()
        assert np.array_equal(c["keys"], pa["keys"]) and np.array_equal(c["ids"], pa["ids"])
        assert np.array_equal(c["data"], pa["data"])
        z = A.similar().parts()
        assert np.array_equal(z["ids"], pa["ids"]) and not np.any(z["data"])
        assert np.array_equal((0.75 * A).parts()["data"], 0.75 * pa["data"])
        s = (A + B).parts()
        r = LO.add(pa, pb)
        assert np.array_equal(s["keys"], r["keys"]) and np.array_equal(s["ids"], r["ids"])
        assert np.array_equal(s["data"], r["data"])
        assert A.dot(B) == LO.dot(pa, pb)
        R = A.copy()
        R.axpby_(0.5, A, -2.0)
        assert np.array_equal(R.parts()["data"], 0.5 * pa["data"] - 2.0 * pa["data"])
        R.fill_(0.25)
        assert np.all(R.parts()["data"] == 0.25)
    # payloads of different length (2 and 3 blocks): the reference's only check
    a = dict(dims=(2, 3, 2), N2=1, keys=np.array([[1, 1], [2, 2]]), ids=np.array([2, 1]), data=np.arange(4.0))
    b = dict(dims=(2, 3, 2), N2=1, keys=np.array([[2, 1], [2, 2], [1, 3]]), ids=np.array([1, 2, 3]),
             data=np.arange(6.0))
    with pytest.raises(sb.SBError) as e:
        dev(a).axpby_(1.0, dev(b), 1.0)
    assert e.value.kind == "ErrorException"
    s = (dev(a) + dev(b)).parts()
    r = LO.add(a, b)
    assert s["keys"].tolist() == r["keys"].tolist() and s["ids"].tolist() == [2, 3, 1, 4]
    assert np.array_equal(s["data"], r["data"])


CHAINS = [
    ("heisenberg-8", lambda: H.heisenberg_model(1.0), 3, 3, (0,), lambda: spin_chain(8, [(1, 1.0, 1.0)])),
    ("j1j2-8", lambda: H.j1j2_model(1.0, 0.5), 3, 3, (0,), lambda: spin_chain(8, [(1, 1.0, 1.0), (2, 0.5, 0.5)])),
    ("hubbard-6", lambda: H.hubbard_model(1.0, 4.0), 2, 2, (6, 0), lambda: hubbard_chain(6, 1.0, 4.0)),
]


@pytest.mark.parametrize("name,model,nl,nr,total,ed", CHAINS, ids=[c[0] for c in CHAINS])
def test_chain_ground_energy_fp64(name, model, nl, nr, total, ed):
    Hm, q = ed()
    E0, _ = ground_energy(Hm, q, total)
    prob = H.chain_problem(model(), nl, nr, total, seed=1)
    prob.upload().outputs()
    plans = prob.make_plans(sb.PREC_FP64)
    print(f"[eigsolve chain] {name}", flush=True)
    vals, vecs, info = sb.eigsolve(plans, prob.psi, krylovdim=30, maxiter=50, tol=1e-12)
    assert info.converged == 1
    assert abs(vals[0] - E0) <= 1e-10, (vals[0], E0)
    # the returned vector is the normalised Ritz vector: |H x - E x| small
    x = prob.psi.parts()["data"]
    assert abs(np.linalg.norm(x) - 1.0) <= 1e-13
    hx = OracleHeff(prob)(x)
    assert np.linalg.norm(hx - vals[0] * x) <= 1e-9


def small_hermitian(name):
    if name == "heisenberg":
        return H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0),
                             H.heisenberg_bond(20, qmax=3, width=4.0), seed=3, hermitian=True)
    return H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(40, n0=0, sigma=1.2),
                         H.hubbard_bond(36, n0=2, sigma=1.2), seed=5, hermitian=True)


@pytest.mark.parametrize("name", ["heisenberg", "hubbard"])
@pytest.mark.parametrize("k,maxiter", [(3, 1), (5, 3), (8, 2)])
def test_trajectory_vs_oracle(name, k, maxiter):
    """no convergence test (tol 0): a fixed number of applications, whose Ritz
    values the GPU and the oracle must agree on"""
    prob = small_hermitian(name)
    x0 = prob.psi_parts["data"].copy()
    e_o, x_o, info_o = LO.eigsolve(OracleHeff(prob), x0, krylovdim=k, maxiter=maxiter, tol=0.0)
    prob.upload().outputs()
    plans = prob.make_plans(sb.PREC_FP64)
    print(f"[eigsolve trajectory] {name} k={k} maxiter={maxiter}", flush=True)
    vals, vecs, info = sb.eigsolve(plans, prob.psi, krylovdim=k, maxiter=maxiter, tol=0.0)
    assert info.numops == info_o["numops"] and info.numiter == info_o["numiter"] and info.converged == 0
    scale = np.abs(info_o["ritz"]).max()
    np.testing.assert_allclose(info.ritz, info_o["ritz"], rtol=0, atol=1e-11 * scale)
    assert abs(info.normres - info_o["normres"]) <= 1e-9 * scale
    x = prob.psi.parts()["data"]
    assert abs(abs(x @ x_o) - 1.0) <= 1e-9


def test_converged_vector_vs_oracle_and_repeatable():
    prob = small_hermitian("hubbard")
    x0 = prob.psi_parts["data"].copy()
    e_o, x_o, info_o = LO.eigsolve(OracleHeff(prob), x0, krylovdim=20, maxiter=100, tol=1e-10)
    prob.upload().outputs()
    plans = prob.make_plans(sb.PREC_FP64)
    vals, _, info = sb.eigsolve(plans, prob.psi, krylovdim=20, maxiter=100, tol=1e-10)
    assert info.converged == 1 and info_o["converged"]
    assert abs(vals[0] - e_o) <= 1e-10 * abs(e_o)
    x = prob.psi.parts()["data"]
    assert abs(abs(x @ x_o) - 1.0) <= 1e-8
    # the same start gives the same numbers bit for bit (fixed reduction trees)
    prob.psi.set_data_(x0)
    vals2, _, info2 = sb.eigsolve(plans, prob.psi, krylovdim=20, maxiter=100, tol=1e-10)
    assert vals2[0] == vals[0] and info2.numops == info.numops
    assert np.array_equal(prob.psi.parts()["data"], x)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_then_fp64_reaches_ed(prec):
    Hm, q = hubbard_chain(6, 1.0, 4.0)
    E0, _ = ground_energy(Hm, q, (6, 0))
    prob = H.chain_problem(H.hubbard_model(1.0, 4.0), 2, 2, (6, 0), seed=1)
    lo = copy.copy(prob).upload().outputs()
    plans_lo = lo.make_plans(sb.PREC_FP32 if prec == "fp32" else sb.PREC_BF16X3)
    hi = prob.upload().outputs()
    plans_hi = hi.make_plans(sb.PREC_FP64)
    print(f"[eigsolve mixed] {prec}", flush=True)
    _, _, info_mixed = sb.eigsolve(plans_lo, lo.psi, krylovdim=30, maxiter=50, tol=1e-5)
    e_mixed = info_mixed.ritz[0]
    assert abs(e_mixed - E0) <= 1e-4 * abs(E0)  # the low-precision operator alone
    lo.psi.set_data_(prob.psi_parts["data"])
    # an FP64 residual of 1e-9 bounds the energy error by about residual^2 / gap
    vals, _, info_lo, info_hi = sb.eigsolve_refined(plans_lo, plans_hi, lo.psi, hi.psi, krylovdim=30, maxiter=50,
                                                    tol_lo=1e-5, tol=1e-9)
    assert info_hi.converged == 1
    assert abs(vals[0] - E0) <= 1e-10, (vals[0], E0, info_lo, info_hi)
    assert info_hi.numops < info_lo.numops  # most applications run in low precision


def test_errors():
    prob = small_hermitian("heisenberg")
    prob.upload().outputs()
    plans = prob.make_plans(sb.PREC_FP64)
    with pytest.raises(sb.SBError) as e:
        sb.eigsolve(plans, prob.T1, y=prob.Hpsi)  # the first plan does not read T1
    assert e.value.kind == "ArgumentError"
    with pytest.raises(sb.SBError) as e:
        sb.eigsolve(plans[:3], prob.psi, y=prob.Hpsi)  # the last plan writes T3
    assert e.value.kind == "ArgumentError"
    with pytest.raises(sb.SBError) as e:
        sb.eigsolve(plans, prob.psi, krylovdim=65)
    assert e.value.kind == "ArgumentError"
    prob.psi.fill_(0.0)
    with pytest.raises(sb.SBError) as e:
        sb.eigsolve(plans, prob.psi)
    assert e.value.kind == "ArgumentError"


def tridiag_operator(n):
    Mat = np.diag(np.full(n, 2.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    Mat[3, 3] = 0.5  # break the uniform spectrum
    return Mat


LAB = lambda s: (int.from_bytes(("t:" + s).encode().ljust(8, b"\0"), "little"), 0)  # noqa: E731


def test_compact_layout_widens():
    """an operator whose reach grows one entry per application (a symmetric
    tridiagonal matrix on a tail axis, applied as contract_shared! with the
    block products of H_eff's psi*L step) from a start vector holding one
    entry: the compact Krylov layout is widened at every step and the result
    equals the matrix's lowest eigenpair"""
    n = 12
    Mat = tridiag_operator(n)
    x0 = np.zeros(n)
    x0[0] = 1.0
    X = sb.NewBlockSparseSorted.from_parts((1, n), 1, np.array([[1]]), np.array([1]), x0)
    Y = sb.NewBlockSparseSorted.new((1, n), 1)
    Mt = sb.NewBlockSparseSorted.from_parts((1, 1, n, n), 2, np.array([[1, 1]]), np.array([1]),
                                            Mat.reshape(-1, order="F"))
    plan = sb.Plan(Y, [LAB("Qp"), LAB("j")], X, [LAB("Q"), LAB("i")], Mt, [LAB("Qp"), LAB("Q"), LAB("i"), LAB("j")])
    print("[eigsolve widening]", flush=True)
    vals, _, info = sb.eigsolve([plan], X, krylovdim=n, maxiter=3, tol=1e-12)
    w, U = np.linalg.eigh(Mat)
    assert info.converged == 1
    assert abs(vals[0] - w[0]) <= 1e-12
    x = X.parts()["data"]
    assert abs(abs(x @ U[:, 0]) - 1.0) <= 1e-10


def test_rank1_plan_replay_follows_values():
    """a single-label BS x BS plan over a dense-tail index (contract_dense_bb!,
    GER per r): replayed with a different zero pattern, every r slice that is
    nonzero now contributes (the block list stays the first call's)"""
    n = 12
    Mat = tridiag_operator(n)
    x0 = np.zeros(n)
    x0[0] = 1.0
    X = sb.NewBlockSparseSorted.from_parts((1, n), 1, np.array([[1]]), np.array([1]), x0)
    Y = sb.NewBlockSparseSorted.new((1, n), 1)
    Mt = sb.NewBlockSparseSorted.from_parts((n, n), 2, np.zeros((1, 0), dtype=np.int64), np.array([1]),
                                            Mat.reshape(-1, order="F"))
    plan = sb.Plan(Y, [LAB("Q"), LAB("j")], X, [LAB("Q"), LAB("i")], Mt, [LAB("i"), LAB("j")])
    v = np.arange(1.0, n + 1.0)
    X.set_data_(v)
    plan.execute()
    np.testing.assert_allclose(Y.parts()["data"], Mat.T @ v, rtol=1e-14, atol=1e-12)
