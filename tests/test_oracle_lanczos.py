"""The oracle's eigensolver and Krylov vector operations, pinned on the CPU:
eigenvalues against numpy's eigh and against Rayleigh-Ritz on an explicitly
built Krylov space, special starts and dimensions, restart accounting, and
exact-chain H_eff ground energies against exact diagonalisation of chain
Hamiltonians built from spin and fermion operators."""
import numpy as np
import pytest

from sparsebackends_amd import heff as H
from oracle import lanczos_oracle as LO
from oracle import oracle as O
from chain_ed import spin_chain, hubbard_chain, ground_energy, OracleHeff


def rand_sym(rng, n, spread=1.0):
    A = rng.standard_normal((n, n))
    return (A + A.T) * 0.5 * spread


def test_lowest_vs_eigh():
    rng = np.random.default_rng(0)
    for n in (5, 40, 200):
        A = rand_sym(rng, n)
        x0 = rng.standard_normal(n)
        e, x, info = LO.eigsolve(lambda v: A @ v, x0, krylovdim=30, maxiter=200, tol=1e-12)
        ref = np.linalg.eigvalsh(A)[0]
        assert info["converged"]
        assert abs(e - ref) <= 1e-10 * max(1.0, abs(ref))
        assert np.linalg.norm(A @ x - e * x) <= 1e-10 * np.linalg.norm(A, 2)
        assert abs(np.linalg.norm(x) - 1.0) <= 1e-14


@pytest.mark.parametrize("k", [2, 3, 5, 8])
def test_first_factorisation_is_krylov_rayleigh_ritz(k):
    """one factorisation of k vectors = Rayleigh-Ritz on span{x0, A x0, ..., A^(k-1) x0}"""
    rng = np.random.default_rng(k)
    n = 60
    A = rand_sym(rng, n)
    x0 = rng.standard_normal(n)
    e, x, info = LO.eigsolve(lambda v: A @ v, x0, krylovdim=k, maxiter=1, tol=0.0)
    K = np.empty((n, k))
    K[:, 0] = x0 / np.linalg.norm(x0)
    for j in range(1, k):
        K[:, j] = A @ K[:, j - 1]
        K[:, j] /= np.linalg.norm(K[:, j])
    Q, _ = np.linalg.qr(K)
    ritz = np.linalg.eigvalsh(Q.T @ A @ Q)
    np.testing.assert_allclose(info["ritz"], ritz, rtol=0, atol=1e-10 * np.abs(ritz).max())
    assert info["numops"] == k and info["numiter"] == 1 and not info["converged"]


def test_eigenvector_start():
    rng = np.random.default_rng(3)
    A = rand_sym(rng, 30)
    w, U = np.linalg.eigh(A)
    e, x, info = LO.eigsolve(lambda v: A @ v, U[:, 4] * 3.0, krylovdim=6, maxiter=5, tol=1e-12)
    assert info["numops"] == 1 and info["converged"]
    assert abs(e - w[4]) <= 1e-13
    assert abs(abs(x @ U[:, 4]) - 1.0) <= 1e-13


def test_dimension_below_krylovdim():
    A = np.array([[2.0, 1.0, 0.0], [1.0, -1.0, 0.5], [0.0, 0.5, 3.0]])
    e, x, info = LO.eigsolve(lambda v: A @ v, np.array([1.0, 1.0, 1.0]), krylovdim=8, maxiter=1, tol=1e-14)
    assert info["converged"] and info["numops"] <= 3
    assert abs(e - np.linalg.eigvalsh(A)[0]) <= 1e-13


def test_restart_accounting():
    """krylovdim 3: a factorisation of 3 applications, then restarts keeping
    max(1, 9 // 5) = 1 Ritz vector and the residual direction: 2 per cycle"""
    rng = np.random.default_rng(4)
    A = rand_sym(rng, 50)
    x0 = rng.standard_normal(50)
    for maxiter, ops in ((1, 3), (2, 5), (4, 9)):
        e, x, info = LO.eigsolve(lambda v: A @ v, x0, krylovdim=3, maxiter=maxiter, tol=0.0)
        assert info["numops"] == ops and info["numiter"] == maxiter and not info["converged"]
    # restarted runs approach the answer monotonically (Ritz values of nested spaces)
    vals = [LO.eigsolve(lambda v: A @ v, x0, krylovdim=5, maxiter=m, tol=0.0)[0] for m in (1, 2, 4, 8)]
    assert all(b <= a + 1e-12 for a, b in zip(vals, vals[1:]))


CHAINS = [
    ("heisenberg-8", lambda: H.heisenberg_model(1.0), 3, 3, (0,), lambda: spin_chain(8, [(1, 1.0, 1.0)])),
    ("xxz-8", lambda: H.heisenberg_model(1.0, 0.6), 3, 3, (2,), lambda: spin_chain(8, [(1, 1.0, 0.6)])),
    ("j1j2-8", lambda: H.j1j2_model(1.0, 0.5), 3, 3, (0,), lambda: spin_chain(8, [(1, 1.0, 1.0), (2, 0.5, 0.5)])),
    ("hubbard-6", lambda: H.hubbard_model(1.0, 4.0), 2, 2, (6, 0), lambda: hubbard_chain(6, 1.0, 4.0)),
    ("hubbard-6-doped", lambda: H.hubbard_model(1.0, 2.0), 2, 2, (4, 2), lambda: hubbard_chain(6, 1.0, 2.0)),
]


@pytest.mark.parametrize("name,model,nl,nr,total,ed", CHAINS, ids=[c[0] for c in CHAINS])
def test_chain_heff_ground_energy(name, model, nl, nr, total, ed):
    """exact boundary environments make H_eff the whole chain's Hamiltonian in
    the sector of the total charge: the solver's energy is the ED ground energy"""
    Hm, q = ed()
    E0, dim = ground_energy(Hm, q, total)
    prob = H.chain_problem(model(), nl, nr, total, seed=1)
    e, x, info = LO.eigsolve(OracleHeff(prob), prob.psi_parts["data"].copy(), krylovdim=30, maxiter=50, tol=1e-12)
    assert info["converged"]
    assert abs(e - E0) <= 1e-10, (e, E0)


@pytest.mark.parametrize("model", [H.heisenberg_model(1.0, 0.7), H.j1j2_model(1.0, 0.3), H.hubbard_model(1.0, 4.0)],
                         ids=["heisenberg", "j1j2", "hubbard"])
def test_link_involution(model):
    c = model.link_conj
    assert sorted(c) == list(range(model.w)) and all(c[c[a]] == a for a in range(model.w))
    for a in range(model.w):
        assert model.Qlink[c[a]] == tuple(-x for x in model.Qlink[a])
        for b in range(model.w):
            assert np.array_equal(model.W[a, :, :, b].T, model.W[c[a], :, :, c[b]])


def test_hermitian_environments_make_heff_symmetric():
    prob = H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(40, n0=0, sigma=1.2),
                         H.hubbard_bond(36, n0=2, sigma=1.2), seed=5, hermitian=True)
    mv = OracleHeff(prob)
    rng = np.random.default_rng(0)
    u = prob.psi_parts["data"]
    v = u * rng.standard_normal(u.size)  # same support (zero padding)
    a, b = u @ mv(v), mv(u) @ v
    assert abs(a - b) <= 1e-12 * abs(a)
    raw = H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(40, n0=0, sigma=1.2),
                        H.hubbard_bond(36, n0=2, sigma=1.2), seed=5)
    mr = OracleHeff(raw)
    assert abs(u @ mr(v) - mr(u) @ v) > 1e-6 * abs(u @ mr(v))  # random environments are not


# ---------------------------------------------------------------- vector operations


def test_add_literal():
    """_ensure_block! order: new keys of B take the next ids in B's key order"""
    dims, N2 = (2, 3, 2), 1
    a = dict(dims=dims, N2=N2, keys=np.array([[1, 1], [2, 2]]), ids=np.array([2, 1]),
             data=np.array([1.0, 2.0, 3.0, 4.0]))
    b = dict(dims=dims, N2=N2, keys=np.array([[2, 1], [2, 2], [1, 3]]), ids=np.array([1, 2, 3]),
             data=np.array([10.0, 20.0, 30.0, 40.0, 50.0, 60.0]))
    c = LO.add(a, b)
    assert c["keys"].tolist() == [[1, 1], [2, 1], [2, 2], [1, 3]]
    assert c["ids"].tolist() == [2, 3, 1, 4]
    assert c["data"].tolist() == [1.0 + 30.0, 2.0 + 40.0, 3.0, 4.0, 10.0, 20.0, 50.0, 60.0]


def test_add_and_dot_vs_dense():
    rng = np.random.default_rng(9)
    for _ in range(10):
        dims, N2 = (3, 4, 2, 3), 2
        parts = []
        for _ in range(2):
            D = rng.integers(-4, 5, size=dims).astype(np.float64) / 4
            D[rng.random((3, 4)) < 0.5] = 0.0
            parts.append(O.BS.from_dense(D, N2).parts())
        c = LO.add(parts[0], parts[1])
        dA = O.BS.from_parts(**{k: parts[0][k] for k in ("dims", "N2", "keys", "ids", "data")}).to_dense()
        dB = O.BS.from_parts(**{k: parts[1][k] for k in ("dims", "N2", "keys", "ids", "data")}).to_dense()
        dC = O.BS.from_parts(c["dims"], c["N2"], c["keys"], c["ids"], c["data"]).to_dense()
        assert np.array_equal(dC, dA + dB)
        assert LO.dot(parts[0], parts[1]) == float(np.sum(dA * dB))
