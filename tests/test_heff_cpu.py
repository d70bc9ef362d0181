"""The H_eff*psi workload on the CPU: site models against their defining
Hamiltonians, and the four-call block chain (run on the oracle) against one
dense einsum over the full bond spaces."""
import numpy as np
import pytest

from sparsebackends_amd import heff as H
from oracle.heff_oracle import HeffOracle
from heff_dense import heff_dense, expand_out


def two_site_from_mpo(m):
    """sum over the link of W[start, s1', s1, a] W[a, s2', s2, done] as a (d*d) x (d*d)
    matrix over (s1, s2) with s1 most significant"""
    W = m.W
    start, done = m.w - 1, 0
    h = np.einsum("xya,azu->xzyu", W[start], W[:, :, :, done])
    d = m.d
    return h.reshape(d * d, d * d)


def test_heisenberg_mpo_two_site():
    J, Jz = 1.3, 0.7
    m = H.heisenberg_model(J, Jz)
    Sz = np.diag([0.5, -0.5])
    Sx = np.array([[0, 0.5], [0.5, 0]])
    Sy2 = np.array([[0, -0.5], [0.5, 0]])  # i*Sy, real
    h = J * (np.kron(Sx, Sx) - np.kron(Sy2, Sy2)) + Jz * np.kron(Sz, Sz)
    np.testing.assert_allclose(two_site_from_mpo(m), h, atol=1e-15)


def test_hubbard_mpo_two_site():
    """Jordan-Wigner with modes ordered (1up, 1dn, 2up, 2dn): the MPO's two-site
    Hamiltonian equals -t sum_s (c+_1s c_2s + h.c.) + U sum_i n_iup n_idn"""
    t, U = 0.8, 3.5
    m = H.hubbard_model(t, U)
    # site basis [0, up, dn, updn], |updn> = c+_up c+_dn |0>
    cu = np.zeros((4, 4))
    cd = np.zeros((4, 4))
    cu[0, 1] = 1
    cu[2, 3] = 1
    cd[0, 2] = 1
    cd[1, 3] = -1
    F = np.diag([1.0, -1, -1, 1])
    I = np.eye(4)
    c = {("1", "u"): np.kron(cu, I), ("1", "d"): np.kron(cd, I), ("2", "u"): np.kron(F, cu), ("2", "d"): np.kron(F, cd)}
    # canonical anticommutation of the global operators pins the JW construction
    for a in c.values():
        for b in c.values():
            assert np.allclose(a @ b + b @ a, 0)
            ac = a @ b.T + b.T @ a
            assert np.allclose(ac, np.eye(16) if a is b else 0)
    h = np.zeros((16, 16))
    for s in "ud":
        hop = c[("1", s)].T @ c[("2", s)]
        h += -t * (hop + hop.T)
    for i in "12":
        h += U * (c[(i, "u")].T @ c[(i, "u")]) @ (c[(i, "d")].T @ c[(i, "d")])
    np.testing.assert_allclose(two_site_from_mpo(m), h, atol=1e-15)


@pytest.mark.parametrize("model", ["heisenberg", "hubbard", "j1j2"])
def test_mpo_charge_rule(model):
    m = {"heisenberg": H.heisenberg_model(), "hubbard": H.hubbard_model(), "j1j2": H.j1j2_model(1.0, 0.5)}[model]
    q, Ql = m.charges, m.Qlink
    for a, sp, s, b in zip(*np.nonzero(m.W)):
        assert H.qsub(q[sp], q[s]) == H.qsub(Ql[a], Ql[b])


def small_problems():
    yield "heisenberg", H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0),
                                      H.heisenberg_bond(20, qmax=3, width=4.0), seed=3)
    Bl = H.hubbard_bond(40, n0=0, sigma=1.2)
    Br = H.hubbard_bond(36, n0=2, sigma=1.2)
    yield "hubbard", H.HeffProblem(H.hubbard_model(1.0, 4.0), Bl, Br, seed=5)
    yield "j1j2", H.HeffProblem(H.j1j2_model(1.0, 0.4), H.heisenberg_bond(18, qmax=2, width=3.0),
                                H.heisenberg_bond(18, qmax=2, width=3.0), seed=7)


@pytest.mark.parametrize("name,prob", list(small_problems()), ids=lambda x: x if isinstance(x, str) else "")
def test_block_chain_equals_dense(name, prob):
    ref = heff_dense(prob)
    hp = HeffOracle(prob).run()
    got, pad = expand_out(prob, hp.parts())
    assert pad == 0.0  # padding of the in-sector axes stays exactly zero
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # the charge rule of the output: blocks only where q(r') = q(l') + q(s1') + q(s2')
    assert np.abs(ref).max() > 0


def test_flop_accounting_matches_pairs():
    """algorithmic_flops counts 2*M*K*N over matched pairs: check it against a
    direct enumeration of the four steps' pairs on a small problem"""
    _, prob = next(iter(small_problems()))
    f = prob.algorithmic_flops()
    d, w = prob.m.d, prob.m.w
    dl, dr = prob.Bl.dims, prob.Br.dims
    psi = [tuple(map(int, k)) for k in prob.psi_parts["keys"]]
    Lk = [tuple(map(int, k)) for k in prob.L_parts["keys"]]
    Rk = [tuple(map(int, k)) for k in prob.R_parts["keys"]]
    s1 = sum(2.0 * d * d * dr[ir - 1] * dl[il - 1] * dl[a - 1] * w for (il, ir) in psi for (a, b) in Lk if b == il)
    T1 = {(a, ir) for (il, ir) in psi for (a, b) in Lk if b == il}
    s2 = sum(2.0 * d * dr[ir - 1] * dl[a - 1] * (d * w) ** 2 for (a, ir) in T1)
    s4 = sum(2.0 * dl[a - 1] * d * d * dr[ir - 1] * w * dr[c - 1] for (a, ir) in T1 for (c, e) in Rk if e == ir)
    np.testing.assert_allclose(f["steps"], [s1, s2, s2, s4])
