"""GPU parity of the two-site H_eff*psi workload (four BS x BS contract! calls).

Small problems: eager contract! and plan replay against the oracle's chain
(keys and ids bit-exact, payload within 1e-12 relative Frobenius, the FP64
bar) and against a dense einsum over the full bonds.  The full-size Hubbard
chi=2048 workload: against the oracle (keys/ids exact, 1e-12) and through
linearity (H(2 psi) = 2 H psi bit for bit, as scaling by 2 is exact).
"""
import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from oracle import oracle as O
from oracle.heff_oracle import HeffOracle
from heff_dense import heff_dense, expand_out
from refs import rel_frobenius

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 2e-6


def small(name):
    if name == "heisenberg":
        return H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0),
                             H.heisenberg_bond(20, qmax=3, width=4.0), seed=3)
    if name == "hubbard":
        return H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(40, n0=0, sigma=1.2),
                             H.hubbard_bond(36, n0=2, sigma=1.2), seed=5)
    if name == "j1j2":
        return H.HeffProblem(H.j1j2_model(1.0, 0.4), H.heisenberg_bond(18, qmax=2, width=3.0),
                             H.heisenberg_bond(18, qmax=2, width=3.0), seed=7)
    if name == "hubbard-chi256":
        return H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(256, n0=0, sigma=3.0),
                             H.hubbard_bond(256, n0=2, sigma=3.0), seed=11)
    raise ValueError(name)


def same_blocks(p, q):
    assert tuple(p["dims"]) == tuple(q["dims"])
    assert np.array_equal(p["keys"], q["keys"])
    assert np.array_equal(p["ids"], q["ids"])


NAMES = ["heisenberg", "hubbard", "j1j2", "hubbard-chi256"]


@pytest.mark.parametrize("name", NAMES)
def test_eager_chain_vs_oracle(name):
    prob = small(name)
    print(f"[heff eager] {name}", flush=True)
    hp = prob.upload().outputs().contract_eager().parts()
    ref = HeffOracle(prob)
    q = ref.run().parts()
    same_blocks(hp, q)
    assert rel_frobenius(hp["data"], q["data"]) <= FP64_TOL
    # the intermediate T1..T3 too
    for dev, orc in ((prob.T1, ref.T1), (prob.T2, ref.T2), (prob.T3, ref.T3)):
        a, b = dev.parts(), orc.parts()
        same_blocks(a, b)
        assert rel_frobenius(a["data"], b["data"]) <= FP64_TOL
    if name != "hubbard-chi256":
        got, pad = expand_out(prob, hp)
        assert pad == 0.0
        d = heff_dense(prob)
        assert np.abs(got - d).max() <= 1e-12 * np.abs(d).max()


@pytest.mark.parametrize("name", NAMES)
def test_plan_replay(name):
    prob = small(name)
    print(f"[heff plan] {name}", flush=True)
    prob.upload().outputs()
    prob.make_plans(0)
    first = prob.Hpsi.parts()
    prob.execute()
    prob.execute()
    again = prob.Hpsi.parts()
    same_blocks(first, again)
    assert np.array_equal(first["data"], again["data"])  # a replay is deterministic
    # change psi's values inside an existing block, replay, compare with the oracle on the changed psi
    pp = prob.psi_parts
    bs = int(np.prod(pp["dims"][2:]))
    blk = pp["data"][:bs]
    j = int(np.flatnonzero(blk)[0])
    tail = np.unravel_index(j, pp["dims"][2:], order="F")
    I = tuple(int(x) for x in pp["keys"][0]) + tuple(int(t) + 1 for t in tail)
    prob.psi[I] = 0.625
    prob.execute()
    got = prob.Hpsi.parts()
    data = pp["data"].copy()
    data[j] = 0.625
    import copy
    p2 = copy.copy(prob)
    p2.psi_parts = dict(pp, data=data)
    q = HeffOracle(p2).run().parts()
    same_blocks(got, q)
    assert rel_frobenius(got["data"], q["data"]) <= FP64_TOL
    assert not np.array_equal(got["data"], first["data"])


@pytest.mark.parametrize("prec", [sb.PREC_FP32, sb.PREC_BF16X3])
@pytest.mark.parametrize("name", ["hubbard", "hubbard-chi256"])
def test_precision_forms(name, prec):
    prob = small(name)
    print(f"[heff prec] {name} {prec}", flush=True)
    ref = HeffOracle(prob).run().parts()
    prob.upload().outputs()
    prob.make_plans(prec)
    prob.execute()
    got = prob.Hpsi.parts()
    same_blocks(got, ref)
    assert rel_frobenius(got["data"], ref["data"]) <= MIXED_TOL


def test_full_hubbard_chi2048():
    """the benchmark workload at its full size"""
    prob = H.hubbard_chi2048_problem(0)
    print("[heff full] hubbard chi=2048", flush=True)
    prob.upload().outputs()
    prob.make_plans(0)
    prob.execute()
    got = prob.Hpsi.parts()
    O.use_blas(True, threads=8)
    try:
        q = HeffOracle(prob).run().parts()
    finally:
        O.use_blas(False)
    same_blocks(got, q)
    assert rel_frobenius(got["data"], q["data"]) <= FP64_TOL
    # linearity, exact: psi -> 2 psi gives 2 Hpsi bit for bit
    import copy
    p2 = copy.copy(prob)
    p2.psi_parts = dict(prob.psi_parts, data=2.0 * prob.psi_parts["data"])
    p2.upload().outputs()
    p2.make_plans(0)
    p2.execute()
    g2 = p2.Hpsi.parts()
    same_blocks(g2, got)
    assert np.array_equal(g2["data"], 2.0 * got["data"])
