"""GPU: the fused two-site H_eff*psi plan (sb_heff_plan_create) against the
oracle's four contract! calls: block list (keys and ids) bit for bit, values
within 1e-12 relative Frobenius (the FP64 bar; the fused kernel sums in an
order of its own), FP32 within 2e-6; replays with new psi values; eigensolves
whose operator is the fused plan against exact diagonalisation; the full
chi=2048 Hubbard workload against the four-call chain on the device."""
import copy

import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from oracle.heff_oracle import HeffOracle
from refs import rel_frobenius
from chain_ed import spin_chain, hubbard_chain, ground_energy
from test_gpu_heff import small, same_blocks, NAMES

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


@pytest.mark.parametrize("name", NAMES)
def test_fused_vs_oracle(name):
    prob = small(name)
    print(f"[heff fused] {name}", flush=True)
    prob.upload()
    prob.make_fused_plan(sb.PREC_FP64)
    got = prob.Hpsi_fused.parts()
    q = HeffOracle(prob).run().parts()
    same_blocks(got, q)
    assert rel_frobenius(got["data"], q["data"]) <= FP64_TOL
    prob.fused.execute()
    again = prob.Hpsi_fused.parts()
    assert np.array_equal(again["data"], got["data"])  # deterministic


@pytest.mark.parametrize("name", ["heisenberg", "hubbard"])
def test_fused_replay_new_psi(name):
    prob = small(name)
    print(f"[heff fused replay] {name}", flush=True)
    prob.upload()
    prob.make_fused_plan(0)
    first = prob.Hpsi_fused.parts()
    rng = np.random.default_rng(17)
    pp = prob.psi_parts
    data = pp["data"] * rng.uniform(0.5, 1.5, size=pp["data"].shape)  # same zero pattern, new values
    prob.psi.set_data_(data)
    prob.fused.execute()
    got = prob.Hpsi_fused.parts()
    p2 = copy.copy(prob)
    p2.psi_parts = dict(pp, data=data)
    q = HeffOracle(p2).run().parts()
    same_blocks(got, q)
    assert rel_frobenius(got["data"], q["data"]) <= FP64_TOL
    assert not np.array_equal(got["data"], first["data"])


def _oracle_for(prob, data):
    p2 = copy.copy(prob)
    p2.psi_parts = dict(prob.psi_parts, data=data)
    return HeffOracle(p2).run().parts()


@pytest.mark.parametrize("name", ["hubbard", "heisenberg"])
def test_fused_rows_change_between_executions(name):
    """The rows of psi that hold nonzeros are found at every execution:
    psi with blocks cleared (their output blocks must return to zero), then
    psi with values in every row of the reach, charge-forbidden rows too (the
    blocks of forbidden charge become nonzero), then the first psi again"""
    prob = small(name)
    print(f"[heff fused rows] {name}", flush=True)
    prob.upload()
    prob.make_fused_plan(0)
    pp = prob.psi_parts
    d, Dl, Dr = prob.m.d, prob.Dl, prob.Dr
    bs = d * d * Dl * Dr
    rng = np.random.default_rng(23)
    half = pp["data"].copy()
    for i in range(0, len(pp["ids"]), 2):
        half[i * bs:(i + 1) * bs] = 0.0
    full = pp["data"].copy()
    for i, (a, b) in enumerate(pp["keys"]):
        blk = np.zeros((d, d, Dl, Dr), order="F")
        blk[:, :, :prob.Bl.dims[a - 1], :prob.Br.dims[b - 1]] = rng.standard_normal(
            (d, d, prob.Bl.dims[a - 1], prob.Br.dims[b - 1]))
        full[i * bs:(i + 1) * bs] = blk.reshape(-1, order="F")
    for data in (half, full, pp["data"], half):
        prob.psi.set_data_(data)
        prob.fused.execute()
        got = prob.Hpsi_fused.parts()
        q = _oracle_for(prob, data)
        same_blocks(got, q)
        assert rel_frobenius(got["data"], q["data"]) <= FP64_TOL


def test_fused_after_output_written_outside():
    """Hpsi written between executions: every block is recomputed or cleared"""
    prob = small("hubbard")
    prob.upload()
    prob.make_fused_plan(0)
    first = prob.Hpsi_fused.parts()
    prob.Hpsi_fused.set_data_(np.random.default_rng(5).standard_normal(first["data"].shape))
    prob.fused.execute()
    assert np.array_equal(prob.Hpsi_fused.parts()["data"], first["data"])


@pytest.mark.parametrize("name", ["hubbard", "hubbard-chi256"])
def test_fused_fp32(name):
    prob = small(name)
    print(f"[heff fused fp32] {name}", flush=True)
    prob.upload()
    prob.make_fused_plan(sb.PREC_FP32)
    got = prob.Hpsi_fused.parts()
    q = HeffOracle(prob).run().parts()
    same_blocks(got, q)
    err = rel_frobenius(got["data"], q["data"])
    assert 0 < err <= FP32_TOL


def test_fused_refuses_changed_environment_and_bf16x3():
    prob = small("heisenberg")
    prob.upload()
    with pytest.raises(sb.SBError) as e:
        prob.make_fused_plan(sb.PREC_BF16X3)
    assert e.value.kind == "ArgumentError"
    plan = prob.make_fused_plan(0)
    k = tuple(int(x) for x in prob.L_parts["keys"][0])
    prob.L[k + (1, 1, 1)] = 0.5  # a library call changes L
    with pytest.raises(sb.SBError) as e:
        plan.execute()
    assert e.value.kind == "ErrorException"


def test_fused_limit_on_large_sectors():
    """sectors of ~52 states: output blocks beyond the fused kernel's reach"""
    B = H.heisenberg_bond(512)
    prob = H.HeffProblem(H.heisenberg_model(1.0), B, B, seed=2)
    prob.upload()
    with pytest.raises(sb.SBError) as e:
        prob.make_fused_plan(0)
    assert e.value.kind == "LimitExceeded"


CHAINS = [
    ("heisenberg-8", lambda: H.heisenberg_model(1.0), 3, 3, (0,), lambda: spin_chain(8, [(1, 1.0, 1.0)])),
    ("hubbard-6", lambda: H.hubbard_model(1.0, 4.0), 2, 2, (6, 0), lambda: hubbard_chain(6, 1.0, 4.0)),
]


@pytest.mark.parametrize("name,model,nl,nr,total,ed", CHAINS, ids=[c[0] for c in CHAINS])
def test_fused_eigsolve_matches_ed(name, model, nl, nr, total, ed):
    Hm, q = ed()
    E0, _ = ground_energy(Hm, q, total)
    prob = H.chain_problem(model(), nl, nr, total, seed=1)
    prob.upload()
    plan = prob.make_fused_plan(0)
    print(f"[heff fused eigsolve] {name}", flush=True)
    vals, vecs, info = sb.eigsolve([plan], prob.psi, krylovdim=30, maxiter=50, tol=1e-12)
    assert info.converged == 1
    assert abs(vals[0] - E0) <= 1e-10, (vals[0], E0)


def test_fused_full_hubbard_chi2048():
    prob = H.hubbard_chi2048_problem(0)
    print("[heff fused full] hubbard chi=2048", flush=True)
    prob.upload().outputs()
    prob.make_plans(0)
    prob.execute()
    chain = prob.Hpsi.parts()
    prob.make_fused_plan(0)
    got = prob.Hpsi_fused.parts()
    same_blocks(got, chain)
    assert rel_frobenius(got["data"], chain["data"]) <= FP64_TOL
    st = prob.fused.stats()
    assert st["nblocks_out"] == len(chain["ids"]) and st["flops"] > 0
    prob.make_fused_plan(sb.PREC_FP32)
    g32 = prob.Hpsi_fused.parts()
    assert rel_frobenius(g32["data"], chain["data"]) <= FP32_TOL


def test_fused_zero_psi_then_values():
    """psi of zeros gives Hpsi of zeros exactly (every block has nothing to
    multiply and is cleared), and the next execution with values is whole"""
    prob = small("hubbard")
    print("[heff fused zero psi]", flush=True)
    prob.upload()
    prob.make_fused_plan(0)
    pp = prob.psi_parts
    prob.psi.set_data_(np.zeros_like(pp["data"]))
    prob.fused.execute()
    z = prob.Hpsi_fused.parts()
    assert len(z["ids"]) > 0 and not np.any(z["data"])
    prob.psi.set_data_(pp["data"])
    prob.fused.execute()
    got = prob.Hpsi_fused.parts()
    q = HeffOracle(prob).run().parts()
    same_blocks(got, q)
    assert rel_frobenius(got["data"], q["data"]) <= FP64_TOL


def test_fused_non_finite_psi_reaches_output():
    """a NaN in psi reaches Hpsi (the rows holding it are multiplied), so an
    eigensolve over the plan sees it; where it spreads may be narrower than in
    the four-call chain, whose dense MPO steps multiply it by every MPO entry"""
    prob = small("hubbard")
    prob.upload()
    prob.make_fused_plan(0)
    data = prob.psi_parts["data"].copy()
    data[int(np.flatnonzero(data)[3])] = np.nan
    prob.psi.set_data_(data)
    prob.fused.execute()
    assert np.isnan(prob.Hpsi_fused.parts()["data"]).any()
