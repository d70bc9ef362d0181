"""GPU parity of batched plan chains (PlanBatch / sb_batch_*): the four
H_eff*psi contractions of several independent problems run stage by stage,
each stage one box product launch over every chain's work list.

A batch must give bit-for-bit what each chain's plans give when executed
alone (same tiles' arithmetic, same K order), in FP64 and in the FP32 form;
one member is also checked against the oracle's contract! chain (keys / ids
exact, 1e-12 relative Frobenius).  Inactive chains are left untouched, stages
that are not one box product run plan by plan, and malformed batches fail.
"""
import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from oracle.heff_oracle import HeffOracle
from refs import rel_frobenius

pytestmark = pytest.mark.gpu


def problems():
    # different models, bond layouts and block sizes in one batch
    return [
        H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0),
                      H.heisenberg_bond(20, qmax=3, width=4.0), seed=3),
        H.HeffProblem(H.j1j2_model(1.0, 0.3), H.heisenberg_bond(40, qmax=3, width=6.0),
                      H.heisenberg_bond(40, qmax=3, width=6.0), seed=4, hermitian=True),
        H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(40, n0=0, sigma=1.2),
                      H.hubbard_bond(36, n0=2, sigma=1.2), seed=5),
        H.HeffProblem(H.j1j2_model(1.0, 0.1), H.heisenberg_bond(96, qmax=4, width=8.0),
                      H.heisenberg_bond(96, qmax=4, width=8.0), seed=6),
        H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(128, n0=0, sigma=2.5),
                      H.hubbard_bond(128, n0=2, sigma=2.5), seed=7),
    ]


def built(flags=0):
    ps = problems()
    for p in ps:
        p.upload().outputs().make_plans(flags)
    return ps


def outputs(p):
    return [t.parts()["data"].copy() for t in (p.T1, p.T2, p.T3, p.Hpsi)]


def scale_psi(p, a):
    p.psi.axpby_(0.0, p.psi, a)  # psi <- a * psi (exact for a power of two)


@pytest.mark.parametrize("flags", [sb.PREC_FP64, sb.PREC_FP32], ids=["fp64", "fp32"])
def test_batch_equals_chains_alone(flags):
    print(f"[batch] equals chains alone flags={flags}", flush=True)
    ps = built(flags)
    batch = sb.PlanBatch([p.plans for p in ps])
    assert batch.stats()["fused_stages"] == 4
    assert batch.stats()["flops"] == pytest.approx(sum(q.stats()["flops"] for p in ps for q in p.plans))
    ref1 = [outputs(p) for p in ps]          # each chain alone, from planning
    for p in ps:
        scale_psi(p, 2.0)
    batch.execute()
    got = [outputs(p) for p in ps]
    for p in ps:
        p.execute()
    alone = [outputs(p) for p in ps]
    for c in range(len(ps)):
        for k in range(4):
            assert np.array_equal(got[c][k], alone[c][k]), (c, k)
            assert np.array_equal(got[c][k], 2.0 * ref1[c][k]), (c, k)
    # a second execution of the batch replays the same work
    batch.execute()
    for c, p in enumerate(ps):
        assert np.array_equal(outputs(p)[3], got[c][3])


def test_batch_member_vs_oracle():
    print("[batch] member vs oracle", flush=True)
    ps = built()
    batch = sb.PlanBatch([p.plans for p in ps])
    for p in ps:
        p.Hpsi.fill_(0.0)
    batch.execute()
    for c in (1, 4):
        q = HeffOracle(ps[c]).run().parts()
        hp = ps[c].Hpsi.parts()
        assert np.array_equal(hp["keys"], q["keys"]) and np.array_equal(hp["ids"], q["ids"])
        assert rel_frobenius(hp["data"], q["data"]) <= 1e-12


def test_batch_active_subset():
    print("[batch] active subset", flush=True)
    ps = built()
    batch = sb.PlanBatch([p.plans for p in ps])
    before = [outputs(p)[3] for p in ps]
    for p in ps:
        scale_psi(p, 4.0)
    batch.execute([1, 3])
    for c, p in enumerate(ps):
        h = outputs(p)[3]
        if c in (1, 3):
            assert np.array_equal(h, 4.0 * before[c])
        else:
            assert np.array_equal(h, before[c])
    batch.execute([0, 2, 4])  # another active set: new job list
    for c, p in enumerate(ps):
        assert np.array_equal(outputs(p)[3], 4.0 * before[c])


def test_batch_unfused_stage():
    """a stage that is not one box product (Hpsi x a dense d x d matrix on s1',
    the BS x Dense family) runs plan by plan inside the batch"""
    print("[batch] unfused stage", flush=True)
    ps = built()[:3]
    t = (int.from_bytes(b"t:newax", "little"), 0)
    chains, extra = [], []
    rng = np.random.default_rng(1)
    for p in ps:
        d = p.m.d
        M = rng.standard_normal((d, d))
        Dm = sb.DenseTensor.from_numpy(M)
        dims = p.Hpsi.dims
        Y = sb.NewBlockSparseSorted.new((dims[0], dims[1], d) + tuple(dims[3:]), 4)
        lab = [H.L_OUT[0], H.L_OUT[1], t] + H.L_OUT[3:]
        extra.append((Y, M))
        chains.append(p.plans + [sb.Plan(Y, lab, p.Hpsi, H.L_OUT, Dm, [H.L_OUT[2], t])])
    batch = sb.PlanBatch(chains)
    assert batch.stats()["fused_stages"] == 4
    for p in ps:
        scale_psi(p, 0.5)
    batch.execute()
    for p, (Y, M) in zip(ps, extra):
        h, y = p.Hpsi.parts(), Y.parts()
        tail = tuple(int(x) for x in h["dims"][2:])
        bs = int(np.prod(tail))
        hid = {tuple(int(v) for v in k): int(i) for k, i in zip(h["keys"], h["ids"])}
        assert len(y["ids"]) > 0
        for k, i_y in zip(y["keys"], y["ids"]):
            i_h = hid[tuple(int(v) for v in k)]
            hb = h["data"][(i_h - 1) * bs:i_h * bs].reshape(tail, order="F")
            yb = y["data"][(i_y - 1) * bs:i_y * bs].reshape(tail, order="F")
            # Y[.., t, s2', l', r'] = sum_s1' Hpsi[.., s1', s2', l', r'] M[s1', t]
            assert np.allclose(yb, np.einsum("abcd,at->tbcd", hb, M), rtol=1e-12, atol=1e-12)
    # the same chains executed alone give the batch's bits
    got = [Y.parts()["data"].copy() for Y, _ in extra]
    for ch in chains:
        for q in ch:
            q.execute()
    for (Y, _), g in zip(extra, got):
        assert np.array_equal(Y.parts()["data"], g)


def test_batch_errors():
    print("[batch] errors", flush=True)
    ps = built()[:2]
    with pytest.raises(ValueError):
        sb.PlanBatch([ps[0].plans, ps[1].plans[:3]])
    batch = sb.PlanBatch([p.plans for p in ps])
    with pytest.raises(sb.SBError):
        batch.execute([1, 0])
    with pytest.raises(sb.SBError):
        batch.execute([2])
