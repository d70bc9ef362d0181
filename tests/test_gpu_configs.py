"""GPU tests at the sizes of the benchmark configurations.

* Batching (64 independent S=1/2 J1-J2 problems, J2 ~ U[0, 0.5], chi
  log-uniform in [64, 1024] in 17 Sz sectors, seeds 0..63): one batch
  execution equals every chain executed alone, bit for bit, in FP64 and FP32;
  two members against the oracle's four contract! calls at 1e-12.
* Precision (Hubbard t=1 U=4, chi=2048 in 520 (n, 2Sz) sectors): the FP32 and
  bf16x3 products of the four-call chain and the fused plan's FP32 form
  within 2e-6 of the oracle's FP64 result.
* Serving (XXZ requests at N=32, chi_max=128 and N=64, chi_max=256): every
  served energy equals the one-at-a-time run's bit for bit.
"""
import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
from oracle import oracle as O
from oracle.heff_oracle import HeffOracle
from refs import rel_frobenius

pytestmark = pytest.mark.gpu


def outputs(p):
    return p.Hpsi.parts()["data"].copy()


@pytest.mark.parametrize("flags", [sb.PREC_FP64, sb.PREC_FP32], ids=["fp64", "fp32"])
def test_batch_64_j1j2_equals_alone(flags):
    print(f"[config batch] 64 J1-J2 problems flags={flags}", flush=True)
    probs = H.j1j2_batch_problems(64)
    chis = [c for _, _, c in probs]
    assert len(probs) == 64 and min(chis) >= 64 and max(chis) <= 1024 and max(chis) > 512
    for p, _, _ in probs:
        p.upload().outputs().make_plans(flags)
    alone = []
    for p, _, _ in probs:
        p.execute()
        alone.append(outputs(p))
        p.Hpsi.fill_(0.0)
    batch = sb.PlanBatch([p.plans for p, _, _ in probs])
    assert batch.stats()["fused_stages"] == 4
    batch.execute()
    for c, (p, _, _) in enumerate(probs):
        assert np.array_equal(outputs(p), alone[c]), c
    if flags == sb.PREC_FP64:
        # two members against the oracle: the smallest bond and one of chi ~ 256
        order = np.argsort(chis)
        pick = [int(order[0]), int(order[np.searchsorted(np.sort(chis), 256)])]
        O.use_blas(True, threads=8)
        try:
            for c in pick:
                p = probs[c][0]
                q = HeffOracle(p).run().parts()
                hp = p.Hpsi.parts()
                assert np.array_equal(hp["keys"], q["keys"]) and np.array_equal(hp["ids"], q["ids"])
                assert rel_frobenius(hp["data"], q["data"]) <= 1e-12, (c, chis[c])
        finally:
            O.use_blas(False)


def test_precision_forms_hubbard_chi2048():
    print("[config precision] hubbard chi=2048", flush=True)
    prob = H.hubbard_chi2048_problem(0)
    O.use_blas(True, threads=8)
    try:
        q = HeffOracle(prob).run().parts()
    finally:
        O.use_blas(False)
    errs = {}
    for name, fl in (("fp32", sb.PREC_FP32), ("bf16x3", sb.PREC_BF16X3)):
        prob.upload().outputs()
        prob.make_plans(fl)
        got = prob.Hpsi.parts()
        assert np.array_equal(got["keys"], q["keys"]) and np.array_equal(got["ids"], q["ids"])
        errs[name] = rel_frobenius(got["data"], q["data"])
    prob.upload()
    prob.make_fused_plan(sb.PREC_FP32)
    got = prob.Hpsi_fused.parts()
    assert np.array_equal(got["keys"], q["keys"]) and np.array_equal(got["ids"], q["ids"])
    errs["fused_fp32"] = rel_frobenius(got["data"], q["data"])
    print("[config precision] relative errors", errs, flush=True)
    for name, e in errs.items():
        assert 0 < e <= 2e-6, (name, e)


@pytest.mark.parametrize("sites,chi", [(32, 128), (64, 256)])
def test_serving_sizes_equal_alone(sites, chi):
    from sparsebackends_amd import serve as S
    print(f"[config serving] N={sites} chi_max={chi}", flush=True)
    rq = [S.SolveRequest(i, sites, chi, delta, 0.0, 1e-10, 8) for i, delta in enumerate((0.7, 1.3))]
    plain = S.run_plain(rq)
    res = S.DMRGServer(slots=2).run(rq)
    for a, b in zip(plain, res):
        assert a.energy == b.energy and a.energies == b.energies, (a.rid, a.energy, b.energy)
        assert a.maxlinkdim <= chi and np.isfinite(a.energy)
    print("[config serving] maxlinkdim", [r.maxlinkdim for r in plain], "sweeps", [r.sweeps for r in plain],
          flush=True)
    assert all(r.maxlinkdim >= 64 for r in plain)  # bonds well beyond the small serving tests' 16-32


def oracle_heff(psi, L, W, R):
    """the oracle's four contract! calls of H_eff*psi on the device tensors' host parts"""
    def bs(t):
        p = t.parts()
        return O.BS.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])
    nl, nr, d1, d2, Dl, Dr = psi.dims
    w, wr, wm = W.dims[0], W.dims[3], W.dims[3]
    opsi, oL, oW, oR = bs(psi), bs(L), bs(W), bs(R)
    T1 = O.BS.new((nl, nr, d1, d2, Dr, w, Dl), 5)
    O.contract(T1, H.L_T1, opsi, H.L_PSI, oL, H.L_L)
    T2 = O.BS.new((nl, nr, d1, wm, d2, Dr, Dl), 5)
    O.contract(T2, H.L_T2, T1, H.L_T1, oW, H.L_W1)
    T3 = O.BS.new((nl, nr, d2, wr, d1, Dr, Dl), 5)
    O.contract(T3, H.L_T3, T2, H.L_T2, oW, H.L_W2)
    Hp = O.BS.new((nl, nr, d1, d2, Dl, Dr), 4)
    O.contract(Hp, H.L_OUT, T3, H.L_T3, oR, H.L_R)
    return Hp.parts()


def test_config1_sweep_from_qr_start():
    """S=1/2 Heisenberg J=1, N=32, chi=256 in Sz sectors, MPS i.i.d. N(0,1)
    from seed 0 left-orthonormalised by QR, Krylov dim 3, cutoff 1e-10"""
    from sparsebackends_amd import dmrg as D
    print("[config 1] N=32 chi=256 sweep from the QR-orthonormalised random MPS", flush=True)
    m = H.heisenberg_model(1.0)
    bonds = D.config_bonds(32, chi=256)
    A, bonds = D.left_orthonormalize(m, D.random_mps(m, bonds, seed=0), bonds)
    assert max(b.dim for b in bonds) == 256
    dm = D.DMRG.from_mps(m, A, bonds)
    # the first two-site step's H_eff*psi against the oracle, at the bond of the largest sectors
    b = 16
    for bb in range(1, b):
        dm.step(bb, +1, 256, 1e-10, 3, 1, 1e-14)
    psi = D.two_site(dm.A[b - 1], dm.A[b])
    plans, Hp = H.heff_plans(psi, dm.L[b - 1], dm.W, dm.W, dm.R[b + 2])
    got = Hp.parts()
    O.use_blas(True, threads=8)
    try:
        q = oracle_heff(psi, dm.L[b - 1], dm.W, dm.R[b + 2])
    finally:
        O.use_blas(False)
    assert np.array_equal(got["keys"], q["keys"]) and np.array_equal(got["ids"], q["ids"])
    assert rel_frobenius(got["data"], q["data"]) <= 1e-12
    # one sweep at chi = 256, and a chi = 512 run from the same start: variational order
    dm = D.DMRG.from_mps(m, A, bonds)
    e256 = dm.sweep(maxdim=256, cutoff=1e-10, krylovdim=3, maxiter=1, tol=1e-14)
    ref = D.DMRG.from_mps(m, A, bonds)
    e512 = None
    for _ in range(3):
        e512 = ref.sweep(maxdim=512, cutoff=1e-10, krylovdim=3, maxiter=1, tol=1e-14)
    print(f"[config 1] E(chi 256, one sweep) = {e256:.12f}, E(chi 512, three sweeps) = {e512:.12f}", flush=True)
    assert -0.45 * 32 < e512 <= e256 + 1e-10
    assert dm.maxlinkdim() <= 256
