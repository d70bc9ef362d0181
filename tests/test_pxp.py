"""The COO workload (PXP / Rydberg-blockade chain, chi = 128, ~30 % fill):
the oracle's psi (COO) x L (dense) and the MPO step against numpy's einsum
and against the nonzero index set of psi (one output block per (s1, s2, r)
that holds a nonzero, contract_coo_dense.jl:113-122); on the device (-m gpu)
the same contractions through the C ABI against the oracle: keys and ids bit
for bit, the exact-order COO x Dense values bit for bit, the MPO step (block
products) to 1e-12 relative."""
import numpy as np
import pytest

from sparsebackends_amd import pxp as X
from oracle import oracle as O


def oracle_steps(pr):
    chi = pr["chi"]
    A = O.COO.from_dense(pr["psi"], atol=1e-12)
    C1 = O.BS.new((2, 2, chi, chi, 3), 2)
    O.contract(C1, X.L_C1, A, X.L_PSI, O.Dense.new(pr["L"]), X.L_ENV)
    C2 = O.BS.new((2, chi, chi, 2, 3), 3)
    O.contract(C2, X.L_C2, C1, X.L_C1, O.Dense.new(pr["W"]), X.L_W)
    return A, C1, C2


@pytest.mark.parametrize("chi,seed", [(128, 0), (32, 5)])
def test_oracle_pxp(chi, seed):
    pr = X.problem(chi, seed)
    assert 0.25 < pr["fill"] < 0.40
    A, C1, C2 = oracle_steps(pr)
    psi = pr["psi"]
    r1 = np.einsum("lxyr,pwl->xyrpw", psi, pr["L"])
    np.testing.assert_allclose(C1.to_dense(), r1, rtol=0, atol=1e-12 * np.abs(r1).max())
    # one block per (s1, s2, r) holding a nonzero of psi, keys in column-major order
    nz = np.argwhere(np.any(psi != 0, axis=0).transpose(0, 1, 2))  # (s1, s2, r)
    want = sorted({(int(a) + 1, int(b) + 1, int(c) + 1) for a, b, c in nz}, key=lambda k: (k[2], k[1], k[0]))
    assert [tuple(k) for k in C1.parts()["keys"]] == want
    r2 = np.einsum("xyrpw,wzxv->yrpzv", r1, pr["W"])
    np.testing.assert_allclose(C2.to_dense(), r2, rtol=0, atol=1e-12 * np.abs(r2).max())


@pytest.mark.gpu
@pytest.mark.parametrize("chi,seed", [(128, 0), (32, 5)])
def test_gpu_pxp(chi, seed):
    import sparsebackends_amd as sb
    print(f"[pxp] chi {chi} seed {seed}", flush=True)
    pr = X.problem(chi, seed)
    _, C1o, C2o = oracle_steps(pr)
    A = sb.coo_from_dense(pr["psi"], atol=1e-12)
    C1 = sb.NewBlockSparseSorted.new((2, 2, chi, chi, 3), 2)
    sb.contract_(C1, X.L_C1, A, X.L_PSI, sb.DenseTensor.from_numpy(pr["L"]), X.L_ENV)
    C2 = sb.NewBlockSparseSorted.new((2, chi, chi, 2, 3), 3)
    sb.contract_(C2, X.L_C2, C1, X.L_C1, sb.DenseTensor.from_numpy(pr["W"]), X.L_W)
    p, q = C1.parts(), C1o.parts()
    assert np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"])
    assert np.array_equal(p["data"], q["data"])
    p, q = C2.parts(), C2o.parts()
    assert np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"])
    assert np.linalg.norm(p["data"] - q["data"]) <= 1e-12 * np.linalg.norm(q["data"])
